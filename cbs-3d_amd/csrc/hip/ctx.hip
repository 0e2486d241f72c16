// C ABI of libcbs3d_hip.so (include/cbs_hip.h): context creation, device
// layouts, field transfer, error reporting and the per-step orchestration of
// Solver::advanceOneStep (src/solver/Solver.cpp:609-680).
#include <cctype>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <map>
#include <stdexcept>
#include <thread>

#include "device_common.hpp"

using namespace cbsdev;

namespace
{
    thread_local std::string g_create_error;
    // tiles per segment aimed at by contexts created from now on
    // (cbs_hip_set_tile_target)
    std::atomic<int32_t> g_tile_target{512};
    // elements from which contexts created from now on store the element
    // passes' terms as whole lines (cbs_hip_set_line_group_min_elems)
    std::atomic<int32_t> g_line_group_min{kLineGroupMinElems};

    struct HipFailure : std::runtime_error
    {
        using std::runtime_error::runtime_error;
    };

    struct CbsFailure : std::runtime_error
    {
        int status;
        CbsFailure(int s, const std::string& m) : std::runtime_error(m), status(s) {}
    };

    template <typename T>
    T* dalloc(cbs_hip_ctx* c, size_t n)
    {
        void* p = nullptr;
        const size_t bytes = (n == 0 ? 1 : n) * sizeof(T);
        hip_check(hipMalloc(&p, bytes), "hipMalloc");
        c->allocations.push_back(p);
        c->device_bytes += bytes;
        hip_check(hipMemsetAsync(p, 0, bytes, c->stream), "hipMemsetAsync");
        return static_cast<T*>(p);
    }

    template <typename T>
    T* upload(cbs_hip_ctx* c, const T* h, size_t n)
    {
        T* d = dalloc<T>(c, n);
        if (n > 0) hip_check(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream), "upload");
        return d;
    }

    template <typename T>
    T* upload(cbs_hip_ctx* c, const std::vector<T>& h)
    {
        return upload(c, h.data(), h.size());
    }

    // Message of an error key, in the reference's wording.
    std::string message_for(unsigned long long key, int* status)
    {
        const uint32_t stage = uint32_t(key >> 48), loop = uint32_t((key >> 40) & 0xff);
        const uint32_t index = uint32_t((key >> 8) & 0xffffffffu), check = uint32_t(key & 0xff);
        const std::string idx = std::to_string(index);
        *status = CBS_ERR_INVALID_GEOMETRY;
        switch (stage)
        {
        case ST_PTERMS:
            return "PressureAssembly::buildElementPressureTerms - invalid detJ at element " + idx;
        case ST_TIMESTEP:
            *status = CBS_ERR_INVALID_TIMESTEP;
            if (loop == 1) return "TimeStep::computeTimeStep - invalid nodal timestep at node " + idx;
            if (loop == 3) return "TimeStep::computeTimeStep - invalid global minimum timestep";
            if (check == 1) return "TimeStep - invalid alen_e at element " + idx;
            if (check == 2) return "TimeStep - invalid thermal diffusivity alpha_e at element " + idx;
            if (check == 3) return "TimeStep - invalid fluid rho/mu at element " + idx;
            if (check == 4) return "TimeStep - invalid timestep diffusivity at element " + idx;
            return "TimeStep - invalid local stability timestep at element " + idx;
        case ST_LHS:
            *status = CBS_ERR_INVALID_TIMESTEP;
            if (loop == 1 && check == 1)
                return "TimeStep::updateLhsDiagonal - non-positive local timestep at element " + idx;
            if (loop == 1) return "TimeStep::updateLhsDiagonal - invalid element-node mass";
            if (loop == 2 && check == 1)
            {
                *status = CBS_ERR_INVALID_GEOMETRY;
                return "TimeStep::updateLhsDiagonal - orphan node detected at node " + idx;
            }
            if (loop == 2 && check == 2)
                return "TimeStep::updateLhsDiagonal - non-positive fluid momentum mass/time diagonal at node " + idx;
            if (loop == 2) return "TimeStep::updateLhsDiagonal - non-positive capacitance/time diagonal at node " + idx;
            return "TimeStep::updateLhsDiagonal - zero pressure diagonal on fluid-connected node " + idx;
        case ST_STEP1:
            if (loop == 0)
            {
                *status = CBS_ERR_INVALID_TIMESTEP;
                return "MomentumAssembly::assembleStep1Rhs - dtreal must be positive before Step 1 assembly";
            }
            if (loop == 2) return "MomentumAssembly::assembleStep1Rhs - invalid detJ at one or more elements";
            if (check == 2)
            {
                *status = CBS_ERR_INVALID_TIMESTEP;
                return "MomentumAssembly::assembleStep1Rhs - invalid element timestep at element " + idx;
            }
            *status = CBS_ERR_INVALID_MATERIAL;
            return "MomentumAssembly::assembleStep1Rhs - invalid fluid rho/mu at element " + idx;
        case ST_STEP2RHS:
            if (loop == 0)
            {
                *status = CBS_ERR_INVALID_TIMESTEP;
                return "PressureAssembly::assembleStep2Rhs - dtreal must be positive before Step 2 RHS assembly";
            }
            return "PressureAssembly::assembleStep2Rhs - invalid detJ at element " + idx;
        case ST_CG:
            if (check == 1)
            {
                *status = CBS_ERR_JACOBI_DIAGONAL;
                return "ConjugateGradient::solvePressure - invalid Jacobi pressure diagonal at active node " + idx;
            }
            if (check == 2)
            {
                *status = CBS_ERR_NONFINITE_RESIDUAL;
                return "ConjugateGradient::solvePressure - invalid pressure residual at node " + idx;
            }
            if (check == 3)
            {
                *status = CBS_ERR_NONFINITE_RESIDUAL;
                return "ConjugateGradient::solvePressure - invalid initial preconditioned residual dot product.";
            }
            if (check == 4)
            {
                *status = CBS_ERR_CG_BREAKDOWN;
                return "ConjugateGradient::solvePressure - breakdown: invalid pAp";
            }
            *status = CBS_ERR_NONFINITE_RESIDUAL;
            return "ConjugateGradient::solvePressure - invalid updated preconditioned residual dot product.";
        case ST_STEP3:
            return "Steps::step3SemiImplicit - invalid detJ at element " + idx;
        case ST_STEP4:
            if (check == 1) return "EnergyAssembly::assembleStep4Rhs - invalid detJ at element " + idx;
            *status = CBS_ERR_INVALID_MATERIAL;
            if (check == 2) return "EnergyAssembly::assembleStep4Rhs - invalid rho*Cp at element " + idx;
            if (check == 3) return "EnergyAssembly::assembleStep4Rhs - invalid thermal conductivity at element " + idx;
            if (check == 4) return "EnergyAssembly::assembleStep4Rhs - invalid volumetric source at element " + idx;
            *status = CBS_ERR_INVALID_TIMESTEP;
            return "EnergyAssembly::assembleStep4Rhs - element timestep must be positive for thermal stabilisation at element " + idx;
        case ST_CONV:
            return "Convergence - orphan node not touched by any material element";
        default:
            return "device error";
        }
    }

    // Runs f; converts exceptions into status codes and the context message.
    template <typename F>
    int guarded(cbs_hip_ctx* c, F&& f)
    {
        if (!c) return CBS_ERR_BAD_ARGUMENT;
        try
        {
            return f();
        }
        catch (const CbsFailure& e)
        {
            c->err = e.what();
            return e.status;
        }
        catch (const HipFailure& e)
        {
            c->err = e.what();
            return CBS_ERR_HIP;
        }
        catch (const std::bad_alloc&)
        {
            c->err = "host allocation failed";
            return CBS_ERR_HIP;
        }
        catch (const std::exception& e)
        {
            c->err = e.what();
            return CBS_ERR_BAD_ARGUMENT;
        }
    }

    bool supported_bc(int bc)
    {
        switch (bc)
        {
        case 500: case 501: case 502: case 503: case 504: case 506: case 507: case 508:
        case 510: case 511: case 520: case 530: case 532: case 901: case 902:
            return true;
        default:
            return false;
        }
    }

    // Boundary.cpp:164-183
    double bfs_profile_u(double y)
    {
        return 0.6624 * std::pow(y, 6) - 7.5547 * std::pow(y, 5) + 33.9 * std::pow(y, 4) -
               75.283 * std::pow(y, 3) + 83.368 * std::pow(y, 2) - 37.793 * y + 2.6959;
    }

    double rectangular_profile_u(double y) { return 6.0 * y * (1.0 - y); }

    // Tiles for the LDS-staged apply.  Recursive coordinate bisection of the
    // fluid elements' centroids into leaves of at most `tmax` elements; a
    // leaf whose node set exceeds `lmax` makes the caller retry smaller.
    // Each tile lists its interior nodes (referenced by no other tile) first
    // and its shared nodes after them, both in ascending node id; its
    // elements are ordered by (colour, element).
    struct Tiles
    {
        std::vector<int32_t> node_ptr{0}, nodes, sidx, col_ptr, order, snode;
        std::vector<ushort4> lidx;
        int32_t lmax = 0, nshared = 0;
    };

    // fluid: the elements to tile, all with nodes in [node_lo, node_hi).
    // The bisection, the per-tile element orders and node sets, and the
    // local ids run on host threads; every range and tile is processed by
    // the same deterministic steps whichever thread takes it, so the tiles
    // do not depend on the thread count.
    bool build_tiles(const cbs_mesh_desc* m, const std::vector<int4>& conn, const std::vector<int32_t>& elem_color,
                     const std::vector<int32_t>& fluid, int node_lo, int node_hi, int ncolor, int tmax,
                     int lmax_allowed, Tiles& out)
    {
        const int N = node_hi - node_lo;
        const size_t F = fluid.size();
        if (F == 0)
        {
            out = Tiles{};
            return true;
        }
        std::vector<float> cen(3 * F);
        constexpr int64_t kGrain = 1 << 16;
        host_parallel_for((int64_t(F) + kGrain - 1) / kGrain, [&](int64_t g) {
            const size_t k1 = std::min(F, size_t(g + 1) * size_t(kGrain));
            for (size_t k = size_t(g) * size_t(kGrain); k < k1; ++k)
            {
                const int4 n = conn[fluid[k]];
                for (int d = 0; d < 3; ++d)
                    cen[3 * k + d] = float(0.25 * (m->coord[3 * size_t(n.x) + d] + m->coord[3 * size_t(n.y) + d] +
                                                   m->coord[3 * size_t(n.z) + d] + m->coord[3 * size_t(n.w) + d]));
            }
        });
        std::vector<int32_t> idx(F);
        for (size_t k = 0; k < F; ++k) idx[k] = int32_t(k);
        // one bisection step of [b, e): the axis of the largest extent, the
        // median by (coordinate, index)
        auto split = [&](size_t b, size_t e) {
            float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
            for (size_t k = b; k < e; ++k)
                for (int d = 0; d < 3; ++d)
                {
                    lo[d] = std::min(lo[d], cen[3 * size_t(idx[k]) + d]);
                    hi[d] = std::max(hi[d], cen[3 * size_t(idx[k]) + d]);
                }
            int ax = 0;
            for (int d = 1; d < 3; ++d)
                if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
            const size_t mid = b + (e - b) / 2;
            std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e, [&](int32_t p, int32_t q) {
                const float a = cen[3 * size_t(p) + ax], bb = cen[3 * size_t(q) + ax];
                return a < bb || (a == bb && p < q);
            });
            return mid;
        };
        // the top of the tree level by level (each level's ranges on
        // threads), then every remaining subtree on one thread
        std::vector<std::pair<size_t, size_t>> leaves, level{{0, F}};
        while (!level.empty() && level.size() < 64)
        {
            std::vector<size_t> mids(level.size());
            host_parallel_for(int64_t(level.size()), [&](int64_t r) {
                const auto [b, e] = level[size_t(r)];
                mids[size_t(r)] = e - b <= size_t(tmax) ? 0 : split(b, e);
            });
            std::vector<std::pair<size_t, size_t>> nextl;
            for (size_t r = 0; r < level.size(); ++r)
            {
                const auto [b, e] = level[r];
                if (e - b <= size_t(tmax)) leaves.push_back({b, e});
                else
                {
                    nextl.push_back({b, mids[r]});
                    nextl.push_back({mids[r], e});
                }
            }
            level.swap(nextl);
        }
        std::vector<std::vector<std::pair<size_t, size_t>>> sub(level.size());
        host_parallel_for(int64_t(level.size()), [&](int64_t r) {
            std::vector<std::pair<size_t, size_t>> stack{level[size_t(r)]};
            while (!stack.empty())
            {
                const auto [b, e] = stack.back();
                stack.pop_back();
                if (e - b <= size_t(tmax))
                {
                    sub[size_t(r)].push_back({b, e});
                    continue;
                }
                const size_t mid = split(b, e);
                stack.push_back({mid, e});
                stack.push_back({b, mid});
            }
        });
        for (const auto& v : sub) leaves.insert(leaves.end(), v.begin(), v.end());
        std::sort(leaves.begin(), leaves.end());
        out = Tiles{};
        const size_t T = leaves.size();
        // pass 1: element order (colour, element) and node set of every tile
        std::vector<std::vector<int32_t>> tile_el(T), tile_nodes(T);
        std::atomic<bool> too_big{false};
        host_parallel_for(int64_t(T), [&](int64_t t) {
            const auto [b, e] = leaves[size_t(t)];
            std::vector<int32_t>& el = tile_el[size_t(t)];
            el.assign(idx.begin() + b, idx.begin() + e);
            for (auto& k : el) k = fluid[size_t(k)];
            std::sort(el.begin(), el.end(), [&](int32_t p, int32_t q) {
                return elem_color[p] < elem_color[q] || (elem_color[p] == elem_color[q] && p < q);
            });
            std::vector<int32_t>& nodes = tile_nodes[size_t(t)];
            nodes.reserve(el.size() * 4);
            for (int32_t ei : el)
            {
                const int4 n = conn[ei];
                nodes.insert(nodes.end(), {n.x, n.y, n.z, n.w});
            }
            std::sort(nodes.begin(), nodes.end());
            nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
            if (int(nodes.size()) > lmax_allowed || nodes.size() > 65535) too_big = true;
        });
        if (too_big) return false;
        out.col_ptr.reserve(T * size_t(ncolor + 1));
        out.lidx.resize(F);
        out.order.reserve(F);
        std::vector<int32_t> tile_count(N, 0);
        std::vector<size_t> el_base(T + 1, 0), nd_base(T + 1, 0);
        for (size_t t = 0; t < T; ++t)
        {
            const std::vector<int32_t>& el = tile_el[t];
            out.lmax = std::max(out.lmax, int32_t(tile_nodes[t].size()));
            for (int32_t nd : tile_nodes[t]) tile_count[nd - node_lo]++;
            const size_t base = out.order.size();
            std::vector<int32_t> cp(size_t(ncolor) + 1, 0);
            for (int32_t ei : el) cp[size_t(elem_color[ei]) + 1]++;
            for (int col = 0; col < ncolor; ++col) cp[size_t(col) + 1] += cp[size_t(col)];
            for (int col = 0; col <= ncolor; ++col) out.col_ptr.push_back(int32_t(base) + cp[size_t(col)]);
            out.order.insert(out.order.end(), el.begin(), el.end());
            el_base[t + 1] = out.order.size();
            nd_base[t + 1] = nd_base[t] + tile_nodes[t].size();
        }
        std::vector<int32_t> srow(N, -1);
        for (int i = 0; i < N; ++i)
            if (tile_count[i] > 1)
            {
                srow[i] = int32_t(out.snode.size());
                out.snode.push_back(node_lo + i);
            }
        out.nshared = int32_t(out.snode.size());
        // pass 2: interior-first node order, local ids (a tile's local id of
        // a node is found by binary search in its sorted node set)
        out.nodes.resize(nd_base[T]);
        out.sidx.resize(nd_base[T]);
        host_parallel_for(int64_t(T), [&](int64_t t) {
            const std::vector<int32_t>& sorted = tile_nodes[size_t(t)];
            std::vector<int32_t> nodes(sorted);
            std::stable_partition(nodes.begin(), nodes.end(),
                                  [&](int32_t nd) { return tile_count[nd - node_lo] == 1; });
            std::vector<uint16_t> lid(sorted.size());
            for (size_t l = 0; l < nodes.size(); ++l)
                lid[size_t(std::lower_bound(sorted.begin(), sorted.end(), nodes[l]) - sorted.begin())] = uint16_t(l);
            auto local = [&](int32_t nd) {
                return lid[size_t(std::lower_bound(sorted.begin(), sorted.end(), nd) - sorted.begin())];
            };
            for (size_t k = el_base[size_t(t)]; k < el_base[size_t(t) + 1]; ++k)
            {
                const int4 n = conn[out.order[k]];
                out.lidx[k] = make_ushort4(local(n.x), local(n.y), local(n.z), local(n.w));
            }
            const size_t nb = nd_base[size_t(t)];
            for (size_t l = 0; l < nodes.size(); ++l)
            {
                out.nodes[nb + l] = nodes[l];
                out.sidx[nb + l] = srow[nodes[l] - node_lo];
            }
        });
        for (size_t t = 0; t < T; ++t) out.node_ptr.push_back(int32_t(nd_base[t + 1]));
        return true;
    }

    // Appends the tiles of one segment to those of the previous segments.
    void append_tiles(Tiles& all, const Tiles& t)
    {
        const int32_t nbase = int32_t(all.nodes.size()), obase = int32_t(all.order.size());
        const int32_t sbase = all.nshared;
        for (size_t k = 1; k < t.node_ptr.size(); ++k) all.node_ptr.push_back(nbase + t.node_ptr[k]);
        all.nodes.insert(all.nodes.end(), t.nodes.begin(), t.nodes.end());
        for (int32_t v : t.sidx) all.sidx.push_back(v < 0 ? v : sbase + v);
        for (int32_t v : t.col_ptr) all.col_ptr.push_back(obase + v);
        all.order.insert(all.order.end(), t.order.begin(), t.order.end());
        all.lidx.insert(all.lidx.end(), t.lidx.begin(), t.lidx.end());
        all.snode.insert(all.snode.end(), t.snode.begin(), t.snode.end());
        all.lmax = std::max(all.lmax, t.lmax);
        all.nshared += t.nshared;
    }

    // Block map of the items [off[s], off[s+1]) of every segment: each
    // segment is cut as a context holding it alone would cut it, into
    // reduce_grid(len) contiguous blocks (per_block == 0), or into blocks of
    // per_block items.
    void build_blocks(cbs_hip_ctx* c, const std::vector<int32_t>& off, int per_block, cbs_hip_blocks& b);

    // Slots of the node sums for the per-node term lists list[ptr[i] ..
    // ptr[i+1]) (entries e*4 + a, in summation order).
    void build_node_slots(cbs_hip_ctx* c, int E, const std::vector<int32_t>& ptr, const std::vector<int32_t>& list,
                          cbs_hip_ctx::NodeSlots& ns, bool line_groups)
    {
        const int N = int(ptr.size()) - 1;
        const int nslice = (N + 63) / 64;
        std::vector<int64_t> off(size_t(nslice) + 1, 0);
        for (int s = 0; s < nslice; ++s)
        {
            int kmax = 0;
            for (int i = s * 64; i < std::min(N, s * 64 + 64); ++i) kmax = std::max(kmax, ptr[size_t(i) + 1] - ptr[size_t(i)]);
            off[size_t(s) + 1] = off[size_t(s)] + int64_t(kmax) * 64;
        }
        if (off.back() >= (int64_t(1) << 32))
            throw CbsFailure(CBS_ERR_UNSUPPORTED, "cbs_hip_create - too many element-node incidences for 32-bit slots");
        std::vector<uint32_t> pos(size_t(E) * 4, 0);
        for (int i = 0; i < N; ++i)
        {
            const int s = i / 64, l = i % 64;
            for (int k = ptr[size_t(i)]; k < ptr[size_t(i) + 1]; ++k)
                pos[size_t(list[size_t(k)])] = uint32_t(off[size_t(s)] + int64_t(k - ptr[size_t(i)]) * 64 + l);
        }
        ns.nslots = off.back();
        ns.off = upload(c, off);
        ns.pos = upload(c, pos);
        if (line_groups)
        {
            // windows of kSlotWindow elements, in parallel over contiguous
            // runs of windows; each window's lines ascending
            const int nwin = (E + kSlotWindow - 1) / kSlotWindow;
            const int nth = std::max(1, std::min({16, int(std::thread::hardware_concurrency()), nwin}));
            std::vector<std::vector<int32_t>> wcount(static_cast<size_t>(nth));
            std::vector<std::vector<uint32_t>> wline(static_cast<size_t>(nth));
            std::vector<std::vector<uint16_t>> wsrc(static_cast<size_t>(nth));
            std::vector<std::thread> pool;
            for (int t = 0; t < nth; ++t)
                pool.emplace_back([&, t] {
                    std::vector<std::pair<uint32_t, uint16_t>> inc;
                    for (int w = int(int64_t(nwin) * t / nth); w < int(int64_t(nwin) * (t + 1) / nth); ++w)
                    {
                        const int b = w * kSlotWindow, e1 = std::min(E, b + kSlotWindow);
                        inc.clear();
                        for (int e = b; e < e1; ++e)
                            for (int a = 0; a < 4; ++a)
                                inc.push_back({pos[size_t(e) * 4 + size_t(a)], uint16_t((e - b) * 4 + a)});
                        std::sort(inc.begin(), inc.end());
                        int32_t ng = 0;
                        for (size_t k = 0; k < inc.size(); ++ng)
                        {
                            const uint32_t line = inc[k].first >> 4;
                            wline[size_t(t)].push_back(line);
                            const size_t g0 = wsrc[size_t(t)].size();
                            wsrc[size_t(t)].resize(g0 + 16, uint16_t(0xFFFF));
                            for (; k < inc.size() && (inc[k].first >> 4) == line; ++k)
                                wsrc[size_t(t)][g0 + (inc[k].first & 15u)] = inc[k].second;
                        }
                        wcount[size_t(t)].push_back(ng);
                    }
                });
            for (auto& th : pool) th.join();
            std::vector<int32_t> gptr(1, 0);
            std::vector<uint32_t> gline;
            std::vector<uint16_t> gsrc;
            for (int t = 0; t < nth; ++t)
            {
                for (const int32_t n : wcount[size_t(t)]) gptr.push_back(gptr.back() + n);
                gline.insert(gline.end(), wline[size_t(t)].begin(), wline[size_t(t)].end());
                gsrc.insert(gsrc.end(), wsrc[size_t(t)].begin(), wsrc[size_t(t)].end());
            }
            ns.grp_ptr = upload(c, gptr);
            ns.grp_line = upload(c, gline);
            ns.grp_src = upload(c, gsrc);
        }
        ns.q = dalloc<double>(c, size_t(ns.nslots) * 3);  // zero-filled: the padding slots
    }

    // Builds every device layout of the context from the mesh descriptor.
    void build_context(cbs_hip_ctx* c, const cbs_mesh_desc* m)
    {
        const int N = m->npoin, E = m->nelem, B = m->nboun;
        c->N = N;
        c->E = E;
        c->B = B;
        c->ncolor = m->ncolor;
        c->bc_fixed = m->bc_fixed;
        c->npoin_wall = m->npoin_wall;
        auto bad = [](const std::string& s) { throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_create - " + s); };
        c->create_seconds.clear();
        auto t_phase = std::chrono::steady_clock::now();
        auto lap = [&] {
            const auto now = std::chrono::steady_clock::now();
            c->create_seconds.push_back(std::chrono::duration<double>(now - t_phase).count());
            t_phase = now;
        };
        if (N < 1 || E < 1 || B < 0 || m->ncolor < 1 || m->bc_fixed < 0 || m->npoin_wall < 0) bad("invalid sizes");
        if (!m->intma || !m->coord || !m->mat_elem || !m->rho_e || !m->mu_e || !m->k_e || !m->rho_cp_e ||
            !m->alpha_e || !m->Qvol_e || !m->dNkdx || !m->detJ || !m->annxf || !m->fedge || !m->elcoe_e ||
            !m->alen_e || (B > 0 && (!m->iside || !m->face_norm)) || !m->color_ptr || !m->color_elem ||
            (m->bc_fixed > 0 && (!m->bc_list || !m->bc_values)) || (m->npoin_wall > 0 && !m->wall_node_list))
            bad("null array");
        const cbs_config& g = c->cfg;
        if (c->h_seg_node.empty())
        {
            c->nseg = 1;
            c->h_seg_node = {0, N};
            c->h_seg_elem = {0, E};
        }
        if (int(c->h_seg_node.size()) != c->nseg + 1 || int(c->h_seg_elem.size()) != c->nseg + 1 ||
            c->h_seg_node.back() != N || c->h_seg_elem.back() != E)
            bad("invalid segments");
        for (int sg = 0; sg < c->nseg; ++sg)
        {
            const int n0 = c->h_seg_node[size_t(sg)], n1 = c->h_seg_node[size_t(sg) + 1];
            for (int e = c->h_seg_elem[size_t(sg)]; e < c->h_seg_elem[size_t(sg) + 1]; ++e)
                for (int a = 0; a < 4; ++a)
                {
                    const int v = m->intma[size_t(e) * 4 + a] - 1;
                    if (v < n0 || v >= n1) bad("element of a segment references a node of another segment");
                }
        }

        lap();  // validation
        // connectivity, flags
        std::vector<int4> conn(E);
        std::vector<uint8_t> tf(N, 0), ts(N, 0);
        std::vector<uint32_t> fe(E);
        for (int e = 0; e < E; ++e)
        {
            int v[4];
            for (int a = 0; a < 4; ++a)
            {
                v[a] = m->intma[size_t(e) * 4 + a] - 1;
                if (v[a] < 0 || v[a] >= N) bad("connectivity node out of range");
                (m->mat_elem[e] == 0 ? tf : ts)[v[a]] = 1;
            }
            conn[e] = make_int4(v[0], v[1], v[2], v[3]);
            uint32_t f = 0;
            for (int s = 0; s < 4; ++s) f |= uint32_t(m->fedge[size_t(e) * 4 + s] & 0xff) << (8 * s);
            fe[e] = f;
        }
        std::vector<uint8_t> fixed(N, 0);
        for (int i = 0; i < m->bc_fixed; ++i)
        {
            const int ip = m->bc_list[i] - 1;
            if (ip < 0 || ip >= N) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "fixed pressure node out of range");
            if (!tf[ip])
                throw CbsFailure(CBS_ERR_INVALID_GEOMETRY,
                                 "ConjugateGradient::solvePressure - fixed pressure node is not fluid-connected: node " +
                                     std::to_string(ip + 1));
            fixed[ip] = 1;
        }
        std::vector<uint8_t> nflag(N);
        for (int i = 0; i < N; ++i) nflag[i] = uint8_t(tf[i] | (ts[i] << 1) | (fixed[i] << 2));

        lap();  // connectivity
        // fluid colour-major slots
        c->h_fcolor_ptr.assign(size_t(c->ncolor) + 1, 0);
        std::vector<int32_t> slot_elem, elem_slot(E, -1);
        slot_elem.reserve(E);
        for (int col = 0; col < c->ncolor; ++col)
        {
            const int beg = m->color_ptr[col], end = m->color_ptr[col + 1];
            if (beg < 0 || end > E || end < beg) bad("invalid color_ptr");
            for (int k = beg; k < end; ++k)
            {
                const int e = m->color_elem[k] - 1;
                if (e < 0 || e >= E) bad("invalid color_elem");
                if (m->mat_elem[e] != 0) continue;
                elem_slot[e] = int32_t(slot_elem.size());
                slot_elem.push_back(e);
            }
            c->h_fcolor_ptr[col + 1] = int32_t(slot_elem.size());
        }
        c->F = int32_t(slot_elem.size());
        std::vector<int4> slot_conn(c->F);
        for (int k = 0; k < c->F; ++k) slot_conn[k] = conn[slot_elem[k]];

        lap();  // colour slots
        // tiles of the LDS-staged apply, sized to the LDS of one workgroup
        // (FP64: x, y and the side-buffer row of every local node)
        Tiles tiles;
        {
            std::vector<int32_t> elem_color(E, -1);
            for (int col = 0; col < c->ncolor; ++col)
                for (int k = m->color_ptr[col]; k < m->color_ptr[col + 1]; ++k) elem_color[m->color_elem[k] - 1] = col;
            int lds = 65536;
            (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
            const int lmax_allowed = std::min(65535, (lds - 1024) / 20);
            std::vector<int32_t> fluid(slot_elem.begin(), slot_elem.end());
            std::sort(fluid.begin(), fluid.end());
            // tiles never straddle segments: each segment is tiled as a
            // context holding it alone would be
            c->h_seg_tile.assign(1, 0);
            size_t f0 = 0;
            for (int sg = 0; sg < c->nseg; ++sg)
            {
                size_t f1 = f0;
                while (f1 < fluid.size() && fluid[f1] < c->h_seg_elem[size_t(sg) + 1]) ++f1;
                const std::vector<int32_t> part(fluid.begin() + long(f0), fluid.begin() + long(f1));
                Tiles t;
                // as large as the LDS allows, but small enough that a
                // segment spreads over about g_tile_target workgroups (tiles
                // of at least kMinTileElems elements): a small mesh in a few
                // LDS-sized tiles would run its apply on a few CUs
                constexpr int kMinTileElems = 512;
                const int64_t target = g_tile_target.load();
                const int occupancy =
                    std::max(kMinTileElems, int((int64_t(part.size()) + target - 1) / target));
                int tmax = std::min(std::max(64, lmax_allowed * 4), occupancy);
                while (!build_tiles(m, conn, elem_color, part, c->h_seg_node[size_t(sg)],
                                    c->h_seg_node[size_t(sg) + 1], c->ncolor, tmax, lmax_allowed, t))
                {
                    tmax /= 2;
                    if (tmax < 16) bad("cannot tile the mesh within the LDS budget");
                }
                append_tiles(tiles, t);
                c->h_seg_tile.push_back(int32_t(tiles.node_ptr.size()) - 1);
                f0 = f1;
            }
            c->ntiles = int32_t(tiles.node_ptr.size()) - 1;
            c->tile_lmax = tiles.lmax;
            c->nshared = tiles.nshared;
        }
        std::vector<int32_t> elem_tslot(E, -1);
        for (size_t k = 0; k < tiles.order.size(); ++k) elem_tslot[tiles.order[k]] = int32_t(k);

        lap();  // tiles
        // node incidences in element order
        std::vector<int32_t> inc_ptr(size_t(N) + 1, 0), inc(size_t(E) * 4);
        for (int e = 0; e < E; ++e)
        {
            inc_ptr[conn[e].x + 1]++;
            inc_ptr[conn[e].y + 1]++;
            inc_ptr[conn[e].z + 1]++;
            inc_ptr[conn[e].w + 1]++;
        }
        for (int i = 0; i < N; ++i) inc_ptr[i + 1] += inc_ptr[i];
        {
            std::vector<int32_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
            for (int e = 0; e < E; ++e)
            {
                const int v[4] = {conn[e].x, conn[e].y, conn[e].z, conn[e].w};
                for (int a = 0; a < 4; ++a) inc[size_t(fill[v[a]]++)] = e * 4 + a;
            }
        }

        lap();  // incidences
        // pressure constraints: inactive nodes and fixed nodes (last writer)
        std::vector<double> fixed_val(N, 0.0);
        std::vector<int32_t> last_pos(N, -1);
        for (int i = 0; i < m->bc_fixed; ++i)
        {
            const int ip = m->bc_list[i] - 1;
            fixed_val[ip] = m->bc_values[i];
            last_pos[ip] = i;
        }
        std::vector<int32_t> nf_node, nf_mode, pbc_node;
        std::vector<double> nf_val, pbc_val;
        for (int i = 0; i < N; ++i)
        {
            if (!tf[i])
            {
                nf_node.push_back(i);
                nf_mode.push_back(0);
                nf_val.push_back(0.0);
            }
            else if (fixed[i])
            {
                nf_node.push_back(i);
                nf_mode.push_back(1);
                nf_val.push_back(fixed_val[i]);
            }
        }
        for (int i = 0; i < m->bc_fixed; ++i)
        {
            const int ip = m->bc_list[i] - 1;
            if (last_pos[ip] != i) continue;  // a later entry for the same node wins
            pbc_node.push_back(ip);
            pbc_val.push_back(m->bc_values[i]);
        }

        // velocity boundary programme (Boundary.cpp:275-508, 672-723; Steps.cpp:102-176)
        for (int ib = 0; ib < B; ++ib)
        {
            const int bc = m->iside[size_t(ib) * 6 + 5];
            if (!supported_bc(bc))
                throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "Boundary - unsupported BC_ID " + std::to_string(bc) +
                                                           " at boundary face " + std::to_string(ib + 1));
            for (int k = 0; k < 3; ++k)
            {
                const int ip = m->iside[size_t(ib) * 6 + k] - 1;
                if (ip < 0 || ip >= N) bad("boundary face node out of range");
            }
        }
        auto unit_normal = [&](int ib, double& nx, double& ny, double& nz) {
            const double* fn = m->face_norm + size_t(ib) * 4;
            const double area = fn[3];
            if (area <= 0.0 || !std::isfinite(area))
                throw CbsFailure(CBS_ERR_INVALID_GEOMETRY,
                                 "Boundary - invalid boundary-face area in face_norm(4,ib). "
                                 "Preprocess::getNormals must be called before boundary application.");
            nx = fn[0] / area;
            ny = fn[1] / area;
            nz = fn[2] / area;
        };
        std::vector<int32_t> vslot(N, -1);
        std::vector<int32_t> vnode, vflag;
        std::vector<double> vval;
        std::vector<std::vector<double>> vsym, vout;
        auto slot_of = [&](int ip) {
            if (vslot[ip] < 0)
            {
                vslot[ip] = int32_t(vnode.size());
                vnode.push_back(ip);
                vflag.push_back(0);
                vval.insert(vval.end(), {0.0, 0.0, 0.0});
                vsym.emplace_back();
                vout.emplace_back();
            }
            return vslot[ip];
        };
        std::vector<int32_t> vsrc;  // BC id of each slot's last prescription (0: none)
        auto prescribe = [&](int ip, double u, double v, double w, int src = 0) {
            const int s = slot_of(ip);
            vflag[s] |= 1;
            vval[3 * size_t(s)] = u;
            vval[3 * size_t(s) + 1] = v;
            vval[3 * size_t(s) + 2] = w;
            if (vsrc.size() <= size_t(s)) vsrc.resize(size_t(s) + 1, 0);
            vsrc[size_t(s)] = src;
        };
        for (int ib = 0; ib < B; ++ib)
        {
            const int bc = m->iside[size_t(ib) * 6 + 5];
            if (bc != 506) continue;
            double nx, ny, nz;
            unit_normal(ib, nx, ny, nz);
            for (int k = 0; k < 3; ++k)
            {
                const int s = slot_of(m->iside[size_t(ib) * 6 + k] - 1);
                vsym[s].insert(vsym[s].end(), {nx, ny, nz});
            }
        }
        for (int ib = 0; ib < B; ++ib)
        {
            const int bc = m->iside[size_t(ib) * 6 + 5];
            double nx = 0.0, ny = 0.0, nz = 0.0;
            if (bc == 511) unit_normal(ib, nx, ny, nz);
            for (int k = 0; k < 3; ++k)
            {
                const int ip = m->iside[size_t(ib) * 6 + k] - 1;
                const double y = m->coord[size_t(ip) * 3 + 1];
                if (bc == 500 || bc == 510) prescribe(ip, g.inlet_u, g.inlet_v, g.inlet_w);
                else if (bc == 501 || bc == 502 || bc == 530 || bc == 532 || bc == 901) prescribe(ip, 0.0, 0.0, 0.0);
                else if (bc == 503) prescribe(ip, 1.0, 0.0, 0.0);
                else if (bc == 507) prescribe(ip, bfs_profile_u(y), 0.0, 0.0);
                else if (bc == 508) prescribe(ip, rectangular_profile_u(y), 0.0, 0.0, 508);
                else if (bc == 511)
                    prescribe(ip, -g.inlet_u_from_massflow * nx, -g.inlet_u_from_massflow * ny,
                              -g.inlet_u_from_massflow * nz);
            }
        }
        for (int iw = 0; iw < m->npoin_wall; ++iw)
        {
            const int ip = m->wall_node_list[iw] - 1;
            if (ip >= 0 && ip < N) prescribe(ip, 0.0, 0.0, 0.0);
        }
        for (int ib = 0; ib < B; ++ib)
        {
            if (m->iside[size_t(ib) * 6 + 5] != 500) continue;
            for (int k = 0; k < 3; ++k) prescribe(m->iside[size_t(ib) * 6 + k] - 1, g.inlet_u, g.inlet_v, g.inlet_w);
        }
        for (int ip = 0; ip < N; ++ip)
        {
            if (!ts[ip]) continue;
            prescribe(ip, 0.0, 0.0, 0.0);
            vflag[slot_of(ip)] |= 2;
        }
        for (int ib = 0; ib < B; ++ib)
        {
            if (m->iside[size_t(ib) * 6 + 5] != 520) continue;
            double nx, ny, nz;
            unit_normal(ib, nx, ny, nz);
            for (int k = 0; k < 3; ++k)
            {
                const int s = slot_of(m->iside[size_t(ib) * 6 + k] - 1);
                vout[s].insert(vout[s].end(), {nx, ny, nz});
            }
        }
        // inflow controls (adjoint workload): slots whose final value is the
        // BC-508 profile, with their (y, z) on the inlet
        vsrc.resize(vnode.size(), 0);
        c->h_bcv_val = vval;
        c->h_inflow_slot.clear();
        c->h_inflow_yz.clear();
        for (size_t sl = 0; sl < vnode.size(); ++sl)
            if (vsrc[sl] == 508)
            {
                c->h_inflow_slot.push_back(int32_t(sl));
                c->h_inflow_yz.push_back(m->coord[size_t(vnode[sl]) * 3 + 1]);
                c->h_inflow_yz.push_back(m->coord[size_t(vnode[sl]) * 3 + 2]);
            }
        std::vector<int32_t> sym_ptr(vnode.size() + 1, 0), out_ptr(vnode.size() + 1, 0);
        std::vector<double> sym_n, out_n;
        for (size_t s = 0; s < vnode.size(); ++s)
        {
            sym_n.insert(sym_n.end(), vsym[s].begin(), vsym[s].end());
            out_n.insert(out_n.end(), vout[s].begin(), vout[s].end());
            sym_ptr[s + 1] = int32_t(sym_n.size() / 3);
            out_ptr[s + 1] = int32_t(out_n.size() / 3);
        }

        // temperature programme (Boundary.cpp:529-626): last writer per node
        std::vector<int32_t> tslot(N, -1), tnode;
        std::vector<double> tval;
        for (int ib = 0; ib < B; ++ib)
        {
            const int bc = m->iside[size_t(ib) * 6 + 5];
            double v;
            if (bc == 501) v = 1.0;
            else if (bc == 502 || bc == 503 || bc == 508) v = 0.0;
            else if (bc == 510 || bc == 511) v = g.inlet_temperature;
            else continue;
            for (int k = 0; k < 3; ++k)
            {
                const int ip = m->iside[size_t(ib) * 6 + k] - 1;
                if (tslot[ip] < 0)
                {
                    tslot[ip] = int32_t(tnode.size());
                    tnode.push_back(ip);
                    tval.push_back(v);
                }
                tval[size_t(tslot[ip])] = v;
            }
        }

        // BC-532 heat flux (EnergyAssembly.cpp:484-532): per node, in face order
        std::vector<int32_t> hf_slot(N, -1), hf_ptr{0};
        std::vector<double> hf_val;
        if (g.heat_flux_bc != 0.0)
        {
            std::vector<std::vector<double>> per;
            std::vector<int32_t> order;
            for (int ib = 0; ib < B; ++ib)
            {
                if (m->iside[size_t(ib) * 6 + 5] != 532) continue;
                const double area = m->face_norm[size_t(ib) * 4 + 3];
                if (area <= 0.0 || !std::isfinite(area))
                    throw CbsFailure(CBS_ERR_INVALID_GEOMETRY,
                                     "EnergyAssembly::assembleStep4Rhs - invalid heat-flux boundary face area");
                const double contribution = g.heat_flux_bc * area / 3.0;
                for (int k = 0; k < 3; ++k)
                {
                    const int ip = m->iside[size_t(ib) * 6 + k] - 1;
                    if (hf_slot[ip] < 0)
                    {
                        hf_slot[ip] = int32_t(per.size());
                        per.emplace_back();
                    }
                    per[size_t(hf_slot[ip])].push_back(contribution);
                }
            }
            for (auto& v : per)
            {
                hf_val.insert(hf_val.end(), v.begin(), v.end());
                hf_ptr.push_back(int32_t(hf_val.size()));
            }
            c->nhf = int32_t(per.size());
        }

        lap();  // boundary programmes
        // uploads
        c->conn = upload(c, conn);
        c->mat = upload(c, m->mat_elem, size_t(E));
        c->rho = upload(c, m->rho_e, size_t(E));
        c->mu = upload(c, m->mu_e, size_t(E));
        c->kcond = upload(c, m->k_e, size_t(E));
        c->rcp = upload(c, m->rho_cp_e, size_t(E));
        c->alpha = upload(c, m->alpha_e, size_t(E));
        c->qvol = upload(c, m->Qvol_e, size_t(E));
        c->dNkdx = upload(c, m->dNkdx, size_t(E) * 12);
        c->detJ = upload(c, m->detJ, size_t(E));
        c->annxf = upload(c, m->annxf, size_t(E) * 16);
        c->elcoe = upload(c, m->elcoe_e, size_t(E) * 4);
        c->alen_e = upload(c, m->alen_e, size_t(E));
        c->fedge = upload(c, fe);
        c->pdiagE = dalloc<double>(c, size_t(E) * 4);
        c->gstifE = dalloc<double>(c, size_t(E) * 6);
        c->delte = dalloc<double>(c, size_t(E));
        c->slot_elem = upload(c, slot_elem);
        c->elem_slot = upload(c, elem_slot);
        c->slot_conn = upload(c, slot_conn);
        c->gstifE_s = dalloc<double>(c, size_t(c->F) * 6);
        c->gstif_s = dalloc<double>(c, size_t(c->F) * 6);
        c->tnode_ptr = upload(c, tiles.node_ptr);
        c->tnodes = upload(c, tiles.nodes);
        c->tsidx = upload(c, tiles.sidx);
        c->tcol_ptr = upload(c, tiles.col_ptr);
        c->tslot_elem = upload(c, tiles.order);
        c->elem_tslot = upload(c, elem_tslot);
        c->snode = upload(c, tiles.snode);
        c->lidx = upload(c, tiles.lidx);
        c->gstifE_t = dalloc<double>(c, size_t(c->F) * 6);
        c->gstif_t = dalloc<double>(c, size_t(c->F) * 6);
        c->sb = dalloc<double>(c, size_t(c->nshared) * size_t(c->ncolor));
        lap();  // uploads
        // segments: block maps of the reductions
        build_blocks(c, c->h_seg_node, 0, c->nblk);
        build_blocks(c, c->h_seg_elem, 0, c->eblk);
        // the assembled operator of the PCG apply, on the element tiles' rows
        {
            SellSource src;
            src.N = N;
            src.F = c->F;
            src.ntiles = c->ntiles;
            src.nseg = c->nseg;
            src.tile_node_ptr = tiles.node_ptr.data();
            src.tile_nodes = tiles.nodes.data();
            src.seg_tile = c->h_seg_tile.data();
            src.seg_node = c->h_seg_node.data();
            src.conn = conn.data();
            src.inc_ptr = inc_ptr.data();
            src.inc = inc.data();
            src.elem_tslot = elem_tslot.data();
            src.nflag = nflag.data();
            int lds = 65536;
            (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
            src.lds_bytes = lds - 1024;
            SellHost h;
            std::string why;
            if (!build_sell_host(src, h, why)) bad(why);
            cbs_hip_sell& m = c->sell;
            m.ntiles = int32_t(h.row_cnt.size());
            m.lmax = h.lmax;
            m.max_slices = h.max_slices;
            m.nnz = h.nnz;
            m.entries = h.slice_off.back();
            m.nstage = int64_t(h.rows.size() + h.halo.size());
            m.ncontrib = int64_t(h.cidx.size());
            m.row_base = upload(c, h.row_base);
            m.row_cnt = upload(c, h.row_cnt);
            m.halo_ptr = upload(c, h.halo_ptr);
            m.halo = upload(c, h.halo);
            m.gid = upload(c, h.gid);
            std::vector<uint8_t> sflag(static_cast<size_t>(N));
            for (int i = 0; i < N; ++i) sflag[size_t(i)] = nflag[size_t(h.gid[size_t(i)])];
            c->sv_flag = upload(c, sflag);
            m.slice_ptr = upload(c, h.slice_ptr);
            m.slice_off = upload(c, h.slice_off);
            m.cols = upload(c, h.cols);
            m.cptr = upload(c, h.cptr);
            m.cidx = upload(c, h.cidx);
            m.v64 = dalloc<double>(c, size_t(m.entries));
            m.v32 = dalloc<float>(c, size_t(m.entries));
            m.base64 = dalloc<double>(c, size_t(m.entries));
            m.h_seg_tile = h.seg_tile;
            m.seg_tile_ptr = upload(c, m.h_seg_tile);
            {
                std::vector<int64_t> se(size_t(c->nseg) + 1);
                std::vector<int32_t> e0(size_t(c->nseg));
                for (int sg = 0; sg <= c->nseg; ++sg)
                    se[size_t(sg)] = h.slice_off[size_t(h.slice_ptr[size_t(m.h_seg_tile[size_t(sg)])])];
                for (int sg = 0; sg < c->nseg; ++sg) e0[size_t(sg)] = c->h_seg_elem[size_t(sg)];
                m.seg_entry = upload(c, se);
                m.seg_elem0 = upload(c, e0);
                m.dt_uniform = dalloc<int32_t>(c, size_t(c->nseg));
            }
            if (c->nseg > 1)
            {
                std::vector<int32_t> tseg(size_t(std::max(m.ntiles, 1)), 0);
                for (int sg = 0; sg < c->nseg; ++sg)
                    for (int t = m.h_seg_tile[size_t(sg)]; t < m.h_seg_tile[size_t(sg) + 1]; ++t) tseg[size_t(t)] = sg;
                m.tile_seg = upload(c, tseg);
            }
        }
        if (c->nseg > 1)
        {
            std::vector<uint16_t> son(N), soe(E);
            for (int sg = 0; sg < c->nseg; ++sg)
            {
                for (int i = c->h_seg_node[size_t(sg)]; i < c->h_seg_node[size_t(sg) + 1]; ++i) son[size_t(i)] = uint16_t(sg);
                for (int e = c->h_seg_elem[size_t(sg)]; e < c->h_seg_elem[size_t(sg) + 1]; ++e) soe[size_t(e)] = uint16_t(sg);
            }
            c->seg_of_node = upload(c, son);
            c->seg_of_elem = upload(c, soe);
            c->seg_ani = upload(c, c->h_seg_ani);
        }
        configure_tiled_apply(c);
        c->dotpart = dalloc<double>(c, size_t(std::max(c->ndotpart, 1)));
        c->inc_ptr = upload(c, inc_ptr);
        c->inc = upload(c, inc);
        {
            // node-sum slots: every incidence in element order (inc), and
            // fluid incidences in slot, i.e. (colour, element), order
            std::vector<int32_t> cptr(size_t(N) + 1, 0), cinc(size_t(c->F) * 4);
            for (int k = 0; k < c->F; ++k)
            {
                const int4 n = conn[size_t(slot_elem[size_t(k)])];
                cptr[size_t(n.x) + 1]++;
                cptr[size_t(n.y) + 1]++;
                cptr[size_t(n.z) + 1]++;
                cptr[size_t(n.w) + 1]++;
            }
            for (int i = 0; i < N; ++i) cptr[size_t(i) + 1] += cptr[size_t(i)];
            std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
            for (int k = 0; k < c->F; ++k)
            {
                const int e = slot_elem[size_t(k)];
                const int4 n = conn[size_t(e)];
                const int v[4] = {n.x, n.y, n.z, n.w};
                for (int a = 0; a < 4; ++a) cinc[size_t(fill[size_t(v[a])]++)] = e * 4 + a;
            }
            lap();  // assembled operator
            build_node_slots(c, E, inc_ptr, inc, c->slots_e, E >= g_line_group_min.load());
            build_node_slots(c, E, cptr, cinc, c->slots_c, false);
            std::vector<int32_t> fl(slot_elem.begin(), slot_elem.end());
            std::sort(fl.begin(), fl.end());
            c->elem_fluid = upload(c, fl);
            lap();  // node slots
        }
        c->nflag = upload(c, nflag);
        const size_t n3 = size_t(N) * 3;
        c->unkno = dalloc<double>(c, n3);
        c->unkn1 = dalloc<double>(c, n3);
        c->rhs = dalloc<double>(c, n3);
        c->unit_vec = dalloc<double>(c, n3);
        for (double** p : {&c->pres, &c->pres1, &c->temp, &c->temp1, &c->rhs1, &c->deltp, &c->deltp1, &c->deltp2,
                           &c->beta1, &c->vvis, &c->velcp, &c->alen, &c->alen_static, &c->elcoe2, &c->elcoe2p,
                           &c->pdiag,
                           &c->velocity, &c->cg_r, &c->cg_p, &c->cg_p2, &c->cg_ap, &c->tmp_x, &c->sv_x, &c->sv_b,
                           &c->sv_pdiag})
            *p = dalloc<double>(c, size_t(N));
        for (float** p : {&c->f_r, &c->f_p, &c->f_p2, &c->f_ap, &c->f_dinv, &c->f_z}) *p = dalloc<float>(c, size_t(N));
        c->partial = dalloc<double>(c, size_t(std::max(c->nblk.n, c->eblk.n)) * 16);
        c->nnonfree = int32_t(nf_node.size());
        c->nf_node = upload(c, nf_node);
        c->nf_mode = upload(c, nf_mode);
        c->nf_val = upload(c, nf_val);
        c->npres_bc = int32_t(pbc_node.size());
        c->pbc_node = upload(c, pbc_node);
        c->pbc_val = upload(c, pbc_val);
        c->nbcv = int32_t(vnode.size());
        c->bcv_node = upload(c, vnode);
        c->bcv_flags = upload(c, vflag);
        c->bcv_val = upload(c, vval);
        c->bcv_sym_ptr = upload(c, sym_ptr);
        c->bcv_sym_n = upload(c, sym_n);
        c->bcv_out_ptr = upload(c, out_ptr);
        c->bcv_out_n = upload(c, out_n);
        c->ntbc = int32_t(tnode.size());
        c->tbc_node = upload(c, tnode);
        c->tbc_val = upload(c, tval);
        c->hf_slot = upload(c, hf_slot);
        c->hf_ptr = upload(c, hf_ptr);
        c->hf_val = upload(c, hf_val);

        const size_t ns = size_t(c->nseg);
        c->scal = dalloc<DevScalars>(c, ns);
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_scal), sizeof(DevScalars) * ns, hipHostMallocDefault),
                  "hipHostMalloc");
        for (size_t k = 0; k < ns; ++k)
        {
            c->h_scal[k] = DevScalars{};
            c->h_scal[k].err = ~0ull;
            c->h_scal[k].pending_err = ~0ull;
        }
        hip_check(hipMemcpyAsync(c->scal, c->h_scal, sizeof(DevScalars) * ns, hipMemcpyHostToDevice, c->stream),
                  "scalars");
        c->seg_result.assign(ns, cbs_cg_result{});
        c->seg_rtime.assign(ns, 0.0);
        launch_static_alen(c);
        hip_check(hipGetLastError(), "launch");
        hip_check(hipStreamSynchronize(c->stream), "create sync");
        lap();  // node vectors, boundary uploads, static element length
    }

    // Disjoint union of the members of a batch: node and element ids of
    // member k are offset by the totals of members 0..k-1; colour c of the
    // union is the members' colour-c lists in member order (elements of
    // different members never share a node, so each node keeps its member's
    // colour-order sequence of additions).
    struct UnionMesh
    {
        cbs_mesh_desc d{};
        std::vector<int32_t> intma, mat, fedge, iside, wall, bcl, cptr, celem;
        std::vector<double> coord, rho, mu, k, rcp, alpha, qvol, dN, detJ, annxf, elcoe, alen, fnorm, bcv;
    };

    template <typename T>
    void cat(std::vector<T>& dst, const T* src, size_t n)
    {
        if (n) dst.insert(dst.end(), src, src + n);
    }

    void build_union(int nm, const cbs_mesh_desc* ms, UnionMesh& u, std::vector<int32_t>& node_off,
                     std::vector<int32_t>& elem_off)
    {
        node_off.assign(1, 0);
        elem_off.assign(1, 0);
        int ncolor = 0;
        for (int k = 0; k < nm; ++k)
        {
            const cbs_mesh_desc& m = ms[k];
            const int N = m.npoin, E = m.nelem, B = m.nboun;
            if (N < 1 || E < 1 || B < 0 || m.ncolor < 1 || m.bc_fixed < 0 || m.npoin_wall < 0)
                throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_create_batch - invalid sizes of member " + std::to_string(k));
            if (!m.intma || !m.coord || !m.mat_elem || !m.rho_e || !m.mu_e || !m.k_e || !m.rho_cp_e || !m.alpha_e ||
                !m.Qvol_e || !m.dNkdx || !m.detJ || !m.annxf || !m.fedge || !m.elcoe_e || !m.alen_e ||
                (B > 0 && (!m.iside || !m.face_norm)) || !m.color_ptr || !m.color_elem ||
                (m.bc_fixed > 0 && (!m.bc_list || !m.bc_values)) || (m.npoin_wall > 0 && !m.wall_node_list))
                throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_create_batch - null array in member " + std::to_string(k));
            const int no = node_off.back(), eo = elem_off.back();
            if (int64_t(no) + N > INT32_MAX || int64_t(eo) + E > INT32_MAX)
                throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_create_batch - batch too large");
            for (size_t i = 0; i < size_t(E) * 4; ++i) u.intma.push_back(m.intma[i] + no);
            cat(u.coord, m.coord, size_t(N) * 3);
            cat(u.mat, m.mat_elem, size_t(E));
            cat(u.rho, m.rho_e, size_t(E));
            cat(u.mu, m.mu_e, size_t(E));
            cat(u.k, m.k_e, size_t(E));
            cat(u.rcp, m.rho_cp_e, size_t(E));
            cat(u.alpha, m.alpha_e, size_t(E));
            cat(u.qvol, m.Qvol_e, size_t(E));
            cat(u.dN, m.dNkdx, size_t(E) * 12);
            cat(u.detJ, m.detJ, size_t(E));
            cat(u.annxf, m.annxf, size_t(E) * 16);
            cat(u.fedge, m.fedge, size_t(E) * 4);
            cat(u.elcoe, m.elcoe_e, size_t(E) * 4);
            cat(u.alen, m.alen_e, size_t(E));
            for (int b = 0; b < B; ++b)
            {
                const int32_t* f = m.iside + size_t(b) * 6;
                u.iside.insert(u.iside.end(), {f[0] + no, f[1] + no, f[2] + no, f[3], f[4] + eo, f[5]});
            }
            cat(u.fnorm, m.face_norm, size_t(B) * 4);
            for (int i = 0; i < m.npoin_wall; ++i) u.wall.push_back(m.wall_node_list[i] + no);
            for (int i = 0; i < m.bc_fixed; ++i) u.bcl.push_back(m.bc_list[i] + no);
            cat(u.bcv, m.bc_values, size_t(m.bc_fixed));
            u.d.nboun += B;
            u.d.npoin_wall += m.npoin_wall;
            u.d.bc_fixed += m.bc_fixed;
            ncolor = std::max(ncolor, m.ncolor);
            node_off.push_back(no + N);
            elem_off.push_back(eo + E);
        }
        u.cptr.assign(1, 0);
        for (int col = 0; col < ncolor; ++col)
        {
            for (int k = 0; k < nm; ++k)
            {
                const cbs_mesh_desc& m = ms[k];
                if (col >= m.ncolor) continue;
                for (int q = m.color_ptr[col]; q < m.color_ptr[col + 1]; ++q)
                    u.celem.push_back(m.color_elem[q] + elem_off[size_t(k)]);
            }
            u.cptr.push_back(int32_t(u.celem.size()));
        }
        cbs_mesh_desc& d = u.d;
        d.npoin = node_off.back();
        d.nelem = elem_off.back();
        d.ncolor = ncolor;
        d.intma = u.intma.data();
        d.coord = u.coord.data();
        d.mat_elem = u.mat.data();
        d.rho_e = u.rho.data();
        d.mu_e = u.mu.data();
        d.k_e = u.k.data();
        d.rho_cp_e = u.rcp.data();
        d.alpha_e = u.alpha.data();
        d.Qvol_e = u.qvol.data();
        d.dNkdx = u.dN.data();
        d.detJ = u.detJ.data();
        d.annxf = u.annxf.data();
        d.fedge = u.fedge.data();
        d.elcoe_e = u.elcoe.data();
        d.alen_e = u.alen.data();
        d.iside = u.iside.data();
        d.face_norm = u.fnorm.data();
        d.wall_node_list = u.wall.data();
        d.bc_list = u.bcl.data();
        d.bc_values = u.bcv.data();
        d.color_ptr = u.cptr.data();
        d.color_elem = u.celem.data();
    }

    // The configurations of a batch's members may differ in ani and re (the
    // Reynolds number of each case), pnode and ntime only.
    void check_member_configs(int nm, const cbs_config* cfgs)
    {
        for (int k = 1; k < nm; ++k)
        {
            cbs_config a = cfgs[k], b = cfgs[0];
            a.ani = b.ani = 0.0;
            a.re = b.re = 0.0;
            a.pnode = b.pnode = 0;
            a.ntime = b.ntime = 0;
            if (std::memcmp(&a, &b, sizeof a) != 0)
                throw CbsFailure(CBS_ERR_BAD_ARGUMENT,
                                 "cbs_hip_create_batch - member " + std::to_string(k) +
                                     " differs from member 0 in a run control other than ani, re, pnode, ntime");
            if (!(cfgs[k].ani > 0.0) || !std::isfinite(cfgs[k].ani))
                throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_create_batch - invalid ani of member " + std::to_string(k));
        }
    }

    // Member (segment) of a union node or element index; 1-based in/out.
    std::string member_note(const cbs_hip_ctx* c, const std::string& msg)
    {
        if (c->nseg <= 1) return msg;
        for (const char* kind : {"node ", "element "})
        {
            const size_t at = msg.rfind(kind);
            if (at == std::string::npos) continue;
            const size_t p = at + std::strlen(kind);
            if (p >= msg.size() || !std::isdigit(static_cast<unsigned char>(msg[p]))) continue;
            const long idx = std::strtol(msg.c_str() + p, nullptr, 10) - 1;
            const std::vector<int32_t>& off = kind[0] == 'n' ? c->h_seg_node : c->h_seg_elem;
            const auto it = std::upper_bound(off.begin(), off.end(), int32_t(idx));
            const long sg = long(it - off.begin()) - 1;
            if (sg < 0 || sg >= c->nseg) break;
            return msg + " (batch member " + std::to_string(sg) + ", " + kind + std::to_string(idx - off[size_t(sg)] + 1) +
                   " of that member)";
        }
        return msg + " (batch context)";
    }

    void build_blocks(cbs_hip_ctx* c, const std::vector<int32_t>& off, int per_block, cbs_hip_blocks& b)
    {
        std::vector<int32_t> beg, end, seg;
        b.h_seg_ptr.assign(1, 0);
        for (size_t sg = 0; sg + 1 < off.size(); ++sg)
        {
            const int64_t lo = off[sg], len = int64_t(off[sg + 1]) - lo;
            if (len > 0)
            {
                const int64_t chunk = per_block > 0 ? per_block : (len + reduce_grid(len) - 1) / reduce_grid(len);
                for (int64_t k = lo; k < lo + len; k += chunk)
                {
                    beg.push_back(int32_t(k));
                    end.push_back(int32_t(std::min(lo + len, k + chunk)));
                    seg.push_back(int32_t(sg));
                }
            }
            b.h_seg_ptr.push_back(int32_t(beg.size()));
        }
        b.n = int32_t(beg.size());
        b.beg = upload(c, beg);
        b.end = upload(c, end);
        b.seg = upload(c, seg);
        b.seg_ptr = upload(c, b.h_seg_ptr);
    }

    struct FieldRef
    {
        double* ptr;
        int64_t count;
    };

    FieldRef field_ref(cbs_hip_ctx* c, int32_t f)
    {
        const int64_t N = c->N, E = c->E;
        switch (f)
        {
        case CBS_F_UNKNO: return {c->unkno, 3 * N};
        case CBS_F_UNKN1: return {c->unkn1, 3 * N};
        case CBS_F_PRES: return {c->pres, N};
        case CBS_F_PRES1: return {c->pres1, N};
        case CBS_F_TEMPERATURE: return {c->temp, N};
        case CBS_F_TEMPERATURE1: return {c->temp1, N};
        case CBS_F_RHS: return {c->rhs, 3 * N};
        case CBS_F_RHS1: return {c->rhs1, N};
        case CBS_F_DELTE: return {c->delte, E};
        case CBS_F_DELTP: return {c->deltp, N};
        case CBS_F_DELTP1: return {c->deltp1, N};
        case CBS_F_DELTP2: return {c->deltp2, N};
        case CBS_F_BETA1: return {c->beta1, N};
        case CBS_F_VVIS: return {c->vvis, N};
        case CBS_F_VELCP: return {c->velcp, N};
        case CBS_F_ALEN: return {c->alen, N};
        case CBS_F_UNKNO_UNIT_VEC: return {c->unit_vec, 3 * N};
        case CBS_F_ELCOE2: return {c->elcoe2, N};
        case CBS_F_ELCOE2P: return {c->elcoe2p, N};
        case CBS_F_PDIAG: return {c->pdiag, N};
        case CBS_F_GSTIF: return {nullptr, 6 * E};
        case CBS_F_GSTIFE: return {c->gstifE, 6 * E};
        case CBS_F_PDIAGE: return {c->pdiagE, 4 * E};
        case CBS_F_VELOCITY: return {c->velocity, N};
        default: throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "unknown field " + std::to_string(f));
        }
    }

    void sync(cbs_hip_ctx* c)
    {
        hip_check(hipGetLastError(), "kernel launch");
        hip_check(hipStreamSynchronize(c->stream), "stream sync");
    }

    int finish(cbs_hip_ctx* c)
    {
        sync(c);
        return check_device_error(c);
    }

    void copy_state(cbs_hip_ctx* c, double* dst, const double* src, size_t n)
    {
        hip_check(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream), "d2d copy");
    }

    int do_solve(cbs_hip_ctx* c, const cbs_cg_options* opts, cbs_cg_result* res)
    {
        const cbs_config& g = c->cfg;
        if (g.cg_max_iter < 1)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "ConjugateGradient::solvePressure - cg_max_iter must be positive");
        if (g.relToler <= 0.0 || g.absToler <= 0.0)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT,
                             "ConjugateGradient::solvePressure - relToler and absToler must be positive");
        if (g.cg_preconditioner < 0 || g.cg_preconditioner > 1)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "ConjugateGradient::solvePressure - cg_preconditioner must be 0 or 1");
        if (g.cg_conv_test < 1 || g.cg_conv_test > 3)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "ConjugateGradient::solvePressure - cg_conv_test must be 1, 2, or 3");
        bool any_active = false;
        for (int col = 0; col < c->ncolor && !any_active; ++col) any_active = c->h_fcolor_ptr[col + 1] > 0;
        if (!any_active)
            throw CbsFailure(CBS_ERR_INVALID_GEOMETRY,
                             "ConjugateGradient::solvePressure - no fluid-connected pressure nodes found");
        cbs_cg_result local{};
        cbs_cg_result* r = res ? res : &local;
        std::memset(r, 0, sizeof *r);
        const int prec = opts ? opts->precision : CBS_PCG_FP64;
        if (c->nseg > 1 && prec == CBS_PCG_AMG)
            throw CbsFailure(CBS_ERR_UNSUPPORTED, "batch contexts have no multigrid-preconditioned solve");
        int rc;
        if (prec == CBS_PCG_FP64) rc = solve_pressure_fp64(c, r);
        else if (prec == CBS_PCG_MIXED) rc = solve_pressure_mixed(c, opts, r);
        else if (prec == CBS_PCG_AMG) rc = solve_pressure_amg(c, r);
        else throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "unknown precision " + std::to_string(prec));
        return rc;
    }

    void validate_timestep_inputs(const cbs_hip_ctx* c)
    {
        const cbs_config& g = c->cfg;
        if (g.csafm <= 0.0 || !std::isfinite(g.csafm))
            throw CbsFailure(CBS_ERR_INVALID_TIMESTEP, "TimeStep - csafm must be positive");
        if (g.dtfix <= 0.0 || !std::isfinite(g.dtfix))
            throw CbsFailure(CBS_ERR_INVALID_TIMESTEP, "TimeStep - dtfix must be positive");
        if (g.htype < 1 || g.htype > 3) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "TimeStep - htype must be 1, 2, or 3");
        if (g.cbs_scheme != 1 && g.cbs_scheme != 0)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT,
                             "TimeStep::computeTimeStep - cbs_scheme must be 1 semi-implicit or 0 explicit");
        if (g.cbs_scheme == 0)
            throw CbsFailure(CBS_ERR_UNSUPPORTED,
                             "Fully explicit CBS3D mode was requested, but explicit CBS3D step kernels are not yet ported.");
    }

    int step2_body(cbs_hip_ctx* c, const cbs_cg_options* opts, cbs_cg_result* res)
    {
        const cbs_config& g = c->cfg;
        if (g.solver_opt == 2)
            throw CbsFailure(CBS_ERR_UNSUPPORTED,
                             "Steps::step2SemiImplicit - banded pressure solver path is not ported for CBS3D yet");
        if (g.solver_opt == 3)
            throw CbsFailure(CBS_ERR_UNSUPPORTED,
                             "Steps::step2SemiImplicit - solver_opt=3 requests PETSc, but this executable was built without CBS3D_USE_PETSC");
        if (g.solver_opt != 1)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "Steps::step2SemiImplicit - unknown pressure solver option");
        launch_step2_rhs(c);
        cbs_cg_result local{};
        cbs_cg_result* r = res ? res : &local;
        int rc = do_solve(c, opts, r);
        if (rc != CBS_OK) return rc;
        c->last_cg_iterations = r->iterations;
        c->last_cg_initial_l2 = r->initial_l2;
        c->last_cg_final_l2 = r->final_l2;
        c->last_cg_relative_l2 = r->final_relative_l2;
        c->last_cg_max_abs = r->final_max_abs;
        if (!r->converged)
        {
            std::string msg = "Steps::step2SemiImplicit - pressure solver failed to converge";
            if (c->nseg > 1)
                for (int k = 0; k < c->nseg; ++k)
                    if (!c->seg_result[size_t(k)].converged)
                    {
                        msg += " (batch member " + std::to_string(k) + ")";
                        break;
                    }
            throw CbsFailure(CBS_ERR_CG_NOT_CONVERGED, msg);
        }
        launch_bc(c, 3);
        return CBS_OK;
    }
}

namespace cbsdev
{
    void hip_check(hipError_t e, const char* what)
    {
        if (e != hipSuccess) throw HipFailure(std::string(what) + ": " + hipGetErrorString(e));
    }

    int set_error(cbs_hip_ctx* c, int status, const std::string& msg)
    {
        c->err = msg;
        return status;
    }

    int check_device_error(cbs_hip_ctx* c)
    {
        // on the context's stream: its kernels raise the errors, and a
        // default-stream copy is not ordered with a non-blocking stream
        unsigned long long key = ~0ull;
        hip_check(hipMemcpyAsync(&key, &c->scal->err, sizeof key, hipMemcpyDeviceToHost, c->stream), "read error");
        hip_check(hipStreamSynchronize(c->stream), "read error");
        if (key == ~0ull) return CBS_OK;
        int status = CBS_ERR_INVALID_GEOMETRY;
        c->err = member_note(c, message_for(key, &status));
        const unsigned long long none = ~0ull;
        hip_check(hipMemcpyAsync(&c->scal->err, &none, sizeof none, hipMemcpyHostToDevice, c->stream), "reset error");
        hip_check(hipStreamSynchronize(c->stream), "reset error");
        return status;
    }
}

extern "C" {

const char* cbs_hip_last_error(const cbs_hip_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int64_t cbs_hip_field_size(const cbs_hip_ctx* ctx, int32_t field)
{
    if (!ctx) return -1;
    try
    {
        return field_ref(const_cast<cbs_hip_ctx*>(ctx), field).count;
    }
    catch (...)
    {
        return -1;
    }
}

int cbs_hip_create(const cbs_config* cfg, const cbs_mesh_desc* mesh, cbs_hip_ctx** out)
{
    if (!cfg || !mesh || !out)
    {
        g_create_error = "cbs_hip_create - null argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    *out = nullptr;
    auto* c = new cbs_hip_ctx();
    c->cfg = *cfg;
    int rc = guarded(c, [&] {
        hip_check(hipGetDevice(&c->device), "hipGetDevice");
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        build_context(c, mesh);
        return int(CBS_OK);
    });
    if (rc != CBS_OK)
    {
        g_create_error = c->err;
        cbs_hip_destroy(c);
        return rc;
    }
    *out = c;
    return CBS_OK;
}

int cbs_hip_create_batch(int32_t nmember, const cbs_config* cfgs, const cbs_mesh_desc* meshes, cbs_hip_ctx** out)
{
    if (!cfgs || !meshes || !out || nmember < 1 || nmember > 65535)
    {
        g_create_error = "cbs_hip_create_batch - null argument or member count outside 1..65535";
        return CBS_ERR_BAD_ARGUMENT;
    }
    *out = nullptr;
    auto* c = new cbs_hip_ctx();
    c->cfg = cfgs[0];
    int rc = guarded(c, [&] {
        check_member_configs(nmember, cfgs);
        UnionMesh u;
        build_union(nmember, meshes, u, c->h_seg_node, c->h_seg_elem);
        c->nseg = nmember;
        c->h_seg_ani.resize(size_t(nmember));
        for (int k = 0; k < nmember; ++k) c->h_seg_ani[size_t(k)] = cfgs[k].ani;
        hip_check(hipGetDevice(&c->device), "hipGetDevice");
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        build_context(c, &u.d);
        return int(CBS_OK);
    });
    if (rc != CBS_OK)
    {
        g_create_error = c->err;
        cbs_hip_destroy(c);
        return rc;
    }
    *out = c;
    return CBS_OK;
}

int cbs_hip_member_count(const cbs_hip_ctx* c, int32_t* n)
{
    if (!c || !n) return CBS_ERR_BAD_ARGUMENT;
    *n = c->nseg;
    return CBS_OK;
}

int cbs_hip_member_range(const cbs_hip_ctx* c, int32_t member, int32_t* node_offset, int32_t* npoin,
                         int32_t* elem_offset, int32_t* nelem)
{
    if (!c || member < 0 || member >= c->nseg) return CBS_ERR_BAD_ARGUMENT;
    const size_t k = size_t(member);
    if (node_offset) *node_offset = c->h_seg_node[k];
    if (npoin) *npoin = c->h_seg_node[k + 1] - c->h_seg_node[k];
    if (elem_offset) *elem_offset = c->h_seg_elem[k];
    if (nelem) *nelem = c->h_seg_elem[k + 1] - c->h_seg_elem[k];
    return CBS_OK;
}

int cbs_hip_get_member_scalars(cbs_hip_ctx* c, int32_t member, cbs_scalars* out)
{
    return guarded(c, [&] {
        if (!out || member < 0 || member >= c->nseg) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "member or output");
        DevScalars h;
        hip_check(hipStreamSynchronize(c->stream), "sync");
        hip_check(hipMemcpy(&h, c->scal + member, sizeof h, hipMemcpyDeviceToHost), "scalars");
        const cbs_cg_result& r = c->seg_result[size_t(member)];
        out->dtreal = h.dtreal;
        out->rtime = c->seg_rtime[size_t(member)];
        for (int i = 0; i < 15; ++i) out->hb[i] = h.hb[i];
        out->last_cg_iterations = r.iterations;
        out->last_cg_initial_l2 = r.initial_l2;
        out->last_cg_final_l2 = r.final_l2;
        out->last_cg_relative_l2 = r.final_relative_l2;
        out->last_cg_max_abs = r.final_max_abs;
        return int(CBS_OK);
    });
}

int cbs_hip_get_member_cg_result(cbs_hip_ctx* c, int32_t member, cbs_cg_result* out)
{
    return guarded(c, [&] {
        if (!out || member < 0 || member >= c->nseg) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "member or output");
        *out = c->seg_result[size_t(member)];
        return int(CBS_OK);
    });
}

int cbs_hip_destroy(cbs_hip_ctx* c)
{
    if (!c) return CBS_ERR_BAD_ARGUMENT;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (cbs_snapshot& sl : c->snap)
    {
        if (sl.copied) (void)hipEventDestroy(sl.copied);
        if (sl.ready) (void)hipEventDestroy(sl.ready);
        if (sl.h) (void)hipHostFree(sl.h);
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    destroy_adjoint(c);
    destroy_amg(c);
    destroy_graphs(c);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    for (void* p : c->allocations) (void)hipFree(p);
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CBS_OK;
}

int cbs_hip_set_field(cbs_hip_ctx* c, int32_t field, const double* host, int64_t count)
{
    return guarded(c, [&] {
        const FieldRef f = field_ref(c, field);
        if (!host || count != f.count) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_set_field - size mismatch");
        if (field == CBS_F_GSTIF)
        {
            // slot-ordered planes (field reads) and tile-ordered planes (the
            // assembled operator is rebuilt from them at every solve)
            const std::pair<int32_t*, double*> orders[2] = {{c->slot_elem, c->gstif_s}, {c->tslot_elem, c->gstif_t}};
            // copies on the context's stream, each complete before its host
            // buffer goes (the stream is non-blocking)
            for (const auto& o : orders)
            {
                std::vector<int32_t> slot(c->F);
                hip_check(hipMemcpyAsync(slot.data(), o.first, sizeof(int32_t) * c->F, hipMemcpyDeviceToHost,
                                         c->stream),
                          "slots");
                hip_check(hipStreamSynchronize(c->stream), "slots");
                std::vector<double> planes(size_t(c->F) * 6);
                for (int k = 0; k < c->F; ++k)
                    for (int p = 0; p < 6; ++p) planes[size_t(p) * c->F + k] = host[size_t(slot[k]) * 6 + p];
                hip_check(hipMemcpyAsync(o.second, planes.data(), planes.size() * 8, hipMemcpyHostToDevice, c->stream),
                          "gstif");
                hip_check(hipStreamSynchronize(c->stream), "gstif");
            }
            c->gstif_slots_stale = false;
            c->gstif_tiles_stale = false;
            c->gstif_manual = true;
            return int(CBS_OK);
        }
        hip_check(hipMemcpyAsync(f.ptr, host, size_t(count) * sizeof(double), hipMemcpyHostToDevice, c->stream),
                  "set field");
        hip_check(hipStreamSynchronize(c->stream), "set field sync");
        return int(CBS_OK);
    });
}

int cbs_hip_create_seconds(const cbs_hip_ctx* c, double* seconds, int32_t cap)
{
    if (!c || (cap > 0 && !seconds)) return -1;
    const int n = int(c->create_seconds.size());
    for (int k = 0; k < std::min(n, int(cap)); ++k) seconds[k] = c->create_seconds[size_t(k)];
    return n;
}

int cbs_hip_coloring(int32_t nelem, int32_t npoin, const int32_t* intma, int32_t* ncolor, int32_t* color_ptr,
                     int32_t* color_elem, int32_t* levels)
{
    if (!intma || !ncolor || !color_ptr || !color_elem || nelem < 1 || npoin < 1 || nelem > INT32_MAX / 4)
    {
        g_create_error = "cbs_hip_coloring - invalid argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    try
    {
        std::string msg;
        int32_t lv = 0;
        const int rc = colour_elements(nelem, npoin, intma, ncolor, color_ptr, color_elem, &lv, msg);
        if (levels) *levels = lv;
        if (rc != CBS_OK) g_create_error = msg;
        return rc;
    }
    catch (const std::exception& e)
    {
        g_create_error = e.what();
        return CBS_ERR_HIP;
    }
}

int32_t cbs_hip_abi_version(void)
{
    return CBS_HIP_ABI_VERSION;
}

int32_t cbs_hip_set_tile_target(int32_t tiles)
{
    if (tiles < 1) return -1;
    return g_tile_target.exchange(tiles);
}

int32_t cbs_hip_set_line_group_min_elems(int32_t elems)
{
    if (elems < 0) return -1;
    return g_line_group_min.exchange(elems);
}

int cbs_hip_coloring_noalloc(int32_t nelem, int32_t npoin, const int32_t* intma_dev, int32_t* ncolor,
                             int32_t* color_ptr_dev, int32_t* color_elem_dev, int32_t* levels)
{
    if (!intma_dev || !ncolor || !color_ptr_dev || !color_elem_dev || nelem < 1 || npoin < 1)
    {
        g_create_error = "cbs_hip_coloring_noalloc - invalid argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    try
    {
        std::string msg;
        int32_t lv = 0;
        const int rc = colour_elements_noalloc(nelem, npoin, intma_dev, ncolor, color_ptr_dev, color_elem_dev, &lv, msg);
        if (levels) *levels = lv;
        if (rc != CBS_OK) g_create_error = msg;
        return rc;
    }
    catch (const std::exception& e)
    {
        g_create_error = e.what();
        return CBS_ERR_HIP;
    }
}

int64_t cbs_hip_coloring_noalloc_scratch_bytes(void)
{
    return colour_noalloc_scratch_bytes();
}

// Output snapshots (the device side of Post::writeSolution, Post.cpp:842-860).
int cbs_hip_snapshot_begin(cbs_hip_ctx* c, cbs_snapshot** out)
{
    if (out) *out = nullptr;
    return guarded(c, [&] {
        if (!out) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_snapshot_begin - null output");
        if (c->nseg != 1)
            throw CbsFailure(CBS_ERR_UNSUPPORTED, "cbs_hip_snapshot_begin - batch contexts have no single solution");
        const size_t N = size_t(c->N);
        if (!c->copy_stream)
            hip_check(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
        cbs_snapshot* sl = nullptr;
        {
            // a slot frees when its writer releases it
            std::unique_lock<std::mutex> lk(c->snap_mu);
            c->snap_cv.wait(lk, [&] { return !c->snap[0].held || !c->snap[1].held; });
            sl = !c->snap[0].held ? &c->snap[0] : &c->snap[1];
            sl->held = true;
        }
        try
        {
            if (!sl->d)
            {
                sl->ctx = c;
                sl->d = dalloc<double>(c, 6 * N);
                hip_check(hipHostMalloc(reinterpret_cast<void**>(&sl->h), sizeof(double) * std::max<size_t>(6 * N, 1),
                                        hipHostMallocDefault),
                          "hipHostMalloc");
                hip_check(hipEventCreateWithFlags(&sl->copied, hipEventDisableTiming), "hipEventCreate");
                hip_check(hipEventCreateWithFlags(&sl->ready, hipEventDisableTiming), "hipEventCreate");
            }
            if (N > 0)
            {
                const std::pair<int32_t, size_t> parts[4] = {
                    {CBS_F_PRES, 0}, {CBS_F_TEMPERATURE, N}, {CBS_F_UNKNO, 2 * N}, {CBS_F_VELOCITY, 5 * N}};
                for (const auto& pt : parts)
                {
                    const FieldRef f = field_ref(c, pt.first);
                    hip_check(hipMemcpyAsync(sl->d + pt.second, f.ptr, size_t(f.count) * sizeof(double),
                                             hipMemcpyDeviceToDevice, c->stream),
                              "snapshot copy");
                }
            }
            hip_check(hipEventRecord(sl->copied, c->stream), "hipEventRecord");
            hip_check(hipStreamWaitEvent(c->copy_stream, sl->copied, 0), "hipStreamWaitEvent");
            if (N > 0)
                hip_check(hipMemcpyAsync(sl->h, sl->d, 6 * N * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream),
                          "snapshot download");
            hip_check(hipEventRecord(sl->ready, c->copy_stream), "hipEventRecord");
        }
        catch (...)
        {
            std::lock_guard<std::mutex> lk(c->snap_mu);
            sl->held = false;
            c->snap_cv.notify_all();
            throw;
        }
        *out = sl;
        return int(CBS_OK);
    });
}

int cbs_hip_snapshot_wait(void* snapshot)
{
    cbs_snapshot* sl = static_cast<cbs_snapshot*>(snapshot);
    if (!sl || !sl->held) return CBS_ERR_BAD_ARGUMENT;
    return hipEventSynchronize(sl->ready) == hipSuccess ? int(CBS_OK) : int(CBS_ERR_HIP);
}

const double* cbs_hip_snapshot_field(const cbs_snapshot* sl, int32_t field)
{
    if (!sl || !sl->held) return nullptr;
    const size_t N = size_t(sl->ctx->N);
    switch (field)
    {
    case CBS_F_PRES: return sl->h;
    case CBS_F_TEMPERATURE: return sl->h + N;
    case CBS_F_UNKNO: return sl->h + 2 * N;
    case CBS_F_VELOCITY: return sl->h + 5 * N;
    default: return nullptr;
    }
}

int cbs_hip_snapshot_release(void* snapshot)
{
    cbs_snapshot* sl = static_cast<cbs_snapshot*>(snapshot);
    if (!sl || !sl->held) return CBS_ERR_BAD_ARGUMENT;
    // the download may still run when the write was abandoned before waiting
    const int rc = hipEventSynchronize(sl->ready) == hipSuccess ? int(CBS_OK) : int(CBS_ERR_HIP);
    cbs_hip_ctx* c = sl->ctx;
    std::lock_guard<std::mutex> lk(c->snap_mu);
    sl->held = false;
    c->snap_cv.notify_all();
    return rc;
}

int cbs_hip_get_field(cbs_hip_ctx* c, int32_t field, double* host, int64_t count)
{
    return guarded(c, [&] {
        const FieldRef f = field_ref(c, field);
        if (!host || count != f.count) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_get_field - size mismatch");
        hip_check(hipStreamSynchronize(c->stream), "get field sync");
        if (field == CBS_F_GSTIF)
        {
            std::vector<int32_t> slot(c->F);
            std::vector<double> planes(size_t(c->F) * 6);
            hip_check(hipMemcpy(slot.data(), c->slot_elem, sizeof(int32_t) * c->F, hipMemcpyDeviceToHost), "slots");
            ensure_gstif_slots(c);
            hip_check(hipStreamSynchronize(c->stream), "gstif slots");
            hip_check(hipMemcpy(planes.data(), c->gstif_s, planes.size() * 8, hipMemcpyDeviceToHost), "gstif");
            std::memset(host, 0, sizeof(double) * size_t(count));
            for (int k = 0; k < c->F; ++k)
                for (int p = 0; p < 6; ++p) host[size_t(slot[k]) * 6 + p] = planes[size_t(p) * c->F + k];
            return int(CBS_OK);
        }
        hip_check(hipMemcpy(host, f.ptr, size_t(count) * sizeof(double), hipMemcpyDeviceToHost), "get field");
        return int(CBS_OK);
    });
}

int cbs_hip_get_scalars(cbs_hip_ctx* c, cbs_scalars* out)
{
    return guarded(c, [&] {
        if (!out) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "null output");
        DevScalars h;
        hip_check(hipStreamSynchronize(c->stream), "sync");
        hip_check(hipMemcpy(&h, c->scal, sizeof h, hipMemcpyDeviceToHost), "scalars");
        out->dtreal = h.dtreal;
        out->rtime = c->rtime;
        for (int i = 0; i < 15; ++i) out->hb[i] = h.hb[i];
        out->last_cg_iterations = c->last_cg_iterations;
        out->last_cg_initial_l2 = c->last_cg_initial_l2;
        out->last_cg_final_l2 = c->last_cg_final_l2;
        out->last_cg_relative_l2 = c->last_cg_relative_l2;
        out->last_cg_max_abs = c->last_cg_max_abs;
        return int(CBS_OK);
    });
}

int cbs_hip_set_dtreal(cbs_hip_ctx* c, double dtreal)
{
    return guarded(c, [&] {
        for (int k = 0; k < c->nseg; ++k)
            hip_check(hipMemcpyAsync(&c->scal[k].dtreal, &dtreal, sizeof dtreal, hipMemcpyHostToDevice, c->stream),
                      "dt");
        hip_check(hipStreamSynchronize(c->stream), "dt sync");
        return int(CBS_OK);
    });
}

int cbs_hip_build_element_pressure_terms(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_build_pressure_terms(c);
        return finish(c);
    });
}

int cbs_hip_compute_timestep(cbs_hip_ctx* c, int32_t iitime)
{
    return guarded(c, [&] {
        validate_timestep_inputs(c);
        launch_timestep(c, iitime);
        return finish(c);
    });
}

int cbs_hip_update_lhs_diagonal(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_lhs(c);
        return finish(c);
    });
}

int cbs_hip_assemble_step1_rhs(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_step1_rhs(c);
        return finish(c);
    });
}

int cbs_hip_step1(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        if (c->cfg.cbs_scheme == 0)
            throw CbsFailure(CBS_ERR_UNSUPPORTED,
                             "Steps::step1 - fully explicit CBS3D mode was requested, but explicit step kernels are not yet ported.");
        launch_step1_rhs(c);
        launch_step1_update(c);
        launch_bc(c, 0);
        return finish(c);
    });
}

int cbs_hip_assemble_step2_rhs(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_step2_rhs(c);
        return finish(c);
    });
}

int cbs_hip_pressure_multiply(cbs_hip_ctx* c, const double* x, double* y)
{
    return guarded(c, [&] {
        if (!x || !y) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "null vector");
        hip_check(hipMemcpyAsync(c->tmp_x, x, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream), "x");
        launch_fill(c, c->cg_ap, c->N, 0.0);
        launch_apply(c, c->tmp_x, c->cg_ap, &c->scal->pad0);
        launch_apply_finish(c, c->tmp_x, c->cg_ap);
        sync(c);
        hip_check(hipMemcpy(y, c->cg_ap, sizeof(double) * c->N, hipMemcpyDeviceToHost), "y");
        return check_device_error(c);
    });
}

int cbs_hip_pressure_multiply_device(cbs_hip_ctx* c, const double* x_dev, double* y_dev)
{
    return guarded(c, [&] {
        if (!x_dev || !y_dev) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "null vector");
        launch_fill(c, y_dev, c->N, 0.0);
        launch_apply(c, x_dev, y_dev, &c->scal->pad0);
        launch_apply_finish(c, x_dev, y_dev);
        return finish(c);
    });
}

int cbs_hip_solve_pressure(cbs_hip_ctx* c, const cbs_cg_options* opts, cbs_cg_result* result)
{
    return guarded(c, [&] { return do_solve(c, opts, result); });
}

int cbs_hip_solve_pressure_multi(cbs_hip_ctx* c, int32_t nrhs, const double* B, double* X,
                                 const cbs_cg_options* opts, cbs_cg_result* results)
{
    return guarded(c, [&] {
        if (nrhs < 1 || nrhs > 64 || !B || !X || !results)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_solve_pressure_multi - nrhs must be 1..64 with non-null arrays");
        if (c->nseg > 1)
            throw CbsFailure(CBS_ERR_UNSUPPORTED, "cbs_hip_solve_pressure_multi - not available on a batch context");
        // passes of at most kMultiMaxCols columns (column j at offset j*N)
        for (int j0 = 0; j0 < nrhs; j0 += kMultiMaxCols)
        {
            const int m = std::min(kMultiMaxCols, nrhs - j0);
            const size_t off = size_t(j0) * size_t(c->N);
            int rc = solve_pressure_multi(c, m, B + off, X + off, opts, results + j0);
            if (rc != CBS_OK) return rc;
            rc = check_device_error(c);
            if (rc != CBS_OK) return rc;
        }
        return int(CBS_OK);
    });
}

int cbs_hip_step2(cbs_hip_ctx* c, const cbs_cg_options* opts, cbs_cg_result* result)
{
    return guarded(c, [&] {
        int rc = step2_body(c, opts, result);
        if (rc != CBS_OK) return rc;
        return finish(c);
    });
}

int cbs_hip_step3(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_step3(c);
        launch_bc(c, 0);
        return finish(c);
    });
}

int cbs_hip_assemble_step4_rhs(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_step4_rhs(c);
        return finish(c);
    });
}

int cbs_hip_step4(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        if (c->cfg.temp_calc < 1) return int(CBS_OK);
        launch_step4_rhs(c);
        launch_step4_update(c);
        launch_bc(c, 2);
        return finish(c);
    });
}

int cbs_hip_apply_boundary(cbs_hip_ctx* c, int32_t which)
{
    return guarded(c, [&] {
        if (which < 0 || which > 3) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "unknown boundary application");
        launch_bc(c, which);
        return finish(c);
    });
}

int cbs_hip_update_velocity_magnitude(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_velocity_magnitude(c);
        return finish(c);
    });
}

int cbs_hip_convergence(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        launch_convergence(c);
        return finish(c);
    });
}

int cbs_hip_bandwidth_probe(cbs_hip_ctx* c, int64_t bytes, int32_t reps, double* gbs)
{
    return guarded(c, [&] {
        if (bytes < 1024 || reps < 1 || !gbs) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "bandwidth probe arguments");
        const int64_t n = bytes / 8;
        double *a = nullptr, *b = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void**>(&a), size_t(n) * 8), "probe alloc");
        hipError_t e2 = hipMalloc(reinterpret_cast<void**>(&b), size_t(n) * 8);
        if (e2 != hipSuccess)
        {
            (void)hipFree(a);
            hip_check(e2, "probe alloc");
        }
        hipEvent_t t0, t1;
        hip_check(hipEventCreate(&t0), "event");
        hip_check(hipEventCreate(&t1), "event");
        launch_fill(c, a, n, 1.0);
        launch_copy(c, b, a, n);  // warm-up
        hip_check(hipEventRecord(t0, c->stream), "event");
        for (int r = 0; r < reps; ++r) launch_copy(c, b, a, n);
        hip_check(hipEventRecord(t1, c->stream), "event");
        hip_check(hipEventSynchronize(t1), "event sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, t0, t1), "elapsed");
        *gbs = 2.0 * double(n) * 8.0 * reps / (double(ms) * 1e-3) / 1e9;
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        (void)hipFree(a);
        (void)hipFree(b);
        return int(CBS_OK);
    });
}

int cbs_hip_info(const cbs_hip_ctx* c, cbs_info* out)
{
    if (!c || !out) return CBS_ERR_BAD_ARGUMENT;
    out->npoin = c->N;
    out->nelem = c->E;
    out->nfluid = c->F;
    out->ncolor = c->ncolor;
    out->device_bytes = int64_t(c->device_bytes);
    return CBS_OK;
}

int cbs_hip_synchronize(cbs_hip_ctx* c)
{
    return guarded(c, [&] {
        sync(c);
        return int(CBS_OK);
    });
}

int cbs_hip_set_profiling(cbs_hip_ctx* c, int32_t on)
{
    return guarded(c, [&] {
        sync(c);
        timer_collect(c);
        c->profile = on != 0;
        return int(CBS_OK);
    });
}

int cbs_hip_set_apply_config(cbs_hip_ctx* c, int32_t precision, int32_t threads, int32_t batch)
{
    return guarded(c, [&] {
        if ((threads != 256 && threads != 512 && threads != 1024) || (batch != 4 && batch != 8))
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_set_apply_config - threads must be 256, 512 or 1024 and "
                                                   "batch 4 or 8");
        sync(c);
        sell_set_config(c, precision == CBS_PCG_MIXED ? 1 : 0, threads, batch);
        destroy_graphs(c);  // captured blocks hold the old launch shape
        return int(CBS_OK);
    });
}

int cbs_hip_direction_fusion(cbs_hip_ctx* c, int32_t precision, int32_t* fused, float* ms_separate, float* ms_fused)
{
    return guarded(c, [&] {
        const cbs_hip_sell::Config& tc = c->sell.cfg[precision == CBS_PCG_MIXED ? 1 : 0];
        if (fused) *fused = tc.fuse;
        if (ms_separate) *ms_separate = tc.fuse_ms[0];
        if (ms_fused) *ms_fused = tc.fuse_ms[1];
        return int(CBS_OK);
    });
}

int cbs_hip_set_direction_fusion(cbs_hip_ctx* c, int32_t precision, int32_t mode)
{
    return guarded(c, [&] {
        if (mode < -1 || mode > 1)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_set_direction_fusion - mode must be -1, 0 or 1");
        sync(c);
        cbs_hip_sell::Config& tc = c->sell.cfg[precision == CBS_PCG_MIXED ? 1 : 0];
        tc.fuse_pin = mode;
        tc.fuse = mode;
        tc.fuse_ms[0] = tc.fuse_ms[1] = 0.0f;
        destroy_graphs(c);  // captured blocks hold the old iteration
        return int(CBS_OK);
    });
}

int cbs_hip_amg_info(const cbs_hip_ctx* c, double* setup_seconds, int32_t* levels, int64_t* stored)
{
    if (!c) return CBS_ERR_BAD_ARGUMENT;
    if (setup_seconds) *setup_seconds = amg_setup_seconds(c);
    if (levels) *levels = amg_levels(c);
    if (stored) *stored = amg_stored(c);
    return CBS_OK;
}

int cbs_hip_set_graphs(cbs_hip_ctx* c, int32_t on)
{
    return guarded(c, [&] {
        sync(c);
        c->use_graphs = on != 0;
        if (!c->use_graphs) destroy_graphs(c);
        return int(CBS_OK);
    });
}

int cbs_hip_graphs_active(const cbs_hip_ctx* c, int32_t* fp64, int32_t* mixed)
{
    if (!c) return CBS_ERR_BAD_ARGUMENT;
    if (fp64) *fp64 = (c->use_graphs && c->pcg_graph[0].exec) ? 1 : 0;
    if (mixed) *mixed = (c->use_graphs && c->pcg_graph[1].exec) ? 1 : 0;
    return CBS_OK;
}

int cbs_hip_apply_config(cbs_hip_ctx* c, int32_t precision, int32_t* threads, int32_t* prefetch, float* candidate_ms)
{
    return guarded(c, [&] {
        const cbs_hip_sell::Config& tc = c->sell.cfg[precision == CBS_PCG_MIXED ? 1 : 0];
        if (threads) *threads = tc.threads;
        if (prefetch) *prefetch = tc.batch;
        if (candidate_ms)
            for (int i = 0; i < 6; ++i) candidate_ms[i] = tc.tune_ms[i];
        return int(CBS_OK);
    });
}

int cbs_hip_operator_info(const cbs_hip_ctx* c, cbs_operator_info* out)
{
    if (!c || !out) return CBS_ERR_BAD_ARGUMENT;
    const cbs_hip_sell& m = c->sell;
    out->row_tiles = m.ntiles;
    out->max_staged = m.lmax;
    out->staged_nodes = m.nstage;
    out->nnz = m.nnz;
    out->entries = m.entries;
    out->contributions = m.ncontrib;
    out->product_parts = c->sell_parts.nparts;
    out->product_max_staged = c->sell_parts.lmax;
    return CBS_OK;
}

int cbs_hip_apply_assembled(cbs_hip_ctx* c, int32_t precision, const double* x, double* y)
{
    return guarded(c, [&] {
        if (!x || !y) throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_apply_assembled - null vector");
        if (precision != CBS_PCG_FP64 && precision != CBS_PCG_MIXED)
            throw CbsFailure(CBS_ERR_BAD_ARGUMENT, "cbs_hip_apply_assembled - unknown precision");
        const size_t n = size_t(c->N);
        hip_check(hipMemcpyAsync(c->tmp_x, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream), "x");
        sell_apply_test(c, precision, c->tmp_x, c->cg_r);
        hip_check(hipMemcpyAsync(y, c->cg_r, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "y");
        hip_check(hipStreamSynchronize(c->stream), "apply sync");
        return int(CBS_OK);
    });
}

int cbs_hip_apply_timer(cbs_hip_ctx* c, int32_t precision, double* ms, int64_t* count, int32_t reset)
{
    return guarded(c, [&] {
        sync(c);
        timer_collect(c);
        const int p = precision == CBS_PCG_MIXED ? 1 : 0;
        if (ms) *ms = c->apply_ms[p];
        if (count) *count = c->apply_n[p];
        if (reset)
        {
            c->apply_ms[0] = c->apply_ms[1] = 0.0;
            c->apply_n[0] = c->apply_n[1] = 0;
        }
        return int(CBS_OK);
    });
}

int cbs_hip_advance_one_step(cbs_hip_ctx* c, int32_t iitime, const cbs_cg_options* opts, cbs_cg_result* cg)
{
    return guarded(c, [&] {
        const size_t N = size_t(c->N);
        copy_state(c, c->unkn1, c->unkno, 3 * N);
        copy_state(c, c->pres1, c->pres, N);
        copy_state(c, c->temp1, c->temp, N);
        validate_timestep_inputs(c);
        launch_timestep(c, iitime);
        launch_lhs(c);
        launch_step1_rhs(c);
        launch_step1_update(c);
        launch_bc(c, 0);
        if (c->cfg.step2_check > 0)
            throw CbsFailure(CBS_ERR_UNSUPPORTED,
                             "TimeStep::applyStep2PressureTimeStepCorrection is not supported on the device path");
        // Errors of the stages so far must surface before the solve reads them.
        sync(c);
        int rc = check_device_error(c);
        if (rc != CBS_OK) return rc;
        rc = step2_body(c, opts, cg);
        if (rc != CBS_OK) return rc;
        if (c->cfg.transient_on > 0)
        {
            std::vector<DevScalars> h(size_t(c->nseg));
            hip_check(hipMemcpy(h.data(), c->scal, sizeof(DevScalars) * h.size(), hipMemcpyDeviceToHost), "dtreal");
            for (int k = 0; k < c->nseg; ++k) c->seg_rtime[size_t(k)] += h[size_t(k)].dtreal;
            c->rtime = c->seg_rtime[0];
        }
        launch_step3(c);
        launch_bc(c, 0);
        if (c->cfg.temp_calc >= 1)
        {
            launch_step4_rhs(c);
            launch_step4_update(c);
            launch_bc(c, 2);
        }
        launch_velocity_magnitude(c);
        launch_convergence(c);
        return finish(c);
    });
}

}  // extern "C"
