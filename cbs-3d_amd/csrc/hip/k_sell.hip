// The assembled pressure operator of the PCG: sliced ELLPACK over row tiles,
// with each tile's x staged in LDS.
//
// The element-by-element apply (MatrixVectorCalc::pressureMultiply,
// MatrixVectorCalc.cpp:125-205) streams six coefficients and four node ids
// per tet.  Every edge of a tet mesh lies in about five tets, so that stream
// is about 2.4x the assembled operator's: ~14 off-diagonal entries per node,
// each an 8-byte value and a 16-bit column.  The PCG needs A x only to
// round-off (the solve is compared with tolerances; pressureMultiply itself
// keeps the bit-exact element form), so its apply runs on the assembled
// operator:
//   (A x)_i = free_i ? sum_j a_ij x_j + pdiag_i x_i : 0,
//   a_ij    = sum over the fluid tets e holding edge (i, j) of gstif_e(i, j),
// with the tets of each sum taken in tile-slot order.
//
// Layout:
//  - Row tiles.  Every element tile (build_tiles in ctx.hip) owns the nodes
//    it lists that no earlier tile lists; those nodes are its rows.  A row
//    tile stages x of its rows (first, in row order) and of their other
//    columns (the halo, ascending node id) in LDS, so a column is a 16-bit
//    local id.  A tile whose staged set would exceed the LDS budget is split
//    into row halves.  Tiles never straddle segments.
//  - Solver order.  The PCG's vectors are numbered so that every row tile's
//    rows are consecutive (row tiles in order, then the segment's nodes that
//    no tile owns, ascending), each segment keeping its node range; gid[s]
//    is the node of solver index s.  Staging the rows, reading pdiag and the
//    node flags and writing y are then contiguous accesses; only the halo
//    is gathered, through its list of solver indices.
//  - Slices.  The rows of a tile are sorted by entry count (descending, then
//    node id) and cut into slices of 64 rows, one wavefront each.  A slice
//    holds K entries per row (its longest row, rounded up to even) as K/2
//    column pairs: pair j of lane l is item j*64 + l of the slice, a 32-bit
//    word with two local ids and a two-vector of values, so every load of
//    the apply is coalesced.  Entry (lane l, column k) of a slice at entry
//    offset off is value off + (k/2)*128 + 2l + (k&1).
//  - Values.  Rebuilt at every solve from the tile-ordered coefficients
//    gstif = gstifE * delte through per-entry contribution lists (plane*F +
//    tile slot, in tile-slot order), or as dt times the gstifE sums where a
//    segment has one time step; the FP32 values are the rounded FP64 sums.
#include <algorithm>
#include <atomic>
#include <thread>

#include "ctx.hpp"
#include "device_common.hpp"

namespace cbsdev
{
    namespace
    {
        // plane of the local pair (a, b), a != b: (1,2)(1,3)(1,4)(2,3)(2,4)(3,4)
        constexpr int kPairPlane[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};

        inline bool host_free(uint8_t f) { return (f & 1u) && !(f & 4u); }

        struct RowTile
        {
            std::vector<int32_t> stage;     // rows, then halo (node ids)
            int32_t nrows = 0;
            std::vector<int32_t> pairs;     // column pairs of each slice
            std::vector<uint32_t> cols;     // packed local-id pairs, slice after slice
            std::vector<uint32_t> count;    // contributions of each entry
            std::vector<uint32_t> contrib;  // plane*F + tile slot, entry after entry
            int64_t nnz = 0;
        };

        struct Item
        {
            int32_t col;
            int32_t slot;
            int32_t plane;
        };

        // Rows `rows` (global ids) of one tile; appends one or more row tiles.
        bool build_row_tile(const SellSource& s, std::vector<int32_t> rows, int cap, std::vector<RowTile>& out,
                            std::string& err)
        {
            struct Row
            {
                int32_t gid, k, e0;
            };
            std::vector<Row> rd;
            rd.reserve(rows.size());
            std::vector<int32_t> ecol;
            std::vector<uint32_t> eptr{0}, contrib;
            std::vector<Item> tmp;
            for (int32_t i : rows)
            {
                Row r{i, 0, int32_t(ecol.size())};
                if (host_free(s.nflag[i]))
                {
                    tmp.clear();
                    for (int k = s.inc_ptr[i]; k < s.inc_ptr[i + 1]; ++k)
                    {
                        const int e = s.inc[k] >> 2, a = s.inc[k] & 3;
                        const int ts = s.elem_tslot[e];
                        if (ts < 0) continue;  // solid
                        const int4 n = s.conn[e];
                        const int nb[4] = {n.x, n.y, n.z, n.w};
                        for (int b = 0; b < 4; ++b)
                            if (b != a) tmp.push_back({nb[b], ts, kPairPlane[a][b]});
                    }
                    std::sort(tmp.begin(), tmp.end(), [](const Item& p, const Item& q) {
                        return p.col < q.col || (p.col == q.col && p.slot < q.slot);
                    });
                    for (size_t u = 0; u < tmp.size(); ++u)
                    {
                        if (u == 0 || tmp[u].col != tmp[u - 1].col)
                        {
                            if (u > 0) eptr.push_back(uint32_t(contrib.size()));
                            ecol.push_back(tmp[u].col);
                            ++r.k;
                        }
                        contrib.push_back(uint32_t(int64_t(tmp[u].plane) * s.F + tmp[u].slot));
                    }
                    if (!tmp.empty()) eptr.push_back(uint32_t(contrib.size()));
                }
                rd.push_back(r);
            }
            std::stable_sort(rd.begin(), rd.end(), [](const Row& p, const Row& q) {
                return p.k > q.k || (p.k == q.k && p.gid < q.gid);
            });
            // staged set: rows in row order, then the halo ascending
            std::vector<int32_t> halo(ecol);
            std::vector<int32_t> own;
            own.reserve(rd.size());
            for (const Row& r : rd) own.push_back(r.gid);
            std::vector<int32_t> own_sorted(own);
            std::sort(own_sorted.begin(), own_sorted.end());
            std::sort(halo.begin(), halo.end());
            halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
            halo.erase(std::remove_if(halo.begin(), halo.end(),
                                      [&](int32_t g) {
                                          return std::binary_search(own_sorted.begin(), own_sorted.end(), g);
                                      }),
                       halo.end());
            const size_t staged = own.size() + halo.size();
            if (staged > size_t(cap))
            {
                if (rows.size() < 2)
                {
                    err = "a node's neighbourhood exceeds the LDS budget of one workgroup";
                    return false;
                }
                const size_t h = rows.size() / 2;
                std::vector<int32_t> lo(rows.begin(), rows.begin() + long(h)), hi(rows.begin() + long(h), rows.end());
                return build_row_tile(s, std::move(lo), cap, out, err) && build_row_tile(s, std::move(hi), cap, out, err);
            }
            RowTile t;
            t.nrows = int32_t(own.size());
            t.stage = own;
            t.stage.insert(t.stage.end(), halo.begin(), halo.end());
            std::vector<std::pair<int32_t, int32_t>> local(staged);
            for (size_t l = 0; l < staged; ++l) local[l] = {t.stage[l], int32_t(l)};
            std::sort(local.begin(), local.end());
            auto lid = [&](int32_t g) {
                auto it = std::lower_bound(local.begin(), local.end(), std::make_pair(g, int32_t(-1)));
                return uint32_t(it->second);
            };
            const int nslice = (t.nrows + 63) / 64;
            for (int sl = 0; sl < nslice; ++sl)
            {
                const int kmax = rd[size_t(sl) * 64].k;
                const int np = (kmax + 1) / 2;
                t.pairs.push_back(np);
                for (int j = 0; j < np; ++j)
                    for (int l = 0; l < 64; ++l)
                    {
                        const int r = sl * 64 + l;
                        uint32_t word = 0;
                        for (int h = 0; h < 2; ++h)
                        {
                            const int kk = 2 * j + h;
                            uint32_t cnt = 0;
                            if (r < t.nrows && kk < rd[size_t(r)].k)
                            {
                                const int32_t e = rd[size_t(r)].e0 + kk;
                                word |= lid(ecol[size_t(e)]) << (16 * h);
                                cnt = eptr[size_t(e) + 1] - eptr[size_t(e)];
                                t.contrib.insert(t.contrib.end(), contrib.begin() + eptr[size_t(e)],
                                                 contrib.begin() + eptr[size_t(e) + 1]);
                                ++t.nnz;
                            }
                            t.count.push_back(cnt);
                        }
                        t.cols.push_back(word);
                    }
            }
            out.push_back(std::move(t));
            return true;
        }
    }

    bool build_sell_host(const SellSource& s, SellHost& out, std::string& err)
    {
        out = SellHost{};
        if (int64_t(s.F) * 12 >= (int64_t(1) << 32))
        {
            err = "more fluid elements than the assembled operator's 32-bit contribution lists address";
            return false;
        }
        // staged FP64 x (8 B a node) and one FP64 slot per 64-row slice
        const int cap = std::min(65535, std::max(64, int(int64_t(s.lds_bytes) * 64 / 520)));
        // owner: the first element tile listing a node
        std::vector<int32_t> owner(size_t(s.N), -1);
        for (int t = 0; t < s.ntiles; ++t)
            for (int k = s.tile_node_ptr[t]; k < s.tile_node_ptr[t + 1]; ++k)
                if (owner[size_t(s.tile_nodes[k])] < 0) owner[size_t(s.tile_nodes[k])] = t;
        std::vector<std::vector<RowTile>> built(size_t(s.ntiles));
        std::vector<std::string> errs(size_t(s.ntiles));
        std::atomic<int> next{0};
        std::atomic<bool> failed{false};
        auto work = [&] {
            for (int t = next++; t < s.ntiles && !failed; t = next++)
            {
                std::vector<int32_t> rows;
                for (int k = s.tile_node_ptr[t]; k < s.tile_node_ptr[t + 1]; ++k)
                    if (owner[size_t(s.tile_nodes[k])] == t) rows.push_back(s.tile_nodes[k]);
                if (rows.empty()) continue;
                if (!build_row_tile(s, std::move(rows), cap, built[size_t(t)], errs[size_t(t)])) failed = true;
            }
        };
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned k = 0; k < nth; ++k) pool.emplace_back(work);
        for (auto& th : pool) th.join();
        for (const auto& e : errs)
            if (!e.empty())
            {
                err = e;
                return false;
            }
        // concatenate in element-tile order (segments stay contiguous)
        std::vector<std::vector<int32_t>> halo_gid;
        int seg = 0;
        for (int t = 0; t < s.ntiles; ++t)
        {
            while (seg < s.nseg && t >= s.seg_tile[seg + 1])
            {
                out.seg_tile.push_back(int32_t(out.row_cnt.size()));
                ++seg;
            }
            for (RowTile& r : built[size_t(t)])
            {
                out.rows.insert(out.rows.end(), r.stage.begin(), r.stage.begin() + r.nrows);
                halo_gid.emplace_back(r.stage.begin() + r.nrows, r.stage.end());
                out.row_cnt.push_back(r.nrows);
                out.lmax = std::max(out.lmax, int32_t(r.stage.size()));
                out.max_slices = std::max(out.max_slices, int32_t(r.pairs.size()));
                for (int32_t np : r.pairs) out.slice_off.push_back(out.slice_off.back() + int64_t(np) * 128);
                out.slice_ptr.push_back(int32_t(out.slice_off.size()) - 1);
                out.cols.insert(out.cols.end(), r.cols.begin(), r.cols.end());
                for (uint32_t cnt : r.count) out.cptr.push_back(out.cptr.back() + cnt);
                out.cidx.insert(out.cidx.end(), r.contrib.begin(), r.contrib.end());
                out.nnz += r.nnz;
                r = RowTile{};
            }
        }
        while (int(out.seg_tile.size()) < s.nseg + 1) out.seg_tile.push_back(int32_t(out.row_cnt.size()));
        // solver order: per segment, the rows of its tiles in tile order,
        // then its nodes no tile owns, ascending
        std::vector<int32_t> sol(size_t(s.N), -1);
        out.gid.assign(size_t(s.N), -1);
        out.row_base.assign(out.row_cnt.size(), 0);
        size_t rpos = 0;
        for (int k = 0; k < s.nseg; ++k)
        {
            int32_t next = s.seg_node[k];
            for (int t = out.seg_tile[k]; t < out.seg_tile[k + 1]; ++t)
            {
                out.row_base[size_t(t)] = next;
                for (int j = 0; j < out.row_cnt[size_t(t)]; ++j)
                {
                    const int32_t g = out.rows[rpos++];
                    if (g < s.seg_node[k] || g >= s.seg_node[k + 1] || sol[size_t(g)] >= 0)
                    {
                        err = "inconsistent row ownership";
                        return false;
                    }
                    sol[size_t(g)] = next;
                    out.gid[size_t(next++)] = g;
                }
            }
            for (int32_t g = s.seg_node[k]; g < s.seg_node[k + 1]; ++g)
                if (sol[size_t(g)] < 0)
                {
                    sol[size_t(g)] = next;
                    out.gid[size_t(next++)] = g;
                }
        }
        for (const auto& h : halo_gid)
        {
            for (int32_t g : h) out.halo.push_back(sol[size_t(g)]);
            out.halo_ptr.push_back(int32_t(out.halo.size()));
        }
        return true;
    }

    // ------------------------------------------------------------------
    // Device
    // ------------------------------------------------------------------
    template <typename T>
    struct Pair2;
    template <>
    struct Pair2<double>
    {
        using type = double2;
        using vec = double __attribute__((ext_vector_type(2)));
    };
    template <>
    struct Pair2<float>
    {
        using type = float2;
        using vec = float __attribute__((ext_vector_type(2)));
    };

    // The operator's values and column pairs are read once per apply:
    // non-temporal loads keep them from displacing the staged x (shared by
    // neighbouring tiles) in L2.
    template <typename T>
    __device__ __forceinline__ typename Pair2<T>::type load_stream(const typename Pair2<T>::type* p)
    {
        using V = typename Pair2<T>::vec;
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
        return typename Pair2<T>::type{v.x, v.y};
    }

    constexpr int kSpmmThreads = 1024;  // workgroup of the multi-column product
    constexpr int kStageBatch = 8;       // staged nodes per thread in flight
    constexpr int kApplySgprs = 56;      // SGPR budget of k_apply_sell (see there)

    struct SellView
    {
        const int32_t* row_base;
        const int32_t* row_cnt;
        const int32_t* halo_ptr;
        const int32_t* halo;
        const int32_t* slice_ptr;
        const int64_t* slice_off;
        const uint32_t* cols;
    };

    // The PCG's direction update (ConjugateGradient.cpp:831-849, p = free ?
    // z + beta p : 0) fused into the apply's staging (kDir): iteration j >= 1
    // stages p_j = free ? z + beta p_{j-1} : 0 for every staged node and
    // writes it for the tile's rows; iteration 0 stages p_0 as the solve's
    // set-up left it.  p_j lives in p0 for even j and in p1 for odd j.  FP64:
    // z = r / pdiag, the reference's Jacobi step (z = r without
    // preconditioner), and the sum in FP64.  Mixed: z = r * dinv in FP64
    // rounded to FP32, then z + float(beta) p_{j-1} in FP32.  beta and j are
    // the segment's DevScalars fields; a tile's rows are written by that tile
    // only, and every tile reads p_{j-1} from the other buffer.
    template <typename T>
    struct DirArgs
    {
        T* p0 = nullptr;
        T* p1 = nullptr;
        const T* r = nullptr;
        const T* zs = nullptr;  // pdiag (FP64) or dinv (FP32), solver order
        const DevScalars* scal = nullptr;
        int precond = 1;
        int force = 0;  // form the direction at iteration 0 too (timing runs on scratch vectors)
    };

    template <typename T>
    __device__ __forceinline__ T direction_of(const DirArgs<T>& d, T r, T zs, T pold, T beta)
    {
        if constexpr (sizeof(T) == 8)
        {
            const double z = (d.precond == 1) ? r / zs : r;
            return z + beta * pold;
        }
        else
        {
            return float(double(r) * double(zs)) + beta * pold;
        }
    }

    // One workgroup per row tile: stage x, then each wavefront takes the
    // tile's slices one at a time (first s0 + wave, then the next one not yet
    // taken); B column pairs are loaded per batch before they are used.  kAp:
    // y = free ? acc + pdiag x : 0 and the tile's partial of (x, y);
    // otherwise y = acc (off-diagonal part).  Vectors, pdiag and flags are in
    // solver order.  The partial depends neither on the launch shape nor on
    // which wavefront takes which slice: each slice's lanes are summed by the
    // wavefront butterfly into the slice's LDS slot, and the first wavefront
    // adds the slots (lane l: slots l, l+64, ..., then the butterfly).  LDS:
    // max_slices doubles, then the staged x.
    // FP32 (kDiff): a pressure row nearly sums to zero (pdiag cancels the
    // off-diagonal sum for interior nodes), so acc + pdiag x loses all of
    // FP32's digits on smooth x, and on uniform meshes the mixed PCG needed
    // up to 5x the FP64 iterations.  The FP32 apply therefore computes
    //   y_i = sum_j a_ij (x_j - x_i) + rs_i x_i,  rs_i = pdiag_i + sum_j a_ij
    // (rs_i formed in FP64 and rounded once: sell_rowsum32), a graph
    // Laplacian with FP32 edge weights plus the row-sum term, without the
    // cancellation; equal to acc + pdiag x in exact arithmetic.
    // SGPRs are capped (kApplySgprs): with the staging loop's 79-83 SGPRs per
    // wavefront, an MI355X CU held one 1,024-thread workgroup of this kernel
    // instead of two (per-workgroup clocks at cavity160), although the
    // compiler's occupancy estimate was 8 wavefronts per SIMD either way.
    // Spilled SGPRs go to VGPR lanes.
    template <typename T, bool kAp, int B, bool kDir>
    __global__ __launch_bounds__(1024) __attribute__((amdgpu_num_sgpr(kApplySgprs))) void k_apply_sell(
        SellView v, const T* __restrict__ vals, const T* __restrict__ x_even, const T* __restrict__ x_odd,
        const int* __restrict__ iter0, T* __restrict__ y, const uint8_t* __restrict__ nflag,
        const T* __restrict__ pdiag, double* __restrict__ dotpart, const int* __restrict__ flag0, int flag_stride,
        const int32_t* __restrict__ tile_seg, int max_slices, DirArgs<T> d)
    {
        using T2 = typename Pair2<T>::type;
        constexpr bool kDiff = sizeof(T) == 4 && kAp;
        const int t = xcd_block(int(gridDim.x));  // neighbouring tiles share halo nodes
        const int sg = tile_seg ? tile_seg[t] : 0;
        if (flag0[sg * flag_stride]) return;  // segment finished
        // the direction of odd iterations lives in its own buffer
        const int it = kDir ? d.scal[sg].iter : 0;
        const T* __restrict__ x = kDir ? ((it & 1) ? d.p1 : d.p0)
                                       : ((x_odd && (iter0[sg * flag_stride] & 1)) ? x_odd : x_even);
        // kDir from iteration 1 on: this iteration's direction is formed here
        const bool form = kDir && (it > 0 || d.force);
        T* __restrict__ pnew = (it & 1) ? d.p1 : d.p0;
        const T* __restrict__ pold = (it & 1) ? d.p0 : d.p1;
        const T beta = kDir ? T(d.scal[sg].beta) : T(0);
        // x, or the direction formed from the previous one, of solver index gi
        auto stage_value = [&](int gi) -> T {
            if (!form) return x[gi];
            const uint8_t f = nflag[gi];
            const T pn = direction_of<T>(d, d.r[gi], d.zs[gi], pold[gi], beta);
            return is_free(f) ? pn : T(0);
        };
        extern __shared__ __align__(16) unsigned char smem[];
        double* sdot = reinterpret_cast<double*>(smem);
        T* lx = reinterpret_cast<T*>(smem + size_t(max_slices) * sizeof(double));
        __shared__ int next_slice;
        const int tid = threadIdx.x, nt = blockDim.x;
        const int rb = v.row_base[t], nrows = v.row_cnt[t];
        const int hb = v.halo_ptr[t], S = nrows + (v.halo_ptr[t + 1] - hb);
        const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
        const int s0 = v.slice_ptr[t], s1 = v.slice_ptr[t + 1];
        if (tid == 0) next_slice = s0 + nw;
        // Staging: kStageBatch nodes per thread in flight (the halo's ids,
        // then all the values), so a thread waits for two memory round trips
        // instead of two per node.  Measured with per-workgroup clocks at
        // cavity160, the node-at-a-time loop took 4-17 us of a ~30 us tile
        // while the other tiles on the CU streamed.  A tile with at most one
        // node per thread keeps the plain form (the batch measured ~1 us
        // slower per apply on the 128x32x32 channel's small tiles).
        // (kDir: four loads per node, so half the batch; six per thread
        // measured no faster)
        constexpr int KB = kDir ? kStageBatch / 2 : kStageBatch;
        if (S <= nt)
        {
            if (tid < S)
            {
                const T xv = stage_value((tid < nrows) ? rb + tid : v.halo[hb + tid - nrows]);
                lx[tid] = xv;
                if (form && tid < nrows) pnew[rb + tid] = xv;
            }
        }
        else
        {
            for (int l0 = tid; l0 < S; l0 += nt * KB)
            {
                int idx[KB];
                T xv[KB];
#pragma unroll
                for (int u = 0; u < KB; ++u)
                {
                    const int l = l0 + u * nt;
                    idx[u] = (l < nrows) ? rb + l : (l < S ? v.halo[hb + l - nrows] : -1);
                }
#pragma unroll
                for (int u = 0; u < KB; ++u) xv[u] = (idx[u] >= 0) ? stage_value(idx[u]) : T(0);
#pragma unroll
                for (int u = 0; u < KB; ++u)
                    if (idx[u] >= 0)
                    {
                        lx[l0 + u * nt] = xv[u];
                        if (form && l0 + u * nt < nrows) pnew[idx[u]] = xv[u];
                    }
            }
        }
        __syncthreads();
        // Slices go to wavefronts as they finish (wavefront w starts with
        // slice s0 + w): memory latency varies from slice to slice, and with
        // a fixed round-robin the last wavefront of a tile ended ~6 us after
        // the first.  Each slice's sum and (x, y) partial do not depend on
        // the wavefront that computes it.
        for (int s = s0 + wave; s < s1;)
        {
            const int64_t off = v.slice_off[s];
            const int np = int((v.slice_off[s + 1] - off) >> 7);
            const uint32_t* cp = v.cols + (off >> 1) + lane;
            const T2* vp = reinterpret_cast<const T2*>(vals + off) + lane;
            const int row = (s - s0) * 64 + lane;
            // FP32: the difference form sum_j a_ij (x_j - x_i), pdiag holding
            // the row sums (sell_rowsum32); see kDiff above
            const T xi = (kDiff && row < nrows) ? lx[row] : T(0);
            // the row's flag and diagonal load with the stream, not after it
            uint8_t fl = 0;
            T pd = T(0);
            if (kAp && row < nrows)
            {
                fl = nflag[rb + row];
                pd = pdiag[rb + row];
            }
            T acc = T(0);
            for (int j0 = 0; j0 < np; j0 += B)
            {
                uint32_t cc[B];
                T2 vv[B];
#pragma unroll
                for (int u = 0; u < B; ++u)
                {
                    cc[u] = 0u;
                    vv[u] = T2{T(0), T(0)};
                    if (j0 + u < np)
                    {
                        cc[u] = __builtin_nontemporal_load(cp + size_t(j0 + u) * 64);
                        vv[u] = load_stream<T>(vp + size_t(j0 + u) * 64);
                    }
                }
#pragma unroll
                for (int u = 0; u < B; ++u)
                    if (j0 + u < np)
                    {
                        if (kDiff)
                        {
                            acc += vv[u].x * (lx[cc[u] & 0xffffu] - xi);
                            acc += vv[u].y * (lx[cc[u] >> 16] - xi);
                        }
                        else
                        {
                            acc += vv[u].x * lx[cc[u] & 0xffffu];
                            acc += vv[u].y * lx[cc[u] >> 16];
                        }
                    }
            }
            double dot = 0.0;
            if (row < nrows)
            {
                const int gi = rb + row;
                if (kAp)
                {
                    T a = T(0);
                    if (is_free(fl))
                    {
                        const T p = lx[row];
                        a = acc + pd * p;
                        dot = double(p) * double(a);
                    }
                    y[gi] = a;
                }
                else
                {
                    y[gi] = acc;
                }
            }
            if (kAp)
            {
                dot = wave_sum(dot);  // all 64 lanes of the wavefront take part
                if (lane == 0) sdot[s - s0] = dot;
            }
            int ns = 0;
            if (lane == 0) ns = atomicAdd(&next_slice, 1);
            s = __builtin_amdgcn_readfirstlane(ns);
        }
        if (kAp)
        {
            // the slots by the first wavefront: lane l adds slots l, l+64,
            // ..., then the butterfly (a fixed order for any launch shape)
            __syncthreads();
            if (tid < 64)
            {
                double sum = 0.0;
                for (int j = tid; j < s1 - s0; j += 64) sum += sdot[j];
                sum = wave_sum(sum);
                if (tid == 0) dotpart[t] = sum;
            }
        }
    }

    namespace
    {
        size_t sell_lds_bytes(const cbs_hip_sell& m, size_t elem)
        {
            return size_t(m.max_slices) * sizeof(double) + size_t(m.lmax) * elem;
        }

        SellView view_of(const cbs_hip_sell& m)
        {
            return {m.row_base, m.row_cnt, m.halo_ptr, m.halo, m.slice_ptr, m.slice_off, m.cols};
        }

        constexpr int kScalarStride = int(sizeof(DevScalars) / sizeof(int));

        template <typename T, bool kAp, int B, bool kDir>
        void launch_sell_b(cbs_hip_ctx* c, int threads, const T* vals, const T* x, T* y, const T* pdiag,
                           const int* flag0, const T* x_odd, const int* iter0, const DirArgs<T>& d)
        {
            const cbs_hip_sell& m = c->sell;
            const size_t lds = sell_lds_bytes(m, sizeof(T));
            if (c->timer_pending)
            {
                // the apply timer's events mark this kernel's own start and end
                c->timer_pending = false;
                hipExtLaunchKernelGGL((k_apply_sell<T, kAp, B, kDir>), dim3(m.ntiles), dim3(threads), lds, c->stream,
                                      c->events[size_t(c->events_used)], c->events[size_t(c->events_used) + 1], 0,
                                      view_of(m), vals, x, x_odd, iter0, y, c->sv_flag, pdiag, c->dotpart, flag0,
                                      kScalarStride, m.tile_seg, m.max_slices, d);
                return;
            }
            hipLaunchKernelGGL((k_apply_sell<T, kAp, B, kDir>), dim3(m.ntiles), dim3(threads), lds, c->stream,
                               view_of(m), vals, x, x_odd, iter0, y, c->sv_flag, pdiag, c->dotpart, flag0,
                               kScalarStride, m.tile_seg, m.max_slices, d);
        }

        template <typename T, bool kAp, bool kDir = false>
        void launch_sell_t(cbs_hip_ctx* c, int threads, int batch, const T* vals, const T* x, T* y, const T* pdiag,
                           const int* flag0, const T* x_odd = nullptr, const int* iter0 = nullptr,
                           const DirArgs<T>& d = DirArgs<T>{})
        {
            if (c->sell.ntiles == 0) return;
            if (batch == 4) launch_sell_b<T, kAp, 4, kDir>(c, threads, vals, x, y, pdiag, flag0, x_odd, iter0, d);
            else launch_sell_b<T, kAp, 8, kDir>(c, threads, vals, x, y, pdiag, flag0, x_odd, iter0, d);
        }

        template <typename T, bool kAp, int B>
        void allow_lds(size_t bytes)
        {
            for (const void* f : {reinterpret_cast<const void*>(&k_apply_sell<T, kAp, B, false>),
                                  reinterpret_cast<const void*>(&k_apply_sell<T, kAp, B, true>)})
                (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
        }

        // Times every (threads, batch) candidate of one precision's apply on
        // the context's operator and keeps the fastest.  The result of the
        // apply does not depend on the choice.
        template <typename T>
        void autotune_t(cbs_hip_ctx* c, const T* vals, const T* x, T* y, const T* pdiag)
        {
            cbs_hip_sell::Config& cf = c->sell.cfg[sizeof(T) == 4 ? 1 : 0];
            if (cf.tuned || c->sell.ntiles == 0) return;
            const int cand[6][2] = {{1024, 8}, {1024, 4}, {512, 8}, {512, 4}, {256, 8}, {256, 4}};
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            const int* never = &c->scal->pad0;
            float best = 1e30f;
            int pick = 0;
            for (int k = 0; k < 6; ++k)
            {
                auto run = [&] { launch_sell_t<T, true>(c, cand[k][0], cand[k][1], vals, x, y, pdiag, never); };
                run();
                hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
                for (int r = 0; r < 3; ++r) run();
                hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                cf.tune_ms[k] = ms / 3.0f;
                if (ms < best)
                {
                    best = ms;
                    pick = k;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            cf.threads = cand[pick][0];
            cf.batch = cand[pick][1];
            cf.tuned = true;
        }
    }

    void sell_configure(cbs_hip_ctx* c)
    {
        const size_t b64 = sell_lds_bytes(c->sell, sizeof(double)), b32 = sell_lds_bytes(c->sell, sizeof(float));
        // Allows dynamic LDS beyond the default 64 KB (best effort, as for
        // the element tiles).
        allow_lds<double, true, 4>(b64);
        allow_lds<double, true, 8>(b64);
        allow_lds<double, false, 4>(b64);
        allow_lds<double, false, 8>(b64);
        allow_lds<float, true, 4>(b32);
        allow_lds<float, true, 8>(b32);
        (void)hipGetLastError();
    }

    // Segment flags for the value refresh: 1 when every element of the
    // segment has the time step of its first element (global or fixed time
    // stepping), 0 otherwise (local time stepping, or a delte set by hand).
    __global__ void k_dt_uniform_reset(int nseg, int32_t* __restrict__ flag)
    {
        const int k = int(blockIdx.x) * kBlock + int(threadIdx.x);
        if (k < nseg) flag[k] = 1;
    }

    __global__ void k_dt_uniform(int E, const double* __restrict__ delte, const uint16_t* __restrict__ seg_of_elem,
                                 const int32_t* __restrict__ seg_elem0, int32_t* __restrict__ flag)
    {
        const int e = int(blockIdx.x) * kBlock + int(threadIdx.x);
        if (e >= E) return;
        const int sg = seg_of_elem ? int(seg_of_elem[e]) : 0;
        if (!(delte[e] == delte[seg_elem0[sg]])) flag[sg] = 0;
    }

    // a_ij of every stored entry.  Its contributions, in list order, are
    // gstif = gstifE * delte (TimeStep.cpp:873-901), so a
    // segment with one time step dt has values dt * (sum of its gstifE
    // contributions): one product per entry instead of a gather per
    // contribution.  The operator differs from the contribution sums by
    // round-off only, as the assembled operator already differs from the
    // reference's element-by-element apply.  Other segments sum the
    // contributions gstifE * delte, each formed as k_gstif forms it (tile
    // slot k of plane p is contribution index p*F + k).  g non-null (gstif
    // set by hand): every segment sums g.
    __global__ void k_sell_values_dt(int64_t n, int nseg, const int64_t* __restrict__ seg_entry,
                                     const int32_t* __restrict__ flag, const int32_t* __restrict__ seg_elem0,
                                     const double* __restrict__ delte, const double* __restrict__ base,
                                     const uint32_t* __restrict__ cptr, const uint32_t* __restrict__ cidx,
                                     const double* __restrict__ g, const double* __restrict__ gE,
                                     const int32_t* __restrict__ tslot_elem, uint32_t F, double* __restrict__ v64,
                                     float* __restrict__ v32)
    {
        const int64_t q = int64_t(xcd_block(int(gridDim.x))) * kBlock + threadIdx.x;
        if (q >= n) return;
        int lo = 0, hi = nseg;  // the segment whose entries hold q
        while (hi - lo > 1)
        {
            const int mid = (lo + hi) >> 1;
            if (q >= seg_entry[mid]) lo = mid;
            else hi = mid;
        }
        double s = 0.0;
        if (g)
        {
            for (uint32_t k = cptr[q]; k < cptr[q + 1]; ++k) s += g[cidx[k]];
        }
        else if (flag[lo])
        {
            s = delte[seg_elem0[lo]] * __builtin_nontemporal_load(base + q);
        }
        else
        {
            for (uint32_t k = cptr[q]; k < cptr[q + 1]; ++k)
            {
                const uint32_t ci = cidx[k];
                s += gE[ci] * delte[tslot_elem[ci % F]];
            }
        }
        __builtin_nontemporal_store(s, v64 + q);
        __builtin_nontemporal_store(float(s), v32 + q);
    }

    __global__ void k_sell_base(int64_t n, const uint32_t* __restrict__ cptr, const uint32_t* __restrict__ cidx,
                                const double* __restrict__ gE, double* __restrict__ base)
    {
        const int64_t q = int64_t(xcd_block(int(gridDim.x))) * kBlock + threadIdx.x;
        if (q >= n) return;
        double s = 0.0;
        for (uint32_t k = cptr[q]; k < cptr[q + 1]; ++k) s += gE[cidx[k]];
        base[q] = s;
    }

    void sell_refresh_values(cbs_hip_ctx* c)
    {
        cbs_hip_sell& m = c->sell;
        if (m.entries == 0) return;
        if (m.base_stale)
        {
            hipLaunchKernelGGL(k_sell_base, dim3(grid_for(m.entries)), dim3(kBlock), 0, c->stream, m.entries, m.cptr,
                               m.cidx, c->gstifE_t, m.base64);
            m.base_stale = false;
        }
        hipLaunchKernelGGL(k_dt_uniform_reset, dim3(grid_for(c->nseg)), dim3(kBlock), 0, c->stream, c->nseg,
                           m.dt_uniform);
        hipLaunchKernelGGL(k_dt_uniform, dim3(grid_for(c->E)), dim3(kBlock), 0, c->stream, c->E, c->delte,
                           c->seg_of_elem, m.seg_elem0, m.dt_uniform);
        hipLaunchKernelGGL(k_sell_values_dt, dim3(grid_for(m.entries)), dim3(kBlock), 0, c->stream, m.entries,
                           c->nseg, m.seg_entry, m.dt_uniform, m.seg_elem0, c->delte, m.base64, m.cptr, m.cidx,
                           c->gstif_manual ? c->gstif_t : static_cast<const double*>(nullptr), c->gstifE_t,
                           c->tslot_elem, uint32_t(c->F), m.v64, m.v32);
    }

    void sell_apply64(cbs_hip_ctx* c, const double* x, double* y, bool ap, const int* flag0, const double* x_odd,
                      const int* iter0)
    {
        const cbs_hip_sell::Config& cf = c->sell.cfg[0];
        if (ap)
            launch_sell_t<double, true>(c, cf.threads, cf.batch, c->sell.v64, x, y, c->sv_pdiag, flag0, x_odd, iter0);
        else
            launch_sell_t<double, false>(c, cf.threads, cf.batch, c->sell.v64, x, y, c->sv_pdiag, flag0, x_odd,
                                         iter0);
    }

    void sell_apply64_dir(cbs_hip_ctx* c, double* p0, double* p1, const double* r, int precond, double* y,
                          const int* flag0, int force)
    {
        const cbs_hip_sell::Config& cf = c->sell.cfg[0];
        DirArgs<double> d;
        d.p0 = p0;
        d.p1 = p1;
        d.r = r;
        d.zs = c->sv_pdiag;
        d.scal = c->scal;
        d.precond = precond;
        d.force = force;
        launch_sell_t<double, true, true>(c, cf.threads, cf.batch, c->sell.v64, p0, y, c->sv_pdiag, flag0, p1,
                                          nullptr, d);
    }

    __global__ void k_rowsum_narrow(const double* __restrict__ acc, const double* __restrict__ pdiag,
                                    float* __restrict__ rs, int n)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n) rs[i] = float(pdiag[i] + acc[i]);
    }

    // rs = pdiag + (sum of the row's off-diagonal FP64 values), in solver
    // order, rounded once to FP32: the diagonal term of the FP32 applies'
    // difference form.  ones and acc are N-double scratch vectors.
    void sell_rowsum32(cbs_hip_ctx* c, double* ones, double* acc, float* rs)
    {
        const int N = c->N;
        launch_fill(c, ones, N, 1.0);
        sell_apply64(c, ones, acc, false, &c->scal->pad0);
        hipLaunchKernelGGL(k_rowsum_narrow, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, acc, c->sv_pdiag, rs, N);
    }

    void sell_apply32(cbs_hip_ctx* c, const float* x, float* y, const float* pdiagf, const int* flag0,
                      const float* x_odd, const int* iter0)
    {
        const cbs_hip_sell::Config& cf = c->sell.cfg[1];
        launch_sell_t<float, true>(c, cf.threads, cf.batch, c->sell.v32, x, y, pdiagf, flag0, x_odd, iter0);
    }

    void sell_apply32_dir(cbs_hip_ctx* c, float* p0, float* p1, const float* r, const float* dinv,
                          const float* pdiagf, float* y, const int* flag0, int force)
    {
        const cbs_hip_sell::Config& cf = c->sell.cfg[1];
        DirArgs<float> d;
        d.p0 = p0;
        d.p1 = p1;
        d.r = r;
        d.zs = dinv;
        d.scal = c->scal;
        d.force = force;
        launch_sell_t<float, true, true>(c, cf.threads, cf.batch, c->sell.v32, p0, y, pdiagf, flag0, p1, nullptr,
                                         d);
    }

    __global__ void k_sell_narrow(const double* __restrict__ x, const double* __restrict__ d, float* __restrict__ xf,
                                  float* __restrict__ df, int n)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        xf[i] = float(x[i]);
        df[i] = float(d[i]);
    }

    __global__ void k_fillf(float* __restrict__ y, int n)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n) y[i] = 0.0f;
    }

    __global__ void k_sell_widen(const float* __restrict__ yf, double* __restrict__ y, int n)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n) y[i] = double(yf[i]);
    }

    __global__ void k_sol_gather(int n, const int32_t* __restrict__ gid, const double* __restrict__ src,
                                 double* __restrict__ dst)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n) dst[i] = src[gid[i]];
    }

    __global__ void k_sol_scatter(int n, const int32_t* __restrict__ gid, const double* __restrict__ src,
                                  double* __restrict__ dst)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n) dst[gid[i]] = src[i];
    }

    // The same permutations for multi-column vectors (mcol_at layout); the
    // padding nodes of the solver-order vector are zero.
    __global__ void k_sol_gather_cols(int n, int rshift, const int32_t* __restrict__ gid,
                                      const double* __restrict__ src, double* __restrict__ dst)
    {
        const size_t q = size_t(blockIdx.x) * kBlock + threadIdx.x;
        if (q >= mcol_size(n, rshift)) return;
        const int i = mcol_node(q, rshift);
        dst[q] = (i < n) ? src[mcol_at(gid[i], mcol_col(q, rshift), rshift)] : 0.0;
    }

    __global__ void k_sol_scatter_cols(int n, int rshift, const int32_t* __restrict__ gid,
                                       const double* __restrict__ src, double* __restrict__ dst)
    {
        const size_t q = size_t(blockIdx.x) * kBlock + threadIdx.x;
        if (q >= mcol_size(n, rshift)) return;
        const int i = mcol_node(q, rshift);
        if (i < n) dst[mcol_at(gid[i], mcol_col(q, rshift), rshift)] = src[q];
    }

    void sol_gather(cbs_hip_ctx* c, const double* node_order, double* solver_order)
    {
        hipLaunchKernelGGL(k_sol_gather, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->sell.gid,
                           node_order, solver_order);
    }

    void sol_scatter(cbs_hip_ctx* c, const double* solver_order, double* node_order)
    {
        hipLaunchKernelGGL(k_sol_scatter, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->sell.gid,
                           solver_order, node_order);
    }

    void sell_apply_test(cbs_hip_ctx* c, int precision, const double* x, double* y)
    {
        const int N = c->N;
        const int* never = &c->scal->pad0;
        sell_refresh_values(c);
        sol_gather(c, x, c->sv_x);
        sol_gather(c, c->pdiag, c->sv_pdiag);
        launch_fill(c, c->sv_b, N, 0.0);
        if (precision == CBS_PCG_MIXED)
        {
            hipLaunchKernelGGL(k_sell_narrow, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->sv_x, c->sv_pdiag,
                               c->f_p, c->f_z, N);
            sell_rowsum32(c, c->cg_p, c->cg_ap, c->f_z);
            hipLaunchKernelGGL(k_fillf, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->f_ap, N);
            sell_apply32(c, c->f_p, c->f_ap, c->f_z, never);
            hipLaunchKernelGGL(k_sell_widen, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->f_ap, c->sv_b, N);
        }
        else
        {
            sell_apply64(c, c->sv_x, c->sv_b, true, never);
        }
        sol_scatter(c, c->sv_b, y);
    }

    void sell_set_config(cbs_hip_ctx* c, int form, int threads, int batch)
    {
        cbs_hip_sell::Config& cf = c->sell.cfg[form];
        cf.threads = threads;
        cf.batch = batch;
        cf.tuned = true;
        for (float& v : cf.tune_ms) v = 0.0f;
    }

    void sell_autotune64(cbs_hip_ctx* c, const double* x, double* y)
    {
        autotune_t<double>(c, c->sell.v64, x, y, c->sv_pdiag);
    }

    void sell_autotune32(cbs_hip_ctx* c, const float* x, float* y, const float* pdiagf)
    {
        autotune_t<float>(c, c->sell.v32, x, y, pdiagf);
    }

    // ------------------------------------------------------------------
    // Multi-column product Y = A X on row tiles (the parts of the operator's
    // row tiles, build_parts below).  X and Y hold R = 2^rshift columns in
    // blocks of 8 nodes (mcol_at) in solver order.  One workgroup per (tile,
    // group of G columns): it stages the group's G values of every staged
    // node (rows, then halo) in LDS, and each lane of a slice keeps G sums,
    // so a column pair is loaded once for G columns and each of the tile's
    // nodes is gathered once per group instead of once per incident entry.
    // The tile's rows are consecutive in solver order, so their staging
    // reads and the writes of Y cover whole blocks of 8 nodes (G=2 in FP64,
    // G=4 in FP32: one 128-byte line per block).  A workgroup stages 16
    // bytes per node, so two of them fit a CU's LDS (waves_per_eu 8: the
    // register budget of two 1024-thread workgroups).  The groups
    // of a tile are consecutive logical blocks on one XCD: after the first,
    // they read the tile's column pairs and values from that XCD's L2
    // (ordinary loads, not streaming ones).  Column q of the sums takes the
    // tile's entries in the order of the single-column apply.  kAp: Y = free
    // ? acc + pdiag x : 0 and the tile's partials of (x_q, y_q) in
    // part[(tile << rshift) + col], the slices' wavefront sums added as in
    // the apply; otherwise Y = acc (the off-diagonal part).  Rows no tile
    // owns are not written.
    // ------------------------------------------------------------------
    template <typename T, int G>
    struct ColGroup
    {
        using type = T __attribute__((ext_vector_type(G)));
    };

    template <typename T, int G, bool kAp>
    __global__ __launch_bounds__(1024) __attribute__((amdgpu_num_sgpr(kApplySgprs), amdgpu_waves_per_eu(8, 8)))
    void k_spmm_sell(SellView v, const T* __restrict__ vals,
                                                       const T* __restrict__ x, T* __restrict__ y,
                                                       const uint8_t* __restrict__ nflag, const T* __restrict__ pdiag,
                                                       double* __restrict__ part, int rshift, int ngroups,
                                                       const int* __restrict__ stop, const int* __restrict__ skip,
                                                       int max_slices)
    {
        using V = typename ColGroup<T, G>::type;
        using T2 = typename Pair2<T>::type;
        constexpr bool kDiff = sizeof(T) == 4 && kAp;  // as k_apply_sell
        constexpr int B = 4;  // column pairs of a slice in flight
        if (*stop) return;
        const int L = xcd_block(int(gridDim.x));
        const int t = L / ngroups, g = L - t * ngroups;
        if (skip)
        {
            // a group whose columns all have skip set is not computed (its Y
            // and partials keep their previous values)
            bool all = true;
#pragma unroll
            for (int q = 0; q < G; ++q) all = all && skip[g * G + q] != 0;
            if (all) return;  // uniform over the workgroup
        }
        extern __shared__ __align__(16) unsigned char smem[];
        double* sdot = reinterpret_cast<double*>(smem);
        V* lx = reinterpret_cast<V*>(smem + size_t(max_slices) * G * sizeof(double));
        const int tid = threadIdx.x, nt = blockDim.x;
        const int rb = v.row_base[t], nrows = v.row_cnt[t];
        const int hb = v.halo_ptr[t], S = nrows + (v.halo_ptr[t + 1] - hb);
        const int c0 = g * G;  // first column of the group
        // kStageBatch staged nodes per thread in flight: their indices, then
        // their loads, then the LDS stores
        for (int l0 = tid; l0 < S; l0 += nt * kStageBatch)
        {
            int idx[kStageBatch];
            V xv[kStageBatch];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u)
            {
                const int l = l0 + u * nt;
                idx[u] = (l < nrows) ? rb + l : (l < S ? v.halo[hb + l - nrows] : -1);
            }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u)
                if (idx[u] >= 0)
                {
#pragma unroll
                    for (int q = 0; q < G; ++q) xv[u][q] = x[mcol_at(idx[u], c0 + q, rshift)];
                }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u)
                if (idx[u] >= 0) lx[l0 + u * nt] = xv[u];
        }
        __syncthreads();
        const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
        const int s0 = v.slice_ptr[t], s1 = v.slice_ptr[t + 1];
        for (int s = s0 + wave; s < s1; s += nw)
        {
            const int64_t off = v.slice_off[s];
            const int np = int((v.slice_off[s + 1] - off) >> 7);
            const uint32_t* cp = v.cols + (off >> 1) + lane;
            const T2* vp = reinterpret_cast<const T2*>(vals + off) + lane;
            const int row = (s - s0) * 64 + lane;
            // FP32: difference form, as k_apply_sell
            const V xi = (kDiff && row < nrows) ? lx[row] : V(T(0));
            V acc = V(T(0));
            for (int j0 = 0; j0 < np; j0 += B)
            {
                uint32_t cc[B];
                T2 vv[B];
#pragma unroll
                for (int u = 0; u < B; ++u)
                {
                    cc[u] = 0u;
                    vv[u] = T2{T(0), T(0)};
                    if (j0 + u < np)
                    {
                        cc[u] = cp[size_t(j0 + u) * 64];
                        vv[u] = vp[size_t(j0 + u) * 64];
                    }
                }
#pragma unroll
                for (int u = 0; u < B; ++u)
                    if (j0 + u < np)
                    {
                        const V xa = lx[cc[u] & 0xffffu], xb = lx[cc[u] >> 16];
#pragma unroll
                        for (int q = 0; q < G; ++q)
                        {
                            if (kDiff)
                            {
                                acc[q] += vv[u].x * (xa[q] - xi[q]);
                                acc[q] += vv[u].y * (xb[q] - xi[q]);
                            }
                            else
                            {
                                acc[q] += vv[u].x * xa[q];
                                acc[q] += vv[u].y * xb[q];
                            }
                        }
                    }
            }
            double dot[G];
#pragma unroll
            for (int q = 0; q < G; ++q) dot[q] = 0.0;
            if (row < nrows)
            {
                const int gi = rb + row;
                V out = acc;
                if (kAp)
                {
                    out = V(T(0));
                    if (is_free(nflag[gi]))
                    {
                        const V p = lx[row];
                        const T d = pdiag[gi];
#pragma unroll
                        for (int q = 0; q < G; ++q)
                        {
                            out[q] = acc[q] + d * p[q];
                            dot[q] = double(p[q]) * double(out[q]);
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < G; ++q) __builtin_nontemporal_store(out[q], y + mcol_at(gi, c0 + q, rshift));
            }
            if (kAp)
            {
#pragma unroll
                for (int q = 0; q < G; ++q)
                {
                    const double d = wave_sum(dot[q]);  // all 64 lanes take part
                    if (lane == 0) sdot[(s - s0) * G + q] = d;
                }
            }
        }
        if (kAp)
        {
            // each column's slots by the first wavefront, as in the apply
            __syncthreads();
            if (tid < 64)
            {
#pragma unroll
                for (int q = 0; q < G; ++q)
                {
                    double sum = 0.0;
                    for (int j = tid; j < s1 - s0; j += 64) sum += sdot[j * G + q];
                    sum = wave_sum(sum);
                    if (tid == 0) part[(size_t(t) << rshift) + c0 + q] = sum;
                }
            }
        }
    }

    namespace
    {
        // bytes per node that size the multi-column product's parts (a
        // workgroup stages half of it: 2 FP64 or 4 FP32 columns, two
        // workgroups a CU; 4 FP64 or 8 FP32 columns in one workgroup a CU
        // measured the same in FP64 and 4% slower in FP32; 64-byte parts
        // measured slower: their larger halo outweighs the fewer re-reads)
        constexpr int kPartNodeBytes = 32;

        template <typename T>
        T* parts_upload(cbs_hip_ctx* c, const std::vector<T>& h)
        {
            void* p = nullptr;
            const size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
            hip_check(hipMalloc(&p, bytes), "hipMalloc product tiles");
            c->allocations.push_back(p);
            c->device_bytes += bytes;
            if (!h.empty())
            {
                // on the context's stream (non-blocking: a default-stream
                // copy is not ordered with its kernels), complete before h goes
                hip_check(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream),
                          "product tiles");
                hip_check(hipStreamSynchronize(c->stream), "product tiles");
            }
            return static_cast<T*>(p);
        }

        // (after the context's stream has been synchronised)
        template <typename T>
        std::vector<T> parts_download(const T* d, size_t n)
        {
            std::vector<T> h(n);
            if (n) hip_check(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "operator tiles");
            return h;
        }

        // Cuts every row tile into the fewest parts of consecutive slices,
        // equal in slice count, whose staging fits the LDS: node_bytes per
        // staged node (the part's rows, then the other columns its entries
        // reference, ascending solver index) and 2 node_bytes per slice (the
        // dot slots).  The parts keep the tile's slices, so the operator's
        // values serve them unchanged; their column pairs are remapped to
        // the part's local ids.  Returns false if a single slice does not fit.
        bool build_parts(cbs_hip_ctx* c, int node_bytes)
        {
            const cbs_hip_sell& m = c->sell;
            hip_check(hipStreamSynchronize(c->stream), "operator tiles");
            const int nt = m.ntiles;
            const auto row_base = parts_download(m.row_base, size_t(nt));
            const auto row_cnt = parts_download(m.row_cnt, size_t(nt));
            const auto halo_ptr = parts_download(m.halo_ptr, size_t(nt) + 1);
            const auto halo = parts_download(m.halo, size_t(halo_ptr[size_t(nt)]));
            const auto slice_ptr = parts_download(m.slice_ptr, size_t(nt) + 1);
            const int nslices = slice_ptr[size_t(nt)];
            const auto slice_off = parts_download(m.slice_off, size_t(nslices) + 1);
            auto cols = parts_download(m.cols, size_t(slice_off[size_t(nslices)] / 2));
            int lds = 65536;
            (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
            struct Part
            {
                int32_t rb, rc, s0;
                std::vector<int32_t> halo;  // solver indices
            };
            std::vector<std::vector<Part>> parts(static_cast<size_t>(nt));
            std::atomic<bool> failed{false};
            host_parallel_for(nt, [&](int64_t tt) {
                const int t = int(tt);
                const int s0 = slice_ptr[size_t(t)], ns = slice_ptr[size_t(t) + 1] - s0;
                const int nrows = row_cnt[size_t(t)], hb = halo_ptr[size_t(t)];
                const int S = nrows + (halo_ptr[size_t(t) + 1] - hb);
                auto solver_of = [&](int f) { return f < nrows ? row_base[size_t(t)] + f : halo[size_t(hb + f - nrows)]; };
                std::vector<int32_t> mark(static_cast<size_t>(S), -1);
                for (int k = 1; k <= ns; ++k)
                {
                    std::vector<Part> cut;
                    std::vector<std::vector<uint32_t>> newcols;
                    bool fits = true;
                    for (int q = 0; q < k && fits; ++q)
                    {
                        const int a = s0 + int(int64_t(ns) * q / k), b = s0 + int(int64_t(ns) * (q + 1) / k);
                        const int r0 = (a - s0) * 64, r1 = std::min(nrows, (b - s0) * 64);
                        Part pt{row_base[size_t(t)] + r0, r1 - r0, a, {}};
                        // local ids: rows of the part first, then the rest
                        std::vector<int32_t> others;
                        for (int sl = a; sl < b; ++sl)
                            for (int64_t w = slice_off[size_t(sl)] / 2; w < slice_off[size_t(sl) + 1] / 2; ++w)
                                for (int h = 0; h < 2; ++h)
                                {
                                    const int f = int((cols[size_t(w)] >> (16 * h)) & 0xffffu);
                                    if ((f < r0 || f >= r1) && mark[size_t(f)] != -2)
                                    {
                                        mark[size_t(f)] = -2;
                                        others.push_back(f);
                                    }
                                }
                        std::sort(others.begin(), others.end(),
                                  [&](int x, int y) { return solver_of(x) < solver_of(y); });
                        const int staged = pt.rc + int(others.size());
                        if (int64_t(staged) * node_bytes + int64_t(b - a) * 2 * node_bytes > lds || staged > 65535)
                            fits = false;
                        for (size_t u = 0; u < others.size(); ++u)
                        {
                            mark[size_t(others[u])] = pt.rc + int32_t(u);
                            pt.halo.push_back(solver_of(others[u]));
                        }
                        if (fits)
                        {
                            std::vector<uint32_t> nc;
                            for (int sl = a; sl < b; ++sl)
                                for (int64_t w = slice_off[size_t(sl)] / 2; w < slice_off[size_t(sl) + 1] / 2; ++w)
                                {
                                    uint32_t word = 0;
                                    for (int h = 0; h < 2; ++h)
                                    {
                                        const int f = int((cols[size_t(w)] >> (16 * h)) & 0xffffu);
                                        const uint32_t l = (f >= r0 && f < r1) ? uint32_t(f - r0) : uint32_t(mark[size_t(f)]);
                                        word |= l << (16 * h);
                                    }
                                    nc.push_back(word);
                                }
                            newcols.push_back(std::move(nc));
                            cut.push_back(std::move(pt));
                        }
                        for (int f : others) mark[size_t(f)] = -1;
                    }
                    if (!fits) continue;
                    // the remapped pairs go where the tile's were (same slices)
                    size_t pos = size_t(slice_off[size_t(s0)] / 2);
                    for (const auto& nc : newcols)
                    {
                        std::copy(nc.begin(), nc.end(), cols.begin() + long(pos));
                        pos += nc.size();
                    }
                    parts[size_t(t)] = std::move(cut);
                    return;
                }
                failed = true;
            });
            if (failed) return false;
            std::vector<int32_t> rb, rc, hp{0}, hl, sp;
            int32_t lmax = 0, max_slices = 0;
            for (int t = 0; t < nt; ++t)
            {
                const auto& pv = parts[size_t(t)];
                for (size_t q = 0; q < pv.size(); ++q)
                {
                    const Part& pt = pv[q];
                    const int32_t s1 = (q + 1 < pv.size()) ? pv[q + 1].s0 : slice_ptr[size_t(t) + 1];
                    rb.push_back(pt.rb);
                    rc.push_back(pt.rc);
                    hl.insert(hl.end(), pt.halo.begin(), pt.halo.end());
                    hp.push_back(int32_t(hl.size()));
                    sp.push_back(pt.s0);
                    lmax = std::max(lmax, pt.rc + int32_t(pt.halo.size()));
                    max_slices = std::max(max_slices, s1 - pt.s0);
                }
            }
            sp.push_back(nslices);
            cbs_hip_sell_parts& P = c->sell_parts;
            P.nparts = int32_t(rb.size());
            P.lmax = lmax;
            P.max_slices = max_slices;
            P.node_bytes = node_bytes;
            P.row_base = parts_upload(c, rb);
            P.row_cnt = parts_upload(c, rc);
            P.halo_ptr = parts_upload(c, hp);
            P.halo = parts_upload(c, hl);
            P.slice_ptr = parts_upload(c, sp);
            P.cols = parts_upload(c, cols);
            return true;
        }

        // The product's tiles: the parts once built, else the operator's row tiles.
        struct SpmmTiles
        {
            SellView v;
            int32_t n, lmax, max_slices;
        };

        SpmmTiles spmm_tiles(const cbs_hip_ctx* c)
        {
            const cbs_hip_sell_parts& P = c->sell_parts;
            if (P.nparts > 0)
                return {{P.row_base, P.row_cnt, P.halo_ptr, P.halo, P.slice_ptr, c->sell.slice_off, P.cols},
                        P.nparts, P.lmax, P.max_slices};
            return {view_of(c->sell), c->sell.ntiles, c->sell.lmax, c->sell.max_slices};
        }

        template <typename T, int G>
        size_t spmm_lds_bytes(const SpmmTiles& m)
        {
            return size_t(m.max_slices) * G * sizeof(double) + size_t(m.lmax) * G * sizeof(T);
        }

        template <typename T, int G>
        bool spmm_allow(const cbs_hip_ctx* c, int rshift)
        {
            if ((1 << rshift) < G) return false;
            int lds = 65536;
            (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
            const size_t bytes = spmm_lds_bytes<T, G>(spmm_tiles(c));
            if (bytes > size_t(lds)) return false;
            for (const void* f : {reinterpret_cast<const void*>(&k_spmm_sell<T, G, true>),
                                  reinterpret_cast<const void*>(&k_spmm_sell<T, G, false>)})
                (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
            (void)hipGetLastError();
            return true;
        }

        template <typename T, int G>
        void spmm_launch(cbs_hip_ctx* c, const T* vals, const T* x, T* y, const T* pdiag, bool ap, double* part,
                         int rshift, const int* stop, const int* skip)
        {
            const SpmmTiles m = spmm_tiles(c);
            const int ngroups = (1 << rshift) / G;
            const dim3 grid(unsigned(m.n) * unsigned(ngroups));
            const size_t lds = spmm_lds_bytes<T, G>(m);
            if (ap)
                hipLaunchKernelGGL((k_spmm_sell<T, G, true>), grid, dim3(kSpmmThreads), lds, c->stream, m.v, vals, x,
                                   y, c->sv_flag, pdiag, part, rshift, ngroups, stop, skip, m.max_slices);
            else
                hipLaunchKernelGGL((k_spmm_sell<T, G, false>), grid, dim3(kSpmmThreads), lds, c->stream, m.v, vals, x,
                                   y, c->sv_flag, pdiag, part, rshift, ngroups, stop, skip, m.max_slices);
        }
    }

    int sell_spmm_groups(cbs_hip_ctx* c, bool fp32, int rshift)
    {
        // parts are built once; a mesh whose single slices exceed the budget
        // keeps the whole row tiles (node_bytes < 0 records the attempt)
        if (c->sell.ntiles > 0 && c->sell_parts.nparts == 0 && c->sell_parts.node_bytes == 0 &&
            !build_parts(c, kPartNodeBytes))
            c->sell_parts.node_bytes = -1;
        if (fp32)
        {
            if (spmm_allow<float, 4>(c, rshift)) return 4;
            if (spmm_allow<float, 2>(c, rshift)) return 2;
            return spmm_allow<float, 1>(c, rshift) ? 1 : 0;
        }
        if (spmm_allow<double, 2>(c, rshift)) return 2;
        return spmm_allow<double, 1>(c, rshift) ? 1 : 0;
    }

    int sell_spmm_tiles(const cbs_hip_ctx* c) { return spmm_tiles(c).n; }

    void sell_spmm64(cbs_hip_ctx* c, int groups, const double* x, double* y, bool ap, double* part, int rshift,
                     const int* stop, const int* skip)
    {
        if (c->sell.ntiles == 0) return;
        if (groups == 4) spmm_launch<double, 4>(c, c->sell.v64, x, y, c->sv_pdiag, ap, part, rshift, stop, skip);
        else if (groups == 2) spmm_launch<double, 2>(c, c->sell.v64, x, y, c->sv_pdiag, ap, part, rshift, stop, skip);
        else spmm_launch<double, 1>(c, c->sell.v64, x, y, c->sv_pdiag, ap, part, rshift, stop, skip);
    }

    void sell_spmm32(cbs_hip_ctx* c, int groups, const float* x, float* y, const float* pdiagf, double* part,
                     int rshift, const int* stop, const int* skip)
    {
        if (c->sell.ntiles == 0) return;
        if (groups == 8) spmm_launch<float, 8>(c, c->sell.v32, x, y, pdiagf, true, part, rshift, stop, skip);
        else if (groups == 4) spmm_launch<float, 4>(c, c->sell.v32, x, y, pdiagf, true, part, rshift, stop, skip);
        else if (groups == 2) spmm_launch<float, 2>(c, c->sell.v32, x, y, pdiagf, true, part, rshift, stop, skip);
        else spmm_launch<float, 1>(c, c->sell.v32, x, y, pdiagf, true, part, rshift, stop, skip);
    }

    void sell_gather_columns(cbs_hip_ctx* c, const double* node_order, double* solver_order, int rshift)
    {
        hipLaunchKernelGGL(k_sol_gather_cols, dim3(grid_for(int64_t(mcol_size(c->N, rshift)))), dim3(kBlock), 0, c->stream,
                           c->N, rshift, c->sell.gid, node_order, solver_order);
    }

    void sell_scatter_columns(cbs_hip_ctx* c, const double* solver_order, double* node_order, int rshift)
    {
        hipLaunchKernelGGL(k_sol_scatter_cols, dim3(grid_for(int64_t(mcol_size(c->N, rshift)))), dim3(kBlock), 0, c->stream,
                           c->N, rshift, c->sell.gid, solver_order, node_order);
    }
}
