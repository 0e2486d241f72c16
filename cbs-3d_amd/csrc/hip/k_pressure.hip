// Matrix-free pressure operator and Jacobi-PCG on the GPU.
//
// FP64 path: MatrixVectorCalc::pressureMultiply (src/linalg/
// MatrixVectorCalc.cpp:125-205) and ConjugateGradient::solvePressure
// (src/linalg/ConjugateGradient.cpp:584-869).
//
// Two forms of the operator:
//  - exact (k_apply_tiled, MatrixVectorCalc::pressureMultiply): applied
//    element by element, tile by tile.  One workgroup stages its tile's node
//    values in LDS and walks the tile's elements colour by colour (no two
//    elements of a colour share a node, so a barrier between colours is the
//    only synchronisation).  Nodes shared with other tiles get one value per
//    colour in a side buffer that a second kernel sums in colour order, so
//    every node accumulates the reference's sequence of additions and the
//    result is bit-identical;
//  - assembled (k_sell.hip, inside the PCG): the same operator with each
//    edge's element coefficients summed once per solve, streamed as sliced
//    ELLPACK with LDS-staged x.  Sums are regrouped (rounding-level
//    difference, deterministic).  The PCG's Ap = free ? acc + pdiag p : 0
//    and the (p, Ap) partial of every tile are fused into its kernel.
//
// The CG vector work is one streaming kernel per phase writing per-block
// partials, followed by a one-block kernel that combines them in a fixed
// order (deterministic) and updates the device-resident scalars (alpha,
// beta, norms, convergence flag).  The host only reads the convergence flag
// between chunks of iterations.  The dot products are tree sums, so they
// differ from the reference's ascending serial sums by rounding only.
//
// Mixed form: FP32 apply, FP32 search direction and Jacobi, FP64 iterate and
// residual, FP64 reductions, with reliable updates of the residual from the
// FP64 operator (see solve_pressure_mixed).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>

#include "device_common.hpp"
#include "pcg_host.hpp"

namespace cbsdev
{
    // Event pairs around apply sweeps; collected after each synchronisation.
    // Every kTimerStride-th apply of a solve is timed (an event record costs
    // a few microseconds of queue time), and `idx` is the apply's position
    // in its chunk of iterations: applies launched after the solve finished
    // return at once and are dropped at collection (timer_collect's
    // keep_below = the iterations the chunk actually performed).
    constexpr int kTimerStride = 8;

    bool timer_begin(cbs_hip_ctx* c, int precision, int idx, int64_t seq)
    {
        if (!c->profile || seq % kTimerStride != 0) return false;
        if (size_t(c->events_used + 2) > c->events.size())
        {
            const size_t old = c->events.size();
            c->events.resize(old + 512);
            for (size_t i = old; i < c->events.size(); ++i) hip_check(hipEventCreate(&c->events[i]), "hipEventCreate");
        }
        // the apply launch records the pair at its kernel's start and end
        c->timer_pending = true;
        c->event_prec.push_back(precision);
        c->event_idx.push_back(idx);
        return true;
    }

    void timer_end(cbs_hip_ctx* c, bool begun)
    {
        if (!begun) return;
        if (c->timer_pending)
        {
            // no apply was launched (an empty operator): an empty pair
            hip_check(hipEventRecord(c->events[size_t(c->events_used)], c->stream), "hipEventRecord");
            hip_check(hipEventRecord(c->events[size_t(c->events_used) + 1], c->stream), "hipEventRecord");
            c->timer_pending = false;
        }
        c->events_used += 2;
    }

    void timer_collect(cbs_hip_ctx* c, int keep_below)
    {
        for (int i = 0; i + 1 < c->events_used; i += 2)
        {
            if (c->event_idx[size_t(i / 2)] >= keep_below) continue;
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, c->events[size_t(i)], c->events[size_t(i) + 1]), "hipEventElapsedTime");
            const int p = c->event_prec[size_t(i / 2)] == CBS_PCG_MIXED ? 1 : 0;
            c->apply_ms[p] += ms;
            c->apply_n[p] += 1;
        }
        c->events_used = 0;
        c->event_prec.clear();
        c->event_idx.clear();
    }

    void phase_mark(cbs_hip_ctx* c, int label, int idx)
    {
        if (size_t(c->ph_used) + 1 > c->ph_events.size())
        {
            const size_t old = c->ph_events.size();
            c->ph_events.resize(old + 256);
            c->ph_label.resize(old + 256);
            c->ph_idx.resize(old + 256);
            for (size_t i = old; i < c->ph_events.size(); ++i)
                hip_check(hipEventCreate(&c->ph_events[i]), "hipEventCreate");
        }
        hip_check(hipEventRecord(c->ph_events[size_t(c->ph_used)], c->stream), "hipEventRecord");
        c->ph_label[size_t(c->ph_used)] = label;
        c->ph_idx[size_t(c->ph_used)] = idx;
        ++c->ph_used;
    }

    void phase_reset(cbs_hip_ctx* c)
    {
        c->ph_used = 0;
        c->ph_sampled_iters = 0;
        for (int p = 0; p < 5; ++p) c->ph_sampled_ms[p] = c->ph_direct_ms[p] = 0.0;
    }

    // keep_below: sampled groups at a chunk position at or past it ran after
    // the solve had finished (they return at once) and are dropped.
    void phase_collect(cbs_hip_ctx* c, int keep_below)
    {
        int open = -1;
        bool keep = false, sampled = false;
        for (int i = 0; i < c->ph_used; ++i)
        {
            const int lab = c->ph_label[size_t(i)];
            if (lab < 0)
            {
                open = i;
                sampled = c->ph_idx[size_t(i)] >= 0;
                keep = !sampled || c->ph_idx[size_t(i)] < keep_below;
                if (keep && sampled) ++c->ph_sampled_iters;
                continue;
            }
            if (open < 0) continue;
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, c->ph_events[size_t(open)], c->ph_events[size_t(i)]),
                      "hipEventElapsedTime");
            if (keep) (sampled ? c->ph_sampled_ms : c->ph_direct_ms)[lab] += ms;
            open = i;
        }
        c->ph_used = 0;
    }

    void phase_fill(cbs_hip_ctx* c, cbs_cg_result* res, double setup_seconds)
    {
        res->setup_seconds = setup_seconds;
        double sec[5] = {0, 0, 0, 0, 0};
        if (c->profile)
        {
            hip_check(hipStreamSynchronize(c->stream), "stream sync");
            phase_collect(c);
            // one-off groups as measured; the iterations' wall time split in
            // proportion to each phase's device time in the sampled
            // iterations (their event marks add launch gaps, so scaled
            // directly they would overstate a launch-bound solve)
            double direct = 0.0, sampled = 0.0;
            for (int p = 0; p < 5; ++p)
            {
                direct += 1e-3 * c->ph_direct_ms[p];
                sampled += c->ph_sampled_ms[p];
            }
            const double iter_wall = std::max(0.0, res->total_seconds - setup_seconds - direct);
            for (int p = 0; p < 5; ++p)
                sec[p] = 1e-3 * c->ph_direct_ms[p] + (sampled > 0.0 ? iter_wall * c->ph_sampled_ms[p] / sampled : 0.0);
        }
        res->matvec_seconds = sec[PH_MATVEC];
        res->dot_seconds = sec[PH_DOT];
        res->preconditioner_seconds = sec[PH_PRECOND];
        res->vector_update_seconds = sec[PH_VECTOR];
        res->constraint_seconds = sec[PH_CONSTRAINT];
    }

    // ------------------------------------------------------------------
    // PCG iterations as HIP graphs.  A block of kGraphIters iterations is
    // captured once per solver form and replayed as often as a chunk needs
    // (a chunk is rounded up to whole blocks; iterations launched after the
    // solve finished return at once, as in the eager loop).  The block holds
    // the same kernels with the same arguments in the same order as the
    // eager loop, so both give the same bits.  No memset, event or host work
    // is captured.  With the apply timer on, one iteration in the middle of
    // a chunk is launched eagerly with the event pair around its apply
    // (events recorded inside a graph cannot be timed on this runtime) and
    // the rest replays the block.  If capture fails, the context falls back to eager launches.
    // ------------------------------------------------------------------
    void destroy_graphs(cbs_hip_ctx* c)
    {
        for (auto& g : c->pcg_graph)
        {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g = cbs_hip_ctx::PcgGraph{};
        }
    }

    // ------------------------------------------------------------------
    // Exact tiled apply.  Slots of colours without a term stay +0.0 (cleared
    // once at allocation and never written); an accumulator that starts at
    // +0.0 can never become -0.0, so adding those zeros leaves every sum
    // bit-identical.  The side buffer is colour-major (sb[c*S + s]) so the
    // colour-order sum reads it coalesced.
    // ------------------------------------------------------------------
    __device__ inline void tile_scatter(double v, int l, size_t col_off, double* ly, const int* ls, double* sb)
    {
        const int srow = ls[l];
        if (srow < 0) ly[l] += v;
        else sb[col_off + size_t(srow)] = v;
    }

    __global__ __launch_bounds__(1024) void k_apply_tiled(const int32_t* __restrict__ tnode_ptr,
                                                         const int32_t* __restrict__ tnodes,
                                                         const int32_t* __restrict__ tsidx,
                                                         const int32_t* __restrict__ tcol_ptr,
                                                         const ushort4* __restrict__ lidx,
                                                         const double* __restrict__ g, int64_t F, int ncolor,
                                                         int S, const double* __restrict__ x,
                                                         double* __restrict__ y, double* __restrict__ sb,
                                                         const int* __restrict__ done)
    {
        if (*done) return;
        extern __shared__ __align__(16) unsigned char smem[];
        const int t = blockIdx.x;
        const int nb = tnode_ptr[t];
        const int L = tnode_ptr[t + 1] - nb;
        double* lx = reinterpret_cast<double*>(smem);
        double* ly = lx + L;
        int* ls = reinterpret_cast<int*>(ly + L);
        for (int l = threadIdx.x; l < L; l += blockDim.x)
        {
            lx[l] = x[tnodes[nb + l]];
            ly[l] = 0.0;
            ls[l] = tsidx[nb + l];
        }
        __syncthreads();
        const int32_t* cp = tcol_ptr + size_t(t) * (ncolor + 1);
        for (int col = 0; col < ncolor; ++col)
        {
            const int e1 = cp[col + 1];
            const size_t off = size_t(col) * size_t(S);
            for (int k = cp[col] + threadIdx.x; k < e1; k += blockDim.x)
            {
                const ushort4 q = lidx[k];
                const double x1 = lx[q.x], x2 = lx[q.y], x3 = lx[q.z], x4 = lx[q.w];
                const double a12 = g[k], a13 = g[F + k], a14 = g[2 * F + k];
                const double a23 = g[3 * F + k], a24 = g[4 * F + k], a34 = g[5 * F + k];
                tile_scatter(a12 * x2 + a13 * x3 + a14 * x4, q.x, off, ly, ls, sb);
                tile_scatter(a12 * x1 + a23 * x3 + a24 * x4, q.y, off, ly, ls, sb);
                tile_scatter(a13 * x1 + a23 * x2 + a34 * x4, q.z, off, ly, ls, sb);
                tile_scatter(a14 * x1 + a24 * x2 + a34 * x3, q.w, off, ly, ls, sb);
            }
            __syncthreads();
        }
        for (int l = threadIdx.x; l < L; l += blockDim.x)
            if (ls[l] < 0) y[tnodes[nb + l]] = ly[l];
    }

    __global__ void k_apply_shared(const int32_t* __restrict__ snode, int S, int ncolor,
                                   const double* __restrict__ sb, double* __restrict__ y,
                                   const int* __restrict__ done)
    {
        if (*done) return;
        const int s = blockIdx.x * kBlock + threadIdx.x;
        if (s >= S) return;
        double acc = 0.0;
        for (int col = 0; col < ncolor; ++col) acc += sb[size_t(col) * size_t(S) + size_t(s)];
        y[snode[s]] = acc;
    }

    // Address of an int flag of segment 0's scalars; segment s's copy is
    // sizeof(DevScalars)/sizeof(int) ints further on per segment.
    inline const int* seg_flag(const cbs_hip_ctx* c, const int DevScalars::*field)
    {
        return &(c->scal->*field);
    }

    void configure_tiled_apply(cbs_hip_ctx* c)
    {
        const int b_exact = c->tile_lmax * int(2 * sizeof(double) + sizeof(int));
        // Allows dynamic LDS beyond the default 64 KB.  Best effort: where the
        // runtime needs no opt-in the call may fail harmlessly; a launch
        // beyond the real limit reports its own error.
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_apply_tiled),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, b_exact);
        (void)hipGetLastError();
        sell_configure(c);
        c->ndotpart = c->sell.ntiles;
    }

    // acc <- the element part of A x, bit-identical to the reference's
    // colour-ordered loop (acc must be zero on nodes no fluid element
    // touches; every other node is overwritten).
    void launch_apply(cbs_hip_ctx* c, const double* x, double* acc, const int* done)
    {
        if (c->ntiles == 0) return;
        ensure_gstif_tiles(c);
        const size_t lds = size_t(c->tile_lmax) * (2 * sizeof(double) + sizeof(int));
        hipLaunchKernelGGL(k_apply_tiled, dim3(c->ntiles), dim3(c->tile_threads), lds, c->stream, c->tnode_ptr,
                           c->tnodes, c->tsidx, c->tcol_ptr, c->lidx, c->gstif_t, int64_t(c->F), c->ncolor,
                           c->nshared, x, acc, c->sb, done);
        if (c->nshared > 0)
            hipLaunchKernelGGL(k_apply_shared, dim3(grid_for(c->nshared)), dim3(kBlock), 0, c->stream, c->snode,
                               c->nshared, c->ncolor, c->sb, acc, done);
    }

    // The off-diagonal part of A x from the assembled operator, x and acc in
    // solver order (rows of non-free nodes are empty: acc = 0 there).
    void launch_apply_fast(cbs_hip_ctx* c, const double* x, double* acc)
    {
        sell_apply64(c, x, acc, false, seg_flag(c, &DevScalars::pad0));
    }

    __global__ void k_fillf_pressure(float* __restrict__ p, int64_t n)
    {
        const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
        if (i < n) p[i] = 0.0f;
    }

    // y = acc + pdiag*x, then identity rows at fixed-pressure nodes.
    __global__ void k_apply_diag(const double* __restrict__ acc, const double* __restrict__ pdiag,
                                 const double* __restrict__ x, double* __restrict__ y, int n)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n) y[i] = acc[i] + pdiag[i] * x[i];
    }

    __global__ void k_fixed_rows(const int32_t* __restrict__ node, int m, const double* __restrict__ x,
                                 double* __restrict__ y)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < m) y[node[i]] = x[node[i]];
    }

    void launch_apply_finish(cbs_hip_ctx* c, const double* x, double* y)
    {
        hipLaunchKernelGGL(k_apply_diag, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, y, c->pdiag, x, y, c->N);
        if (c->npres_bc > 0)
        {
            hipLaunchKernelGGL(k_fixed_rows, dim3(grid_for(c->npres_bc)), dim3(kBlock), 0, c->stream, c->pbc_node,
                               c->npres_bc, x, y);
        }
    }

    // gstif = gstifE * delte (TimeStep.cpp:873-901) in one element order: the
    // colour-sweep slots (with the FP32 copy of the multi-RHS mixed form) or
    // the tiles.
    __global__ void k_gstif(const int32_t* __restrict__ slot_elem, const double* __restrict__ gE,
                            const double* __restrict__ delte, double* __restrict__ g, float* __restrict__ gf,
                            int64_t F)
    {
        const int64_t k = int64_t(blockIdx.x) * kBlock + threadIdx.x;
        if (k >= F) return;
        const double dt = delte[slot_elem[k]];
#pragma unroll
        for (int p = 0; p < 6; ++p)
        {
            const double v = gE[p * F + k] * dt;
            g[p * F + k] = v;
            if (gf) gf[p * F + k] = float(v);
        }
    }

    // An LHS update makes both orders of gstif stale.  They are formed when
    // next read: the tile-order planes by the exact (element) apply, the
    // slot-order planes by field reads.  The assembled operator does not
    // read them: its values are dt times the gstifE sums, or sums of
    // gstifE * delte formed as it gathers them.
    void launch_gstif(cbs_hip_ctx* c)
    {
        c->gstif_tiles_stale = true;
        c->gstif_slots_stale = true;
        c->gstif_manual = false;
    }

    void ensure_gstif_tiles(cbs_hip_ctx* c)
    {
        if (!c->gstif_tiles_stale || c->F == 0) return;
        hipLaunchKernelGGL(k_gstif, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->tslot_elem, c->gstifE_t,
                           c->delte, c->gstif_t, static_cast<float*>(nullptr), int64_t(c->F));
        c->gstif_tiles_stale = false;
    }

    void ensure_gstif_slots(cbs_hip_ctx* c)
    {
        if (!c->gstif_slots_stale || c->F == 0) return;
        hipLaunchKernelGGL(k_gstif, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->slot_elem, c->gstifE_s,
                           c->delte, c->gstif_s, static_cast<float*>(nullptr), int64_t(c->F));
        c->gstif_slots_stale = false;
    }

    // ------------------------------------------------------------------
    // Constraints: impose_pressure_constraints (ConjugateGradient.cpp:250-270)
    // on the non-free nodes only (free nodes are untouched by it).
    // ------------------------------------------------------------------
    __global__ void k_constraints(const int32_t* __restrict__ node, const int32_t* __restrict__ mode,
                                  const double* __restrict__ val, int m, double* __restrict__ pres,
                                  double* __restrict__ rhs1, double* __restrict__ pdiag)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= m) return;
        const int ip = node[i];
        if (mode[i] == 0)
        {
            pres[ip] = 0.0;
            rhs1[ip] = 0.0;
            pdiag[ip] = 0.0;
        }
        else
        {
            const double v = val ? val[i] : 0.0;  // val null: homogeneous (adjoint) constraints
            pres[ip] = v;
            rhs1[ip] = v;
            pdiag[ip] = 1.0;
        }
    }

    void launch_constraints(cbs_hip_ctx* c)
    {
        if (c->nnonfree > 0)
            hipLaunchKernelGGL(k_constraints, dim3(grid_for(c->nnonfree)), dim3(kBlock), 0, c->stream, c->nf_node,
                               c->nf_mode, c->constraint_zero ? nullptr : c->nf_val, c->nnonfree, c->pres, c->rhs1,
                               c->pdiag);
    }

    // Preconditioner checks of apply_preconditioner (ConjugateGradient.cpp:302-369).
    // (the node's id is read only on a failure)
    __device__ inline void jacobi_check(int precond, double diag, double r, const int32_t* __restrict__ gid, int i,
                                        unsigned long long* pend)
    {
        if (precond == 1 && (!isfinite(diag) || diag <= 1.0e-30))
            atomicMin(pend, err_key(ST_CG, 1, uint32_t(gid[i] + 1), CK_BAD_DIAG));
        else if (!isfinite(r))
            atomicMin(pend, err_key(ST_CG, 1, uint32_t(gid[i] + 1), CK_BAD_R));
    }

    // ------------------------------------------------------------------
    // FP64 Jacobi-PCG (ConjugateGradient.cpp:584-869).  Per iteration: the
    // apply (Ap and the (p,Ap) tile partials; when fused, tune_fusion, it
    // also forms the direction p = z + beta p while it stages it:
    // sell_apply64_dir), k_cg_update (alpha from the tile partials, then the
    // updates), k_cg_update_fin, and k_cg_direction when not fused.
    // Partials: part[b*4 + q].
    // ------------------------------------------------------------------

    // Partials of block b are part[b*4 + q]; the *_fin kernels run one block
    // per segment over that segment's block range (BlockMap::seg_ptr).  The
    // update and direction kernels take logical block b = xcd_block(...) as
    // the apply takes its row tiles, so each XCD works on one contiguous
    // node range in all of them and reads from its own L2 part of what the
    // previous kernel wrote there (the mixed iteration 1% faster at
    // cavity160; the same blocks, partials and sums).

    // Initial residual, preconditioned residual and first direction
    // (ConjugateGradient.cpp:634-723):
    //   r = free ? rhs1 - (acc + pdiag*pres) : 0,  z = r/pdiag,  p = z.
    __global__ __launch_bounds__(kBlock) void k_cg_init(BlockMap bm, const uint8_t* __restrict__ nflag,
                                                         const double* __restrict__ rhs1,
                                                         const double* __restrict__ pdiag,
                                                         const double* __restrict__ pres, double* __restrict__ acc,
                                                         double* __restrict__ r, double* __restrict__ p,
                                                         int precond, double* __restrict__ part, DevScalars* s,
                                                         const int32_t* __restrict__ gid)
    {
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        DevScalars* ss = s + bm.seg[blockIdx.x];
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
        {
            double ri = 0.0, pi = 0.0;
            if (is_free(nflag[i]))
            {
                const double ap = acc[i] + pdiag[i] * pres[i];
                ri = rhs1[i] - ap;
                const double d = pdiag[i];
                jacobi_check(precond, d, ri, gid, i, &ss->pending_err);
                const double zi = (precond == 1) ? ri / d : ri;
                pi = zi;
                v[0] += ri * ri;
                v[1] += ri * zi;
                const double a = fabs(ri);
                m = (m < a) ? a : m;
            }
            r[i] = ri;
            p[i] = pi;
            acc[i] = 0.0;
        }
        block_sum<2>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x == 0)
        {
            part[blockIdx.x * 4 + 0] = v[0];
            part[blockIdx.x * 4 + 1] = v[1];
            part[blockIdx.x * 4 + 2] = m;
        }
    }

    __global__ __launch_bounds__(kBlock) void k_cg_init_fin(const double* __restrict__ part,
                                                             const int32_t* __restrict__ seg_ptr, DevScalars* sbase)
    {
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int b0 = seg_ptr[blockIdx.x], nb = seg_ptr[blockIdx.x + 1] - b0;
        const double* pp = part + size_t(b0) * 4;
        double rr, rz, mx;
        reduce3_partials(pp, nb, lds, ldm, rr, rz, mx);
        if (threadIdx.x != 0) return;
        DevScalars* s = sbase + blockIdx.x;
        s->initial_l2 = sqrt(rr);
        s->max_abs = mx;
        s->iter = 0;
        if (s->initial_l2 <= 1.0e-300)
        {
            s->final_l2 = s->initial_l2;
            s->rel_l2 = 0.0;
            s->converged = 1;
            s->done = 1;
            return;
        }
        s->final_l2 = s->initial_l2;
        s->rel_l2 = 1.0;
        const bool c1 = mx <= s->abs_tol, c2 = 1.0 <= s->rel_tol;
        const bool conv = (s->conv_test == 1) ? c1 : (s->conv_test == 2) ? c2 : (c1 || c2);
        if (conv)
        {
            s->converged = 1;
            s->done = 1;
            return;
        }
        if (s->pending_err != ~0ull)
        {
            raise_error(sbase, s->pending_err);
            s->done = 1;
            return;
        }
        if (rz <= 1.0e-300 || !isfinite(rz))
        {
            raise_error(sbase, err_key(ST_CG, 2, 0, CK_RZ_INIT));
            s->done = 1;
            return;
        }
        s->rz_old = rz;
    }

    // pAp of the block's segment from the fused apply's tile partials and
    // alpha = rz_old / pAp (ConjugateGradient.cpp:727-758).  Every block of
    // an update kernel computes them in the same order (identical bits); the
    // segment's first block records pAp and alpha and, on breakdown, raises
    // the error and ends the solve (kMixed: the inner iteration too).
    // Returns false on breakdown.  Called by all threads of the block.
    template <bool kMixed>
    __device__ inline bool block_alpha(BlockMap bm, const double* __restrict__ dotpart,
                                       const int32_t* __restrict__ seg_tile, DevScalars* sbase, double* alpha,
                                       double (*lds)[kBlock / 64])
    {
        const int blk = xcd_block(int(gridDim.x));  // logical block (same XCD as the apply tiles of its rows)
        const int sg = bm.seg[blk];
        DevScalars* s = sbase + sg;
        const int t0 = seg_tile[sg];
        const double pap = reduce_partials(dotpart + t0, seg_tile[sg + 1] - t0, 1, 0, lds);
        const bool bad = fabs(pap) <= 1.0e-300 || !isfinite(pap);
        const bool first = blk == 0 || bm.seg[blk - 1] != sg;
        if (first && threadIdx.x == 0)
        {
            s->pAp = pap;
            if (bad)
            {
                raise_error(sbase, err_key(ST_CG, 3, uint32_t(s->iter + 1), CK_BREAKDOWN));
                if (kMixed) s->f_done = 1;
                s->done = 1;
            }
            else
            {
                s->alpha = s->rz_old / pap;
            }
        }
        *alpha = s->rz_old / pap;
        return !bad;
    }

    // alpha (block_alpha), then pres += alpha p, r -= alpha Ap on free nodes;
    // ||r||2, max|r|; z = r/pdiag, (r,z) (ConjugateGradient.cpp:727-800).
    // The iterate is updated every other iteration: when the previous
    // iteration's alpha p is pending, pres = (pres + alpha_prev p_prev) +
    // alpha p, the same two additions in the same order as updating at every
    // iteration; otherwise this iteration's term is left pending
    // (k_cg_update_fin records it).  The direction of iteration j is p0 for
    // even j, p1 for odd j.
    __global__ __launch_bounds__(kBlock) void k_cg_update(BlockMap bm, const uint8_t* __restrict__ nflag,
                                                           const double* __restrict__ pdiag,
                                                           const double* __restrict__ p0,
                                                           const double* __restrict__ p1,
                                                           const double* __restrict__ ap,
                                                           double* __restrict__ pres, double* __restrict__ r,
                                                           int precond, double* __restrict__ part, DevScalars* s,
                                                           const int32_t* __restrict__ gid,
                                                           const double* __restrict__ dotpart,
                                                           const int32_t* __restrict__ seg_tile)
    {
        const int blk = xcd_block(int(gridDim.x));  // logical block (same XCD as the apply tiles of its rows)
        DevScalars* ss = s + bm.seg[blk];
        if (ss->done) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const bool odd = ss->iter & 1, pending = ss->x_pending;
        const double alpha_prev = ss->alpha_prev;
        const double* __restrict__ p = odd ? p1 : p0;
        const double* __restrict__ pp = odd ? p0 : p1;
        // Every load of a node is issued before its free flag is tested, so
        // they go out together; the thread's first kPre nodes are loaded
        // before block_alpha, whose reduction latency they overlap.
        struct Vals
        {
            uint8_t f;
            double r0, api, d, x0, pi, ppi;
        };
        auto load = [&](int i) {
            Vals n{nflag[i], r[i], ap[i], pdiag[i], 0.0, 0.0, 0.0};
            if (pending)
            {
                n.x0 = pres[i];
                n.pi = p[i];
                n.ppi = pp[i];
            }
            return n;
        };
        constexpr int kPre = 2;
        const int i0 = bm.beg[blk] + int(threadIdx.x), iend = bm.end[blk];
        Vals pre[kPre];
#pragma unroll
        for (int k = 0; k < kPre; ++k)
            if (i0 + k * kBlock < iend) pre[k] = load(i0 + k * kBlock);
        double alpha;
        if (!block_alpha<false>(bm, dotpart, seg_tile, s, &alpha, lds)) return;
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        auto step = [&](int i, const Vals& n) {
            if (!is_free(n.f)) return;
            if (pending) pres[i] = (n.x0 + alpha_prev * n.ppi) + alpha * n.pi;
            const double ri = n.r0 - alpha * n.api;
            r[i] = ri;
            jacobi_check(precond, n.d, ri, gid, i, &ss->pending_err);
            const double zi = (precond == 1) ? ri / n.d : ri;
            v[0] += ri * ri;
            v[1] += ri * zi;
            const double a = fabs(ri);
            m = (m < a) ? a : m;
        };
#pragma unroll
        for (int k = 0; k < kPre; ++k)
            if (i0 + k * kBlock < iend) step(i0 + k * kBlock, pre[k]);
#pragma unroll 2
        for (int i = i0 + kPre * kBlock; i < iend; i += kBlock) step(i, load(i));
        block_sum<2>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x == 0)
        {
            part[blk * 4 + 0] = v[0];
            part[blk * 4 + 1] = v[1];
            part[blk * 4 + 2] = m;
        }
    }

    // After an iteration's update: the pending term was added with this
    // iteration's (clear it), or this iteration's alpha p is left pending.
    __device__ inline void defer_iterate_update(DevScalars* s)
    {
        if (s->x_pending)
        {
            s->x_pending = 0;
        }
        else
        {
            s->x_pending = 1;
            s->alpha_prev = s->alpha;
        }
    }

    // Convergence test and beta (ConjugateGradient.cpp:796-852).
    __global__ __launch_bounds__(kBlock) void k_cg_update_fin(const double* __restrict__ part,
                                                               const int32_t* __restrict__ seg_ptr,
                                                               DevScalars* sbase)
    {
        DevScalars* s = sbase + blockIdx.x;
        if (s->done) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int b0 = seg_ptr[blockIdx.x], nb = seg_ptr[blockIdx.x + 1] - b0;
        const double* pp = part + size_t(b0) * 4;
        double rr, rz, mx;
        reduce3_partials(pp, nb, lds, ldm, rr, rz, mx);
        if (threadIdx.x != 0) return;
        s->iter += 1;
        defer_iterate_update(s);
        s->final_l2 = sqrt(rr);
        s->max_abs = mx;
        s->rel_l2 = s->final_l2 / s->initial_l2;
        const bool c1 = mx <= s->abs_tol, c2 = s->rel_l2 <= s->rel_tol;
        const bool conv = (s->conv_test == 1) ? c1 : (s->conv_test == 2) ? c2 : (c1 || c2);
        if (conv)
        {
            s->converged = 1;
            s->done = 1;
            return;
        }
        if (s->pending_err != ~0ull)
        {
            raise_error(sbase, s->pending_err);
            s->done = 1;
            return;
        }
        if (rz <= 1.0e-300 || !isfinite(rz))
        {
            raise_error(sbase, err_key(ST_CG, 4, uint32_t(s->iter), CK_RZ_UPD));
            s->done = 1;
            return;
        }
        s->beta = rz / s->rz_old;
        s->rz_old = rz;
        if (s->iter >= s->max_iter) s->done = 1;
    }

    // p_new = free ? r/pdiag + beta p : 0, from the direction of the
    // iteration just completed (iter - 1) into the buffer of the next (iter)
    // (ConjugateGradient.cpp:831-849), when the apply does not form it.
    __global__ __launch_bounds__(kBlock) void k_cg_direction(BlockMap bm, const uint8_t* __restrict__ nflag,
                                                              const double* __restrict__ pdiag,
                                                              const double* __restrict__ r, double* __restrict__ p0,
                                                              double* __restrict__ p1, int precond,
                                                              const DevScalars* s)
    {
        const int blk = xcd_block(int(gridDim.x));  // logical block (same XCD as the apply tiles of its rows)
        const DevScalars* ss = s + bm.seg[blk];
        if (ss->done) return;
        const double beta = ss->beta;
        const bool odd = ss->iter & 1;
        const double* __restrict__ p = odd ? p0 : p1;
        double* __restrict__ pn = odd ? p1 : p0;
#pragma unroll 2
        for (int i = bm.beg[blk] + int(threadIdx.x); i < bm.end[blk]; i += kBlock)
        {
            const uint8_t f = nflag[i];
            const double ri = r[i], di = pdiag[i], po = p[i];
            double pi = 0.0;
            if (is_free(f))
            {
                const double zi = (precond == 1) ? ri / di : ri;
                pi = zi + beta * po;
            }
            pn[i] = pi;
        }
    }

    // Times the two ways of forming the PCG iteration's direction on the
    // solve's scratch vectors and keeps the faster (once per context and
    // form, unless pinned by cbs_hip_set_direction_fusion): sep() launches
    // the apply and the direction kernel, fused() the apply that forms the
    // direction while it stages it.  The fused apply saves a launch and a
    // pass over r and p, but gathers four values per staged node instead of
    // one; on small operators the launch dominates (cavity20: 231 against
    // 282 FP64 steps/s), on the 160^3 cavity the two measure alike within
    // 2%, either way from run to run.  Fused is kept only when it is at
    // least 3% faster, so that such ties settle on the separate kernels.
    // Both give the same directions bit for bit.
    template <typename Sep, typename Fused>
    void tune_fusion(cbs_hip_ctx* c, int form, Sep sep, Fused fused)
    {
        cbs_hip_sell::Config& cf = c->sell.cfg[form];
        if (cf.fuse_pin >= 0)
        {
            cf.fuse = cf.fuse_pin;
            return;
        }
        if (cf.fuse >= 0 || c->sell.ntiles == 0) return;
        hipEvent_t e[3];
        for (hipEvent_t& ev : e) hip_check(hipEventCreate(&ev), "hipEventCreate");
        sep();
        fused();
        hip_check(hipEventRecord(e[0], c->stream), "hipEventRecord");
        for (int r = 0; r < 3; ++r) sep();
        hip_check(hipEventRecord(e[1], c->stream), "hipEventRecord");
        for (int r = 0; r < 3; ++r) fused();
        hip_check(hipEventRecord(e[2], c->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e[2]), "hipEventSynchronize");
        float ms[2] = {0.0f, 0.0f};
        hip_check(hipEventElapsedTime(&ms[0], e[0], e[1]), "hipEventElapsedTime");
        hip_check(hipEventElapsedTime(&ms[1], e[1], e[2]), "hipEventElapsedTime");
        for (hipEvent_t& ev : e) (void)hipEventDestroy(ev);
        cf.fuse_ms[0] = ms[0] / 3.0f;
        cf.fuse_ms[1] = ms[1] / 3.0f;
        cf.fuse = (ms[1] < 0.97f * ms[0]) ? 1 : 0;
    }

    // x += alpha_prev p on free nodes where an iterate update is pending (p:
    // the direction of the last completed iteration, iter - 1).
    template <typename TP>
    __global__ __launch_bounds__(kBlock) void k_flush_iterate(BlockMap bm, const uint8_t* __restrict__ nflag,
                                                               const TP* __restrict__ p0, const TP* __restrict__ p1,
                                                               double* __restrict__ x, const DevScalars* s)
    {
        const DevScalars* ss = s + bm.seg[blockIdx.x];
        if (!ss->x_pending) return;
        const double alpha = ss->alpha_prev;
        const TP* __restrict__ p = ((ss->iter - 1) & 1) ? p1 : p0;
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
            if (is_free(nflag[i])) x[i] += alpha * double(p[i]);
    }

    __global__ void k_clear_pending(DevScalars* s, int nseg)
    {
        const int k = blockIdx.x * kBlock + threadIdx.x;
        if (k < nseg) s[k].x_pending = 0;
    }

    // Algorithmic bytes of one PCG iteration: the assembled apply,
    // nnz*(s + 2) + N*3s (an s-byte value and a 16-bit column per
    // off-diagonal entry; x, pdiag and y once per node), plus the minimal
    // fused vector schedule N*(8s + 1).
    static double bytes_apply(const cbs_hip_ctx* c, double s)
    {
        return double(c->sell.nnz) * (s + 2.0) + double(c->N) * 3.0 * s;
    }

    static double bytes_fp64_iter(const cbs_hip_ctx* c)
    {
        return bytes_apply(c, 8.0) + double(c->N) * 65.0;
    }

    static double bytes_fp32_iter(const cbs_hip_ctx* c)
    {
        return bytes_apply(c, 4.0) + double(c->N) * 33.0;
    }

    // Adds every segment's pending alpha p to x (before x is read).
    template <typename TP>
    void flush_iterate(cbs_hip_ctx* c, const TP* p0, const TP* p1)
    {
        hipLaunchKernelGGL(k_flush_iterate<TP>, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, c->nblk.map(), c->sv_flag,
                           p0, p1, c->sv_x, c->scal);
        hipLaunchKernelGGL(k_clear_pending, dim3(grid_for(c->nseg)), dim3(kBlock), 0, c->stream, c->scal, c->nseg);
    }

    int solve_pressure_fp64(cbs_hip_ctx* c, cbs_cg_result* res)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const int N = c->N;
        const int pre = c->cfg.cg_preconditioner;
        phase_reset(c);
        reset_cg_scalars(c);
        upload_cg_scalars(c);
        timed_group(c, PH_CONSTRAINT, [&] { launch_constraints(c); });
        sol_gather(c, c->pres, c->sv_x);
        sol_gather(c, c->rhs1, c->sv_b);
        sol_gather(c, c->pdiag, c->sv_pdiag);
        sell_refresh_values(c);
        sell_autotune64(c, c->sv_x, c->cg_ap);
        const BlockMap nb = c->nblk.map();
        const int32_t* gid = c->sell.gid;
        {
            // on scratch: cg_p, cg_p2 and cg_ap are set before they are read
            const int* never = &c->scal->pad0;
            tune_fusion(
                c, 0,
                [&] {
                    sell_apply64(c, c->cg_p, c->cg_ap, true, never);
                    hipLaunchKernelGGL(k_cg_direction, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       c->sv_pdiag, c->cg_r, c->cg_p, c->cg_p2, pre, c->scal);
                },
                [&] { sell_apply64_dir(c, c->cg_p, c->cg_p2, c->cg_r, pre, c->cg_ap, never, 1); });
        }
        const bool fuse = c->sell.cfg[0].fuse == 1;
        launch_fill(c, c->cg_ap, N, 0.0);
        launch_apply_fast(c, c->sv_x, c->cg_ap);
        hipLaunchKernelGGL(k_cg_init, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag, c->sv_b,
                           c->sv_pdiag, c->sv_x, c->cg_ap, c->cg_r, c->cg_p, pre, c->partial, c->scal, gid);
        hipLaunchKernelGGL(k_cg_init_fin, dim3(c->nseg), dim3(kBlock), 0, c->stream, c->partial, c->nblk.seg_ptr,
                           c->scal);
        download_scalars(c);
        const double setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const int* done0 = seg_flag(c, &DevScalars::done);
        const int* iter0 = seg_flag(c, &DevScalars::iter);
        int chunk = 4;
        int64_t seq = 0;
        while (!all_done(c, &DevScalars::done) && c->h_scal[0].err == ~0ull)
        {
            const int iter_before = max_iter_of(c);
            launch_iterations(
                c, 0, CBS_PCG_FP64, chunk, c->nseg == 1, seq,
                // fused: the apply forms the iteration's direction while it
                // stages it; otherwise k_cg_direction after the update
                [&] {
                    if (fuse) sell_apply64_dir(c, c->cg_p, c->cg_p2, c->cg_r, pre, c->cg_ap, done0);
                    else sell_apply64(c, c->cg_p, c->cg_ap, true, done0, c->cg_p2, iter0);
                },
                [&](const PhaseMark& mark) {
                    hipLaunchKernelGGL(k_cg_update, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       c->sv_pdiag, c->cg_p, c->cg_p2, c->cg_ap, c->sv_x, c->cg_r, pre, c->partial,
                                       c->scal, gid, c->dotpart, c->sell.seg_tile_ptr);
                    mark(PH_VECTOR);
                    hipLaunchKernelGGL(k_cg_update_fin, dim3(c->nseg), dim3(kBlock), 0, c->stream, c->partial,
                                       c->nblk.seg_ptr, c->scal);
                    mark(PH_DOT);
                    if (!fuse)
                    {
                        hipLaunchKernelGGL(k_cg_direction, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb,
                                           c->sv_flag, c->sv_pdiag, c->cg_r, c->cg_p, c->cg_p2, pre, c->scal);
                        mark(PH_VECTOR);
                    }
                });
            hip_check(hipGetLastError(), "CG launch");
            download_scalars(c, iter_before);
            if (all_done(c, &DevScalars::done) || c->h_scal[0].err != ~0ull) break;
            // Size the next chunk from the slowest segment's convergence rate.
            int next = 1;
            int remaining = 1;
            for (int k = 0; k < c->nseg; ++k)
            {
                const DevScalars& h = c->h_scal[k];
                if (h.done) continue;
                remaining = std::max(remaining, c->cfg.cg_max_iter - h.iter);
                int nk = 64;
                if (h.iter > 0 && h.rel_l2 > 0.0 && h.rel_l2 < 1.0 && c->cfg.cg_conv_test == 2)
                {
                    const double rate = std::log(h.rel_l2) / h.iter;
                    const double need = std::log(c->cfg.relToler / h.rel_l2) / rate;
                    nk = int(std::ceil(need * 0.9));
                }
                next = std::max(next, nk);
            }
            chunk = std::max(1, std::min(next, std::min(256, remaining)));
        }
        flush_iterate(c, c->cg_p, c->cg_p2);
        sol_scatter(c, c->sv_x, c->pres);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        collect_results(c, res, secs, bytes_fp64_iter(c));
        phase_fill(c, res, setup_s);
        if (c->nseg == 1) c->seg_result[0] = *res;
        return check_device_error(c);
    }

    // ------------------------------------------------------------------
    // Mixed precision with reliable updates (Sleijpen and van der Vorst 1996:
    // the residual is refreshed in high precision at chosen points).  The
    // iterate x is FP64; the recursive residual r, the direction p, Ap and
    // the operator coefficients are FP32.  Each update computes in FP64 and
    // rounds once; every reduction accumulates in FP64.  Whenever the
    // recursive residual has fallen by a decade since the last true residual
    // (or meets the convergence test), the host replaces r by the FP64 true
    // residual b - A x, keeping p (no Krylov restart).  Convergence is only
    // ever declared on a true residual.  (An FP32 iterate correction between
    // those points does not work here: its rounding, relative to the
    // correction and amplified by the condition number, is as large as the
    // decade the residual is meant to drop.)
    // ------------------------------------------------------------------
    constexpr double kReliableDrop = 0.1;

    // r = free ? b - (acc + pdiag x) : 0 in FP64 with FP64 apply output acc;
    // the true norms from it; the FP32 recursive residual restarts from its
    // rounding rf, with (rf, z), z = rf*dinv.  start: also p = z (FP32).
    // Only segments being refreshed (f_refresh) are touched.
    __global__ __launch_bounds__(kBlock) void k_rl_true(BlockMap bm, const uint8_t* __restrict__ nflag,
                                                         const double* __restrict__ rhs1,
                                                         const double* __restrict__ pdiag,
                                                         const double* __restrict__ x, const double* __restrict__ acc,
                                                         float* __restrict__ rf, float* __restrict__ pf,
                                                         const float* __restrict__ dinv, int start,
                                                         double* __restrict__ part, const DevScalars* s)
    {
        if (!s[bm.seg[blockIdx.x]].f_refresh) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
        {
            float rfi = 0.0f;
            double zi = 0.0;
            if (is_free(nflag[i]))
            {
                const double ri = rhs1[i] - (acc[i] + pdiag[i] * x[i]);
                rfi = float(ri);
                zi = double(rfi) * double(dinv[i]);
                v[0] += ri * ri;
                v[1] += double(rfi) * zi;
                const double a = fabs(ri);
                m = (m < a) ? a : m;
            }
            rf[i] = rfi;
            if (start) pf[i] = float(zi);
        }
        block_sum<2>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x == 0)
        {
            part[blockIdx.x * 4 + 0] = v[0];
            part[blockIdx.x * 4 + 1] = v[1];
            part[blockIdx.x * 4 + 2] = m;
        }
    }

    // start: the scalars of a new solve; otherwise beta = (r,z)/rz_old for
    // the direction update along the refreshed residual.
    __global__ __launch_bounds__(kBlock) void k_rl_true_fin(const double* __restrict__ part,
                                                             const int32_t* __restrict__ seg_ptr, int start,
                                                             DevScalars* sbase)
    {
        if (!sbase[blockIdx.x].f_refresh) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int b0 = seg_ptr[blockIdx.x], nb = seg_ptr[blockIdx.x + 1] - b0;
        const double* pp = part + size_t(b0) * 4;
        double rr, rz, mx;
        reduce3_partials(pp, nb, lds, ldm, rr, rz, mx);
        if (threadIdx.x != 0) return;
        DevScalars* s = sbase + blockIdx.x;
        s->true_l2 = sqrt(rr);
        s->true_max = mx;
        if (start)
        {
            s->initial_l2 = s->true_l2;
            s->rz_old = rz;
        }
        else
        {
            s->beta = rz / s->rz_old;
            s->rz_old = rz;
        }
        s->f_initial = s->true_l2;  // reference of the next drop by f_tol
        s->f_done = 0;
    }

    // alpha, then x += alpha p (FP64), r -= alpha Ap (FP64 arithmetic, FP32
    // storage); ||r||2, max|r|, (r, z) with z = r*dinv.
    // As k_cg_update: the FP64 iterate is updated every other iteration with
    // the same additions in the same order.
    __global__ __launch_bounds__(kBlock) void k_rl_update(BlockMap bm, const uint8_t* __restrict__ nflag,
                                                           const float* __restrict__ dinv,
                                                           const float* __restrict__ p0, const float* __restrict__ p1,
                                                           const float* __restrict__ ap, double* __restrict__ x,
                                                           float* __restrict__ r, double* __restrict__ part,
                                                           DevScalars* s, const double* __restrict__ dotpart,
                                                           const int32_t* __restrict__ seg_tile)
    {
        const int blk = xcd_block(int(gridDim.x));  // logical block (same XCD as the apply tiles of its rows)
        const DevScalars* ss = s + bm.seg[blk];
        if (ss->f_done) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const bool odd = ss->iter & 1, pending = ss->x_pending;
        const double alpha_prev = ss->alpha_prev;
        const float* __restrict__ p = odd ? p1 : p0;
        const float* __restrict__ pp = odd ? p0 : p1;
        // loads issued before the free flag is tested, and the first kPre
        // nodes' before block_alpha, as in k_cg_update
        struct Vals
        {
            uint8_t f;
            float r0, api, di, pi, ppi;
            double x0;
        };
        auto load = [&](int i) {
            Vals n{nflag[i], r[i], ap[i], dinv[i], 0.0f, 0.0f, 0.0};
            if (pending)
            {
                n.x0 = x[i];
                n.pi = p[i];
                n.ppi = pp[i];
            }
            return n;
        };
        constexpr int kPre = 2;
        const int i0 = bm.beg[blk] + int(threadIdx.x), iend = bm.end[blk];
        Vals pre[kPre];
#pragma unroll
        for (int k = 0; k < kPre; ++k)
            if (i0 + k * kBlock < iend) pre[k] = load(i0 + k * kBlock);
        double alpha;
        if (!block_alpha<true>(bm, dotpart, seg_tile, s, &alpha, lds)) return;
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        auto step = [&](int i, const Vals& n) {
            if (!is_free(n.f)) return;
            if (pending) x[i] = (n.x0 + alpha_prev * double(n.ppi)) + alpha * double(n.pi);
            const float rfi = float(double(n.r0) - alpha * double(n.api));
            r[i] = rfi;
            const double ri = double(rfi);
            const double zi = ri * double(n.di);
            v[0] += ri * ri;
            v[1] += ri * zi;
            const double a = fabs(ri);
            m = (m < a) ? a : m;
        };
#pragma unroll
        for (int k = 0; k < kPre; ++k)
            if (i0 + k * kBlock < iend) step(i0 + k * kBlock, pre[k]);
#pragma unroll 2
        for (int i = i0 + kPre * kBlock; i < iend; i += kBlock) step(i, load(i));
        block_sum<2>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x == 0)
        {
            part[blk * 4 + 0] = v[0];
            part[blk * 4 + 1] = v[1];
            part[blk * 4 + 2] = m;
        }
    }

    __global__ __launch_bounds__(kBlock) void k_rl_update_fin(const double* __restrict__ part,
                                                               const int32_t* __restrict__ seg_ptr,
                                                               DevScalars* sbase)
    {
        DevScalars* s = sbase + blockIdx.x;
        if (s->f_done) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int b0 = seg_ptr[blockIdx.x], nb = seg_ptr[blockIdx.x + 1] - b0;
        const double* pp = part + size_t(b0) * 4;
        double rr, rz, mx;
        reduce3_partials(pp, nb, lds, ldm, rr, rz, mx);
        if (threadIdx.x != 0) return;
        s->iter += 1;
        defer_iterate_update(s);
        const double l2 = sqrt(rr);
        s->final_l2 = l2;
        s->max_abs = mx;
        s->rel_l2 = l2 / s->initial_l2;
        const bool c1 = mx <= s->abs_tol, c2 = s->rel_l2 <= s->rel_tol;
        const bool conv = (s->conv_test == 1) ? c1 : (s->conv_test == 2) ? c2 : (c1 || c2);
        if (conv || l2 <= s->f_tol * s->f_initial || s->iter >= s->max_iter || rz <= 1.0e-300 ||
            !isfinite(rz))
        {
            s->f_done = 1;  // pause: the host replaces r by the true residual
            return;
        }
        s->beta = rz / s->rz_old;
        s->rz_old = rz;
    }

    // the direction of the next iteration (iter) from that of the one just
    // completed (iter - 1), as k_cg_direction, when the apply does not form
    // it; only_refresh: only for the segments whose residual was just
    // refreshed (continuing along p)
    __global__ __launch_bounds__(kBlock) void k_rl_direction(BlockMap bm, const uint8_t* __restrict__ nflag,
                                                              const float* __restrict__ dinv,
                                                              const float* __restrict__ r, float* __restrict__ p0,
                                                              float* __restrict__ p1, const DevScalars* s,
                                                              int only_refresh)
    {
        const int blk = xcd_block(int(gridDim.x));  // logical block (same XCD as the apply tiles of its rows)
        const DevScalars* ss = s + bm.seg[blk];
        if (ss->f_done || (only_refresh && !ss->f_refresh)) return;
        const float beta = float(ss->beta);
        const bool odd = ss->iter & 1;
        const float* __restrict__ p = odd ? p0 : p1;
        float* __restrict__ pn = odd ? p1 : p0;
#pragma unroll 2
        for (int i = bm.beg[blk] + int(threadIdx.x); i < bm.end[blk]; i += kBlock)
        {
            const uint8_t f = nflag[i];
            const float ri = r[i], di = dinv[i], po = p[i];
            float pi = 0.0f;
            if (is_free(f)) pi = float(double(ri) * double(di)) + beta * po;
            pn[i] = pi;
        }
    }

    // FP32 inverse Jacobi diagonal of the mixed form.
    __global__ void k_mx_prepare(const uint8_t* __restrict__ nflag, const double* __restrict__ pdiag,
                                 float* __restrict__ dinv, int n, int precond, DevScalars* s,
                                 const int32_t* __restrict__ gid)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        const double d = pdiag[i];
        float di = 0.0f;
        if (is_free(nflag[i]))
        {
            if (precond == 1 && (!isfinite(d) || d <= 1.0e-30))
                raise_error(s, err_key(ST_CG, 1, uint32_t(gid[i] + 1), CK_BAD_DIAG));
            di = (precond == 1) ? float(1.0 / d) : 1.0f;
        }
        dinv[i] = di;
    }

    // Every segment (batch member) runs its own reliable-update cycle: a
    // segment pauses when its recursive residual has dropped by f_tol, and
    // the paused segments are refreshed together with their FP64 true
    // residuals while the others are left as they are; a finished segment
    // stays paused.  With one segment this is the single-system solve.
    int solve_pressure_mixed(cbs_hip_ctx* c, const cbs_cg_options* o, cbs_cg_result* res)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const int N = c->N, S = c->nseg;
        const BlockMap nb = c->nblk.map();
        const double drop = (o && o->inner_rel_tol > 0.0 && o->inner_rel_tol < 1.0) ? o->inner_rel_tol
                                                                                    : kReliableDrop;
        phase_reset(c);
        reset_cg_scalars(c);
        for (int k = 0; k < S; ++k)
        {
            c->h_scal[k].f_tol = drop;
            c->h_scal[k].f_refresh = 1;
        }
        upload_cg_scalars(c);
        timed_group(c, PH_CONSTRAINT, [&] { launch_constraints(c); });
        sol_gather(c, c->pres, c->sv_x);
        sol_gather(c, c->rhs1, c->sv_b);
        sol_gather(c, c->pdiag, c->sv_pdiag);
        float* pdiagf = c->f_z;  // row sums of the FP32 apply's difference form
        hipLaunchKernelGGL(k_mx_prepare, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->sv_flag, c->sv_pdiag,
                           c->f_dinv, N, c->cfg.cg_preconditioner, c->scal, c->sell.gid);
        sell_refresh_values(c);
        sell_rowsum32(c, c->cg_p, c->cg_ap, pdiagf);
        sell_autotune64(c, c->sv_x, c->cg_ap);
        sell_autotune32(c, c->f_p, c->f_ap, pdiagf);
        {
            // on scratch: f_r, f_p, f_p2 and f_ap are set before they are read
            const int* never = &c->scal->pad0;
            tune_fusion(
                c, 1,
                [&] {
                    sell_apply32(c, c->f_p, c->f_ap, pdiagf, never);
                    hipLaunchKernelGGL(k_rl_direction, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       c->f_dinv, c->f_r, c->f_p, c->f_p2, c->scal, 0);
                },
                [&] { sell_apply32_dir(c, c->f_p, c->f_p2, c->f_r, c->f_dinv, pdiagf, c->f_ap, never, 1); });
        }
        const bool fuse = c->sell.cfg[1].fuse == 1;
        launch_fill(c, c->cg_ap, N, 0.0);
        hipLaunchKernelGGL(k_fillf_pressure, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->f_ap, int64_t(N));
        const size_t nS = static_cast<size_t>(S);
        std::vector<int> refreshes(nS, 0);
        std::vector<char> finished(nS, 0), conv(nS, 0);
        std::vector<double> initial(nS), final_l2(nS), final_rel(nS, 1.0), final_max(nS);
        const int ct = c->cfg.cg_conv_test;
        // FP64 true residuals of the segments with f_refresh set (all at start)
        auto true_residual = [&](int start) {
            timed_group(c, PH_MATVEC, [&] {
                flush_iterate(c, c->f_p, c->f_p2);
                launch_apply_fast(c, c->sv_x, c->cg_ap);
            });
            timed_group(c, PH_VECTOR, [&] {
                hipLaunchKernelGGL(k_rl_true, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag, c->sv_b,
                                   c->sv_pdiag, c->sv_x, c->cg_ap, c->f_r, c->f_p, c->f_dinv, start, c->partial,
                                   c->scal);
            });
            timed_group(c, PH_DOT, [&] {
                hipLaunchKernelGGL(k_rl_true_fin, dim3(S), dim3(kBlock), 0, c->stream, c->partial, c->nblk.seg_ptr,
                                   start, c->scal);
            });
            download_scalars(c);
        };
        auto criterion = [&](double rel, double mx) {
            const bool c1 = mx <= c->cfg.absToler, c2 = rel <= c->cfg.relToler;
            return (ct == 1) ? c1 : (ct == 2) ? c2 : (c1 || c2);
        };
        auto failed = [&] { return c->h_scal[0].err != ~0ull; };
        // finished segments stay paused; the refreshed ones that continue
        // keep f_refresh for their continuation direction
        auto settle_flags = [&](const std::vector<int>& touched) {
            for (int k : touched)
            {
                DevScalars& h = c->h_scal[k];
                h.f_refresh = finished[size_t(k)] ? 0 : 1;
                if (finished[size_t(k)]) h.f_done = 1;
            }
            upload_cg_scalars(c);
        };
        true_residual(1);
        const double setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<int> all(nS);
        for (int k = 0; k < S; ++k) all[size_t(k)] = k;
        if (!failed())
        {
            for (int k = 0; k < S; ++k)
            {
                const DevScalars& h = c->h_scal[k];
                initial[size_t(k)] = final_l2[size_t(k)] = h.true_l2;
                final_max[size_t(k)] = h.true_max;
                if (h.true_l2 <= 1.0e-300)
                {
                    conv[size_t(k)] = finished[size_t(k)] = 1;
                    final_rel[size_t(k)] = 0.0;
                }
                else if (criterion(1.0, h.true_max))
                {
                    conv[size_t(k)] = finished[size_t(k)] = 1;
                }
                else if (h.rz_old <= 1.0e-300 || !std::isfinite(h.rz_old))
                {
                    return set_error(c, CBS_ERR_NONFINITE_RESIDUAL,
                                     "ConjugateGradient::solvePressure - invalid initial preconditioned residual dot product.");
                }
            }
            settle_flags(all);
        }
        auto running = [&] {
            for (int k = 0; k < S; ++k)
                if (!finished[size_t(k)]) return true;
            return false;
        };
        int chunk = 16;
        int64_t seq = 0;
        while (running() && !failed())
        {
            const int iter_before = max_iter_of(c);
            launch_iterations(
                c, 1, CBS_PCG_MIXED, chunk, S == 1, seq,
                // fused: the apply forms the iteration's direction while it
                // stages it, also right after a refresh; otherwise
                // k_rl_direction after the update
                [&] {
                    if (fuse)
                        sell_apply32_dir(c, c->f_p, c->f_p2, c->f_r, c->f_dinv, pdiagf, c->f_ap,
                                         seg_flag(c, &DevScalars::f_done));
                    else
                        sell_apply32(c, c->f_p, c->f_ap, pdiagf, seg_flag(c, &DevScalars::f_done), c->f_p2,
                                     seg_flag(c, &DevScalars::iter));
                },
                [&](const PhaseMark& mark) {
                    hipLaunchKernelGGL(k_rl_update, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       c->f_dinv, c->f_p, c->f_p2, c->f_ap, c->sv_x, c->f_r, c->partial, c->scal,
                                       c->dotpart, c->sell.seg_tile_ptr);
                    mark(PH_VECTOR);
                    hipLaunchKernelGGL(k_rl_update_fin, dim3(S), dim3(kBlock), 0, c->stream, c->partial,
                                       c->nblk.seg_ptr, c->scal);
                    mark(PH_DOT);
                    if (!fuse)
                    {
                        hipLaunchKernelGGL(k_rl_direction, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb,
                                           c->sv_flag, c->f_dinv, c->f_r, c->f_p, c->f_p2, c->scal, 0);
                        mark(PH_VECTOR);
                    }
                });
            hip_check(hipGetLastError(), "mixed CG launch");
            download_scalars(c, iter_before);
            if (failed()) break;
            std::vector<int> paused;
            for (int k = 0; k < S; ++k)
                if (!finished[size_t(k)] && c->h_scal[k].f_done) paused.push_back(k);
            if (paused.empty())
            {
                // size the next chunk from the running segments' rates of
                // the recursive residual: 0.45 of the predicted iterations
                // to the next pause.  A chunk that overruns the pause runs
                // its remaining graph blocks as early-exit launches, which
                // cost more than the extra host round trip of a short chunk
                // (measured: 0.9 of the prediction was 1-6% slower per step
                // than 0.45 on cavity20, channel128 and cavity160).
                int next = 1;
                for (int k = 0; k < S; ++k)
                {
                    const DevScalars& h = c->h_scal[k];
                    if (finished[size_t(k)]) continue;
                    int nk = 64;
                    if (h.iter > 0 && h.final_l2 > 0.0 && h.final_l2 < h.f_initial)
                    {
                        const double rate = std::log(h.final_l2 / initial[size_t(k)]) / h.iter;
                        if (rate < 0.0) nk = int(std::ceil(0.45 * std::log(drop * h.f_initial / h.final_l2) / rate));
                    }
                    next = std::max(next, nk);
                }
                chunk = std::max(1, std::min(next, 256));
                continue;
            }
            // the paused segments' recursive residuals replaced by their
            // true residuals
            for (int k = 0; k < S; ++k) c->h_scal[k].f_refresh = 0;
            for (int k : paused)
            {
                c->h_scal[k].f_refresh = 1;
                ++refreshes[size_t(k)];
            }
            upload_cg_scalars(c);
            true_residual(0);
            if (failed()) break;
            bool cont = false;
            for (int k : paused)
            {
                const DevScalars& h = c->h_scal[k];
                final_l2[size_t(k)] = h.true_l2;
                final_max[size_t(k)] = h.true_max;
                final_rel[size_t(k)] = h.true_l2 / initial[size_t(k)];
                if (criterion(final_rel[size_t(k)], final_max[size_t(k)]))
                    conv[size_t(k)] = finished[size_t(k)] = 1;
                else if (h.iter >= c->cfg.cg_max_iter)
                    finished[size_t(k)] = 1;
                else
                    cont = true;
            }
            settle_flags(paused);
            // the segments that continue do so along p with the refreshed
            // residual, p = z + beta p with the beta of k_rl_true_fin: formed
            // by their next apply when fused, here otherwise
            if (cont && !fuse)
                timed_group(c, PH_VECTOR, [&] {
                    hipLaunchKernelGGL(k_rl_direction, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       c->f_dinv, c->f_r, c->f_p, c->f_p2, c->scal, 1);
                });
            chunk = 8;
        }
        flush_iterate(c, c->f_p, c->f_p2);
        sol_scatter(c, c->sv_x, c->pres);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        c->seg_result.assign(nS, cbs_cg_result{});
        int worst = 0;
        bool all_conv = true;
        for (int k = 0; k < S; ++k)
        {
            const DevScalars& h = c->h_scal[k];
            cbs_cg_result& r = c->seg_result[size_t(k)];
            const int mv = refreshes[size_t(k)] + 1;  // FP64 applies of its true residuals
            r.iterations = h.iter;
            r.converged = conv[size_t(k)];
            r.initial_l2 = initial[size_t(k)];
            r.final_l2 = final_l2[size_t(k)];
            r.final_relative_l2 = initial[size_t(k)] > 0.0
                                      ? (conv[size_t(k)] && initial[size_t(k)] <= 1.0e-300 ? 0.0 : final_rel[size_t(k)])
                                      : 0.0;
            r.final_max_abs = final_max[size_t(k)];
            r.matvec_calls = mv + h.iter;
            r.dot_calls = 3 * h.iter + 2 * mv;
            r.preconditioner_calls = h.iter + mv;
            r.refinements = refreshes[size_t(k)];
            r.algorithmic_bytes = bytes_fp32_iter(c) * h.iter + bytes_fp64_iter(c) * mv;
            r.total_seconds = secs;
            all_conv = all_conv && conv[size_t(k)];
            if (h.iter > c->h_scal[worst].iter) worst = k;
        }
        *res = c->seg_result[size_t(worst)];
        res->converged = all_conv ? 1 : 0;
        if (S > 1)
        {
            double bytes = 0.0;
            for (const auto& r : c->seg_result) bytes += r.algorithmic_bytes;
            res->algorithmic_bytes = bytes;
        }
        phase_fill(c, res, setup_s);
        if (S == 1) c->seg_result[0] = *res;
        return check_device_error(c);
    }
}
