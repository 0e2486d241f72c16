// Multi-right-hand-side Jacobi-PCG on the context's pressure operator.
//
// R independent systems A x_j = b_j (the same A: the assembled operator,
// pdiag and constraints of the context) are solved together, up to 32 per
// pass.  Vectors hold the R columns (R a power of two) in blocks of 8 nodes
// (mcol_at in device_common.hpp) and in the assembled operator's solver
// order, so the product runs on its row tiles (sell_spmm64/32 in
// k_sell.hip): each workgroup stages G columns of its tile's nodes in LDS
// and streams the tile's column pairs and values once for those G columns,
// and the groups of a tile share the stream through L2.  The operator's
// bytes per iteration are shared by the R systems; the vector bytes grow
// with R.
//
// Each column runs the reference's Jacobi-PCG recurrence (ConjugateGradient.cpp:
// 584-869) with its own alpha, beta and convergence test; per-column dot
// products are FP64 sums in a fixed order.  The FP32 form keeps vectors and
// operator values in FP32 with FP64 reductions inside FP64 iterative
// refinement, column by column, exactly like the single-system mixed form.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "device_common.hpp"

namespace cbsdev
{
    constexpr int kMaxRhs = kMultiMaxCols;  // columns of one pass
    // decade of the mixed form's true-residual refreshes (as the
    // single-system mixed solve)
    constexpr double kMultiReliableDrop = 0.1;

    struct MultiScalars
    {
        double rz_old[kMaxRhs], alpha[kMaxRhs], beta[kMaxRhs];
        double initial[kMaxRhs], final_l2[kMaxRhs], rel[kMaxRhs], max_abs[kMaxRhs];
        double tol_scale[kMaxRhs];  // mixed form: the last true residual norm
        int iters[kMaxRhs], done[kMaxRhs], converged[kMaxRhs];
        int refresh[kMaxRhs];  // mixed form: the columns of a true-residual refresh
        int all_done, max_iter, conv_test, inner;  // inner: the mixed form's iterations
        double rel_tol, abs_tol;
        double drop;  // mixed form: recursive residual drop between true-residual refreshes
    };

    // Per-column block sums over the threads of a grid-stride loop on
    // multi-column vectors whose threads take 8/TPR consecutive elements
    // each: a run of 8 elements (8 nodes of one column) spans TPR threads,
    // so thread t holds column (t/TPR) % R and column c is held by threads
    // TPR(c + kR) + j (j < TPR), summed in that order.  On return thread
    // c < R holds column c's sums.
    template <int NV, int TPR = 8>
    __device__ inline void column_block_sum(double (&v)[NV], int rshift, double* lds /* NV*kBlock */)
    {
        for (int q = 0; q < NV; ++q) lds[q * kBlock + threadIdx.x] = v[q];
        __syncthreads();
        const int R = 1 << rshift;
        const int c = int(threadIdx.x);
        if (c < R)
        {
            for (int q = 0; q < NV; ++q)
            {
                double s = 0.0;
                for (int k = c; k < kBlock / TPR; k += R)
                    for (int j = 0; j < TPR; ++j) s += lds[q * kBlock + TPR * k + j];
                v[q] = s;
            }
        }
        __syncthreads();
    }

    template <int TPR = 8>
    __device__ inline void column_block_max(double& m, int rshift, double* lds)
    {
        lds[threadIdx.x] = m;
        __syncthreads();
        const int R = 1 << rshift;
        const int c = int(threadIdx.x);
        if (c < R)
        {
            double s = 0.0;
            for (int k = c; k < kBlock / TPR; k += R)
                for (int j = 0; j < TPR; ++j) s = (s < lds[TPR * k + j]) ? lds[TPR * k + j] : s;
            m = s;
        }
        __syncthreads();
    }

    // Elements per thread of the vectorised update kernels: 16 bytes of TV,
    // consecutive nodes of one column.
    template <typename TV>
    struct MVec
    {
        static constexpr int kN = int(16 / sizeof(TV));
        static constexpr int kTpr = 8 / kN;  // threads per run of 8 nodes
        using V = TV __attribute__((ext_vector_type(kN)));
        using D = double __attribute__((ext_vector_type(kN)));
    };

    __device__ inline double sum_column_partials(const double* part, int nblocks, int stride, int comp)
    {
        double s = 0.0;
        for (int b = 0; b < nblocks; ++b) s += part[size_t(b) * stride + comp];
        return s;
    }

    // Start of a solve (or of a refinement cycle): r = free ? b - (acc + pdiag*x) : 0,
    // z = r/pdiag, p = z, (r,r), (r,z), max|r|; clears acc.  For the FP32
    // inner solve, r and p are written in FP32 and x_out (the correction) is
    // cleared.
    template <typename TV>
    __global__ __launch_bounds__(kBlock) void k_minit(const uint8_t* __restrict__ nflag,
                                                      const double* __restrict__ b, const double* __restrict__ pdiag,
                                                      const double* __restrict__ x, double* __restrict__ acc,
                                                      TV* __restrict__ r, TV* __restrict__ p,
                                                      const TV* __restrict__ dinv, double* __restrict__ dx, int n,
                                                      int rshift, int precond, double* __restrict__ part,
                                                      MultiScalars* s)
    {
        __shared__ double lds[3 * kBlock];
        const int R = 1 << rshift;
        const int col = mcol_col(threadIdx.x, rshift);
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        const size_t total = mcol_size(n, rshift);
        for (size_t idx = size_t(blockIdx.x) * kBlock + threadIdx.x; idx < total; idx += size_t(gridDim.x) * kBlock)
        {
            const int i = mcol_node(idx, rshift);
            TV ri = TV(0), pi = TV(0);
            if (i < n && (nflag[i] & 1u) && !(nflag[i] & 4u) && !s->done[col])
            {
                const double rr = b[idx] - (acc[idx] + pdiag[i] * x[idx]);
                ri = TV(rr);
                const TV zi = dinv ? TV(ri * dinv[i]) : TV((precond == 1) ? rr / pdiag[i] : rr);
                pi = zi;
                v[0] += rr * rr;
                v[1] += double(ri) * double(zi);
                const double a = fabs(rr);
                m = (m < a) ? a : m;
            }
            r[idx] = ri;
            p[idx] = pi;
            acc[idx] = 0.0;
            if (dx) dx[idx] = 0.0;
        }
        column_block_sum<2>(v, rshift, lds);
        column_block_max<>(m, rshift, lds + 2 * kBlock);
        const int c = int(threadIdx.x);
        if (c < R)
        {
            part[size_t(blockIdx.x) * 3 * R + c] = v[0];
            part[size_t(blockIdx.x) * 3 * R + R + c] = v[1];
            part[size_t(blockIdx.x) * 3 * R + 2 * R + c] = m;
        }
    }

    __global__ __launch_bounds__(kBlock) void k_minit_fin(const double* __restrict__ part, int nb, int rshift,
                                                           MultiScalars* s)
    {
        const int R = 1 << rshift;
        const int col = threadIdx.x;
        if (int(threadIdx.x) < R)
        {
            const int c = col;
            const double rr = sum_column_partials(part, nb, 3 * R, c);
            const double rz = sum_column_partials(part, nb, 3 * R, R + c);
            double mx = 0.0;
            for (unsigned bb = 0; bb < unsigned(nb); ++bb)
            {
                const double w = part[size_t(bb) * 3 * R + 2 * R + c];
                mx = (mx < w) ? w : mx;
            }
            if (!s->done[c])
            {
                const double l2 = sqrt(rr);
                s->final_l2[c] = l2;
                s->max_abs[c] = mx;
                if (s->inner)
                {
                    s->tol_scale[c] = l2;
                    s->rel[c] = 1.0;
                }
                s->rz_old[c] = rz;
                if (rz <= 1.0e-300 || !isfinite(rz)) s->done[c] = 1;
            }
        }
    }

    // True residual of the mixed form's columns in the refresh mask: r =
    // b - (acc + pdiag x) in FP64, stored in FP32, z = r dinv; per-column
    // (r,r), (r,z), max|r|; start: p = z.  The other columns are not touched
    // (at the start they are zeroed).  Clears acc.
    __global__ __launch_bounds__(kBlock) void k_mtrue(const uint8_t* __restrict__ nflag,
                                                      const double* __restrict__ b, const double* __restrict__ pdiag,
                                                      const double* __restrict__ x, double* __restrict__ acc,
                                                      float* __restrict__ rf, float* __restrict__ pf,
                                                      const float* __restrict__ dinv, int n, int rshift, int start,
                                                      double* __restrict__ part, const MultiScalars* s)
    {
        __shared__ double lds[3 * kBlock];
        const int R = 1 << rshift;
        const bool active = s->refresh[mcol_col(threadIdx.x, rshift)] != 0;
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        const size_t total = mcol_size(n, rshift);
        for (size_t idx = size_t(blockIdx.x) * kBlock + threadIdx.x; idx < total; idx += size_t(gridDim.x) * kBlock)
        {
            const int i = mcol_node(idx, rshift);
            if (active)
            {
                float ri = 0.0f;
                double zi = 0.0;
                if (i < n && is_free(nflag[i]))
                {
                    const double rr = b[idx] - (acc[idx] + pdiag[i] * x[idx]);
                    ri = float(rr);
                    zi = double(ri) * double(dinv[i]);
                    v[0] += rr * rr;
                    v[1] += double(ri) * zi;
                    const double a = fabs(rr);
                    m = (m < a) ? a : m;
                }
                rf[idx] = ri;
                if (start) pf[idx] = float(zi);
            }
            else if (start)
            {
                rf[idx] = 0.0f;  // columns beyond nrhs: finite values for the product
                pf[idx] = 0.0f;
            }
            acc[idx] = 0.0;
        }
        column_block_sum<2>(v, rshift, lds);
        column_block_max<>(m, rshift, lds + 2 * kBlock);
        const int c = int(threadIdx.x);
        if (c < R)
        {
            part[size_t(blockIdx.x) * 3 * R + c] = v[0];
            part[size_t(blockIdx.x) * 3 * R + R + c] = v[1];
            part[size_t(blockIdx.x) * 3 * R + 2 * R + c] = m;
        }
    }

    // Per column of the refresh mask: the true residual norms (final_l2, max_abs),
    // tol_scale = ||r||; start: initial = ||r||, rz_old = (r,z); otherwise
    // beta = (r,z) / rz_old for p = z + beta p along the refreshed residual.
    __global__ __launch_bounds__(kBlock) void k_mtrue_fin(const double* __restrict__ part, int nb, int rshift,
                                                           int start, MultiScalars* s)
    {
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int R = 1 << rshift;
        const int c = blockIdx.x;  // one block per column, fixed order
        if (!s->refresh[c]) return;  // block-uniform
        double v[2] = {0.0, 0.0};
        double m = 0.0;
#pragma unroll 4
        for (int b = threadIdx.x; b < nb; b += kBlock)
        {
            const double* pb = part + size_t(b) * 3 * R;
            v[0] += pb[c];
            v[1] += pb[R + c];
            const double w = pb[2 * R + c];
            m = (m < w) ? w : m;
        }
        block_sum<2>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x != 0) return;
        const double rr = v[0], rz = v[1], mx = m;
        const double l2 = sqrt(rr);
        s->final_l2[c] = l2;
        s->max_abs[c] = mx;
        s->tol_scale[c] = l2;
        if (start)
        {
            s->initial[c] = l2;
        }
        else
        {
            s->beta[c] = rz / s->rz_old[c];
        }
        s->rz_old[c] = rz;
    }

    // x += alpha p (x in FP64), r -= alpha Ap (FP64 arithmetic, stored as
    // TV), norms, (r,z), convergence per column.  Each thread takes
    // MVec<TV>::kN consecutive nodes of one column per step.
    template <typename TV>
    __global__ __launch_bounds__(kBlock) void k_mupdate(const uint8_t* __restrict__ nflag,
                                                        const double* __restrict__ pdiag64,
                                                        const TV* __restrict__ dinv, const TV* __restrict__ p,
                                                        const TV* __restrict__ ap, double* __restrict__ x,
                                                        TV* __restrict__ r, int n, int rshift, int precond,
                                                        double* __restrict__ part, MultiScalars* s)
    {
        using M = MVec<TV>;
        using V = typename M::V;
        using D = typename M::D;
        if (s->all_done) return;
        __shared__ double lds[3 * kBlock];
        const int R = 1 << rshift;
        const size_t q0 = (size_t(blockIdx.x) * kBlock + threadIdx.x) * M::kN;
        const int col = mcol_col(q0, rshift);
        const bool active = !s->done[col];
        const double alpha = s->alpha[col];
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        const size_t total = mcol_size(n, rshift);
        const size_t stride = size_t(gridDim.x) * kBlock * M::kN;
        for (size_t q = q0; active && q < total; q += stride)
        {
            const int i0 = mcol_node(q, rshift);
            // the vectors stream through once per iteration (several GB,
            // no reuse): non-temporal loads and stores
            const V pv = __builtin_nontemporal_load(reinterpret_cast<const V*>(p + q));
            const V av = __builtin_nontemporal_load(reinterpret_cast<const V*>(ap + q));
            V rv = __builtin_nontemporal_load(reinterpret_cast<const V*>(r + q));
            D xv = __builtin_nontemporal_load(reinterpret_cast<const D*>(x + q));
            bool any = false;
#pragma unroll
            for (int u = 0; u < M::kN; ++u)
            {
                const int i = i0 + u;
                if (i >= n || !is_free(nflag[i])) continue;
                any = true;
                xv[u] += alpha * double(pv[u]);
                const TV ri = TV(double(rv[u]) - alpha * double(av[u]));
                rv[u] = ri;
                const double zi = dinv ? double(ri) * double(dinv[i])
                                       : ((precond == 1) ? double(ri) / pdiag64[i] : double(ri));
                v[0] += double(ri) * double(ri);
                v[1] += double(ri) * zi;
                const double a = fabs(double(ri));
                m = (m < a) ? a : m;
            }
            if (any)
            {
                __builtin_nontemporal_store(xv, reinterpret_cast<D*>(x + q));
                __builtin_nontemporal_store(rv, reinterpret_cast<V*>(r + q));
            }
        }
        column_block_sum<2, M::kTpr>(v, rshift, lds);
        column_block_max<M::kTpr>(m, rshift, lds + 2 * kBlock);
        const int c = int(threadIdx.x);
        if (c < R)
        {
            part[size_t(blockIdx.x) * 3 * R + c] = v[0];
            part[size_t(blockIdx.x) * 3 * R + R + c] = v[1];
            part[size_t(blockIdx.x) * 3 * R + 2 * R + c] = m;
        }
    }

    // Per-column norms, convergence and beta from the update partials: one
    // block per column, fixed order.
    __global__ __launch_bounds__(kBlock) void k_mupdate_fin(const double* __restrict__ part, int nb, int rshift,
                                                             MultiScalars* s)
    {
        if (s->all_done) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int R = 1 << rshift;
        const int c = blockIdx.x;
        if (s->done[c]) return;  // block-uniform
        double v[2] = {0.0, 0.0};
        double m = 0.0;
#pragma unroll 4
        for (int b = threadIdx.x; b < nb; b += kBlock)
        {
            const double* pb = part + size_t(b) * 3 * R;
            v[0] += pb[c];
            v[1] += pb[R + c];
            const double w = pb[2 * R + c];
            m = (m < w) ? w : m;
        }
        block_sum<2>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x != 0) return;
        const double rr = v[0], rz = v[1], mx = m;
        s->iters[c] += 1;
        const double l2 = sqrt(rr);
        s->final_l2[c] = l2;
        s->max_abs[c] = mx;
        const double rel = l2 / s->tol_scale[c];
        s->rel[c] = rel;
        bool conv;
        if (s->inner)
        {
            // mixed form: the column pauses for a true-residual refresh once
            // its recursive residual meets the criterion (relative to the
            // solve's initial residual) or has dropped a decade below the
            // last true residual (tol_scale); convergence is decided on the
            // true residual
            const bool c1 = mx <= s->abs_tol, c2 = l2 / s->initial[c] <= s->rel_tol;
            conv = ((s->conv_test == 1) ? c1 : (s->conv_test == 2) ? c2 : (c1 || c2)) ||
                   l2 <= s->drop * s->tol_scale[c];
        }
        else
        {
            const bool c1 = mx <= s->abs_tol, c2 = rel <= s->rel_tol;
            conv = (s->conv_test == 1) ? c1 : (s->conv_test == 2) ? c2 : (c1 || c2);
        }
        if (conv)
        {
            s->converged[c] = s->inner ? 0 : 1;
            s->done[c] = 1;
        }
        else if (rz <= 1.0e-300 || !isfinite(rz) || s->iters[c] >= s->max_iter)
        {
            s->done[c] = 1;
        }
        else
        {
            s->beta[c] = rz / s->rz_old[c];
            s->rz_old[c] = rz;
        }
    }

    __global__ void k_mall_done(MultiScalars* s)
    {
        if (threadIdx.x != 0) return;
        int all = 1;
        for (int c = 0; c < kMaxRhs; ++c) all &= s->done[c];
        s->all_done = all;
    }

    // p = z + beta p per running (or refreshed) column; grid-stride with a stride that is a
    // multiple of 8R, so each thread keeps one column (its beta and flag).
    // The direction of a finished or paused column is left as it is: a
    // paused column continues along it after its true-residual refresh.
    template <typename TV>
    __global__ __launch_bounds__(kBlock) void k_mdirection(const uint8_t* __restrict__ nflag,
                                                           const double* __restrict__ pdiag64,
                                                           const TV* __restrict__ dinv, const TV* __restrict__ r,
                                                           TV* __restrict__ p, int n, int rshift, int precond,
                                                           int refreshed, const MultiScalars* s)
    {
        using M = MVec<TV>;
        using V = typename M::V;
        if (s->all_done) return;
        const size_t total = mcol_size(n, rshift);
        const size_t stride = size_t(gridDim.x) * kBlock * M::kN;
        const size_t q0 = (size_t(blockIdx.x) * kBlock + threadIdx.x) * M::kN;
        const int col = mcol_col(q0, rshift);
        // refreshed: the columns of the refresh mask continue along p from
        // their true residual; otherwise every running column
        const bool active = refreshed ? s->refresh[col] != 0 : !s->done[col];
        const TV beta = TV(s->beta[col]);
        if (!active) return;  // finished, or paused for a refresh: p is kept
        for (size_t q = q0; q < total; q += stride)
        {
            const int i0 = mcol_node(q, rshift);
            const V rv = __builtin_nontemporal_load(reinterpret_cast<const V*>(r + q));
            const V pv = __builtin_nontemporal_load(reinterpret_cast<const V*>(p + q));
            V out;
#pragma unroll
            for (int u = 0; u < M::kN; ++u)
            {
                const int i = i0 + u;
                TV pi = TV(0);
                if (i < n && is_free(nflag[i]))
                {
                    const TV zi = dinv ? TV(rv[u] * dinv[i])
                                       : TV((precond == 1) ? double(rv[u]) / pdiag64[i] : double(rv[u]));
                    pi = zi + beta * pv[u];
                }
                out[u] = pi;
            }
            __builtin_nontemporal_store(out, reinterpret_cast<V*>(p + q));
        }
    }

    // Column-major host layout <-> the multi-column device layout (node
    // order); the padding nodes of the last block of 8 hold zeros.
    __global__ void k_mpack(const double* __restrict__ colmajor, double* __restrict__ blocked, int n, int nrhs,
                            int rshift)
    {
        const size_t idx = size_t(blockIdx.x) * kBlock + threadIdx.x;
        if (idx >= mcol_size(n, rshift)) return;
        const int i = mcol_node(idx, rshift);
        const int col = mcol_col(idx, rshift);
        blocked[idx] = (i < n && col < nrhs) ? colmajor[size_t(col) * n + i] : 0.0;
    }

    __global__ void k_munpack(const double* __restrict__ blocked, double* __restrict__ colmajor, int n, int nrhs,
                              int rshift)
    {
        const size_t idx = size_t(blockIdx.x) * kBlock + threadIdx.x;
        if (idx >= mcol_size(n, rshift)) return;
        const int i = mcol_node(idx, rshift);
        const int col = mcol_col(idx, rshift);
        if (i < n && col < nrhs) colmajor[size_t(col) * n + i] = blocked[idx];
    }

    __global__ void k_mconstraints(const int32_t* __restrict__ node, const int32_t* __restrict__ mode,
                                   const double* __restrict__ val, int m, double* __restrict__ x,
                                   double* __restrict__ b, int rshift)
    {
        const size_t idx = size_t(blockIdx.x) * kBlock + threadIdx.x;
        if (idx >= (size_t(m) << rshift)) return;
        const int t = int(idx >> rshift);
        const size_t at = mcol_at(node[t], int(idx & ((size_t(1) << rshift) - 1)), rshift);
        const double v = (mode[t] == 0) ? 0.0 : val[t];
        x[at] = v;
        b[at] = v;
    }

    __global__ void k_f32_dinv(const uint8_t* __restrict__ nflag, const double* __restrict__ pdiag,
                               float* __restrict__ dinv, int n, int precond)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        const bool fr = (nflag[i] & 1u) && !(nflag[i] & 4u);
        dinv[i] = fr ? ((precond == 1) ? float(1.0 / pdiag[i]) : 1.0f) : 0.0f;
    }

    __global__ void k_fillf(float* __restrict__ p, int64_t n)
    {
        const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
        if (i < n) p[i] = 0.0f;
    }

    // pAp of column blockIdx.x from the row tiles' partials of the
    // multi-column product (tile order), then alpha.
    __global__ __launch_bounds__(kBlock) void k_spmm_fin(const double* __restrict__ part, int nb, int rshift,
                                                          MultiScalars* s)
    {
        if (s->all_done) return;
        __shared__ double lds[1][kBlock / 64];
        const int col = blockIdx.x;
        double v[1] = {0.0};
#pragma unroll 4
        for (int b = threadIdx.x; b < nb; b += kBlock) v[0] += part[(size_t(b) << rshift) + col];
        block_sum<1>(v, lds);
        if (threadIdx.x != 0 || s->done[col]) return;
        const double pap = v[0];
        if (fabs(pap) <= 1.0e-300 || !isfinite(pap)) s->done[col] = 1;
        else s->alpha[col] = s->rz_old[col] / pap;
    }

    struct MultiWork
    {
        size_t n = 0;       // elements of a multi-column vector (mcol_size)
        double *x = nullptr, *b = nullptr, *acc = nullptr, *r = nullptr, *p = nullptr, *dx = nullptr;
        float *rf = nullptr, *pf = nullptr, *apf = nullptr, *dinv = nullptr, *pdf = nullptr;
        double* part = nullptr;
        double* spart = nullptr;  // per-tile column partials of the product: ntiles << rshift
        MultiScalars* s = nullptr;
        MultiScalars* hs = nullptr;
        std::vector<void*> extra;  // tiled-apply buffers
    };

    static void mfree(MultiWork& w)
    {
        for (void* q : {(void*)w.x, (void*)w.b, (void*)w.acc, (void*)w.r, (void*)w.p, (void*)w.dx, (void*)w.rf,
                        (void*)w.pf, (void*)w.apf, (void*)w.dinv, (void*)w.pdf, (void*)w.part, (void*)w.spart, (void*)w.s})
            if (q) (void)hipFree(q);
        for (void* q : w.extra) (void)hipFree(q);
        if (w.hs) (void)hipHostFree(w.hs);
    }

    template <typename T>
    static void mdalloc(T*& p, size_t n)
    {
        hip_check(hipMalloc(reinterpret_cast<void**>(&p), (n ? n : 1) * sizeof(T)), "hipMalloc multi");
    }

    static void download_multi(cbs_hip_ctx* c, MultiWork& w)
    {
        hip_check(hipMemcpyAsync(w.hs, w.s, sizeof(MultiScalars), hipMemcpyDeviceToHost, c->stream), "multi scalars");
        hip_check(hipStreamSynchronize(c->stream), "multi sync");
        timer_collect(c);
    }

    // The iterations' product: column groups whose columns are all done
    // (finished, or paused for a refresh) are skipped.
    static void multi_product(cbs_hip_ctx* c, int groups, const double* x, double* y, const double*, double* part,
                              int rshift, MultiScalars* s)
    {
        sell_spmm64(c, groups, x, y, true, part, rshift, &s->all_done, s->done);
    }

    static void multi_product(cbs_hip_ctx* c, int groups, const float* x, float* y, const float* pdiagf,
                              double* part, int rshift, MultiScalars* s)
    {
        sell_spmm32(c, groups, x, y, pdiagf, part, rshift, &s->all_done, s->done);
    }

    // Runs PCG iterations until every column is done (FP64 vectors when
    // TV=double, FP32 vectors and operator values when TV=float).  Vectors,
    // flags and pdiag are in solver order.
    // finished (mixed form): the host's mask of finished columns; the loop
    // also returns, at the end of a chunk, once some unfinished column has
    // paused for a true-residual refresh.
    template <typename TV>
    static void iterate(cbs_hip_ctx* c, MultiWork& w, int rshift, int groups, TV* r, TV* p, TV* acc,
                        const TV* dinv, const TV* pd, double* xacc, int prec_tag, const int* finished = nullptr)
    {
        const int N = c->N;
        // up to kReduceBlocks blocks: k_mupdate_fin sums the partials with one block per column
        const int rg = reduce_grid(int64_t(w.n));
        const int pre = c->cfg.cg_preconditioner;
        int chunk = 8;
        int64_t seq = 0;
        for (;;)
        {
            for (int k = 0; k < chunk; ++k)
            {
                const bool timed = timer_begin(c, prec_tag, 0, seq++);
                // Ap and the (p, Ap) partials per row tile, then alpha per column
                multi_product(c, groups, p, acc, pd, w.spart, rshift, w.s);
                timer_end(c, timed);
                hipLaunchKernelGGL(k_spmm_fin, dim3(1 << rshift), dim3(kBlock), 0, c->stream, w.spart,
                                   sell_spmm_tiles(c), rshift, w.s);
                hipLaunchKernelGGL(k_mupdate<TV>, dim3(rg), dim3(kBlock), 0, c->stream, c->sv_flag, c->sv_pdiag,
                                   dinv, p, acc, xacc, r, N, rshift, pre, w.part, w.s);
                hipLaunchKernelGGL(k_mupdate_fin, dim3(1 << rshift), dim3(kBlock), 0, c->stream, w.part, rg, rshift,
                                   w.s);
                hipLaunchKernelGGL(k_mall_done, dim3(1), dim3(64), 0, c->stream, w.s);
                hipLaunchKernelGGL(k_mdirection<TV>, dim3(std::min(grid_for(int64_t(w.n) / MVec<TV>::kN), 4096)), dim3(kBlock), 0,
                                   c->stream, c->sv_flag, c->sv_pdiag, dinv, r, p, N, rshift, pre, 0, w.s);
            }
            hip_check(hipGetLastError(), "multi CG launch");
            download_multi(c, w);
            if (w.hs->all_done) break;
            if (finished)
            {
                bool paused = false;
                for (int j = 0; j < kMaxRhs; ++j) paused = paused || (!finished[j] && w.hs->done[j]);
                if (paused) break;
            }
            // a paused column waits for the end of the chunk: at most 32
            // iterations in the mixed form
            chunk = std::min(2 * chunk, finished ? 32 : 128);
        }
    }

    int solve_pressure_multi(cbs_hip_ctx* c, int nrhs, const double* B, double* X, const cbs_cg_options* o,
                             cbs_cg_result* res)
    {
        const auto t0 = std::chrono::steady_clock::now();
        int rshift = 0;
        while ((1 << rshift) < nrhs) ++rshift;
        const int R = 1 << rshift;
        const int N = c->N;
        const bool mixed = o && o->precision == CBS_PCG_MIXED;
        MultiWork w;
        w.n = mcol_size(N, rshift);
        int rc = CBS_OK;
        try
        {
            mdalloc(w.x, w.n);
            mdalloc(w.b, w.n);
            mdalloc(w.acc, w.n);
            mdalloc(w.part, size_t(kReduceBlocks) * 3 * kMaxRhs);
            mdalloc(w.s, 1);
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&w.hs), sizeof(MultiScalars), hipHostMallocDefault),
                      "hipHostMalloc");
            if (mixed)
            {
                mdalloc(w.dx, w.n);
                mdalloc(w.rf, w.n);
                mdalloc(w.pf, w.n);
                mdalloc(w.apf, w.n);
                mdalloc(w.dinv, size_t(N));
                mdalloc(w.pdf, size_t(N));
            }
            else
            {
                mdalloc(w.r, w.n);
                mdalloc(w.p, w.n);
            }
            // the assembled operator's current values and pdiag in solver order
            sell_refresh_values(c);
            sol_gather(c, c->pdiag, c->sv_pdiag);
            const int g64 = sell_spmm_groups(c, false, rshift);
            const int g32 = mixed ? sell_spmm_groups(c, true, rshift) : 1;
            if (c->sell.ntiles > 0 && (g64 == 0 || g32 == 0))
                throw std::runtime_error("multi-RHS product: the row tiles' staging exceeds the LDS");
            mdalloc(w.spart, std::max<size_t>(1, size_t(sell_spmm_tiles(c))) << rshift);
            // host column-major -> device node-major in node order, the
            // constraints of impose_pressure_constraints on the non-free
            // nodes, then solver order
            double* tmp = mixed ? w.dx : w.r;
            double* stage = w.acc;
            hip_check(hipMemcpyAsync(stage, B, sizeof(double) * size_t(nrhs) * N, hipMemcpyHostToDevice, c->stream), "B");
            hipLaunchKernelGGL(k_mpack, dim3(grid_for(int64_t(w.n))), dim3(kBlock), 0, c->stream, stage, w.x, N, nrhs,
                               rshift);
            hip_check(hipMemcpyAsync(stage, X, sizeof(double) * size_t(nrhs) * N, hipMemcpyHostToDevice, c->stream), "X");
            hipLaunchKernelGGL(k_mpack, dim3(grid_for(int64_t(w.n))), dim3(kBlock), 0, c->stream, stage, tmp, N, nrhs,
                               rshift);
            if (c->nnonfree > 0)
            {
                launch_constraints(c);
                hipLaunchKernelGGL(k_mconstraints, dim3(grid_for(int64_t(c->nnonfree) << rshift)), dim3(kBlock), 0,
                                   c->stream, c->nf_node, c->nf_mode, c->nf_val, c->nnonfree, tmp, w.x, rshift);
            }
            sell_gather_columns(c, w.x, w.b, rshift);
            sell_gather_columns(c, tmp, w.x, rshift);
            MultiScalars& h = *w.hs;
            h = MultiScalars{};
            for (int j = 0; j < kMaxRhs; ++j) h.done[j] = (j >= nrhs) ? 1 : 0;
            h.max_iter = c->cfg.cg_max_iter;
            h.conv_test = c->cfg.cg_conv_test;
            h.rel_tol = c->cfg.relToler;
            h.abs_tol = c->cfg.absToler;
            const int max_ref = (o && o->max_refinements > 0) ? o->max_refinements : 50;
            double initial[kMaxRhs] = {0};
            int converged[kMaxRhs] = {0};
            int refinements[kMaxRhs] = {0};
            int iters[kMaxRhs] = {0};
            double final_l2[kMaxRhs] = {0}, final_rel[kMaxRhs] = {0}, final_max[kMaxRhs] = {0};
            const int rg = std::min(reduce_grid(int64_t(w.n)), 1024);  // column sums are serial per column in the _fin kernels
            if (!mixed)
            {
                hip_check(hipMemcpyAsync(w.s, w.hs, sizeof(MultiScalars), hipMemcpyHostToDevice, c->stream), "scal");
                launch_fill(c, w.acc, int64_t(w.n), 0.0);
                sell_spmm64(c, g64, w.x, w.acc, false, w.spart, rshift, &w.s->all_done);
                hipLaunchKernelGGL(k_minit<double>, dim3(rg), dim3(kBlock), 0, c->stream, c->sv_flag, w.b, c->sv_pdiag, w.x,
                                   w.acc, w.r, w.p, static_cast<const double*>(nullptr), static_cast<double*>(nullptr),
                                   N, rshift, c->cfg.cg_preconditioner, w.part, w.s);
                hipLaunchKernelGGL(k_minit_fin, dim3(1), dim3(kBlock), 0, c->stream, w.part, rg, rshift, w.s);
                download_multi(c, w);
                for (int j = 0; j < nrhs; ++j)
                {
                    initial[j] = h.final_l2[j];
                    h.tol_scale[j] = initial[j];
                    h.rel[j] = 1.0;
                    const bool c1 = h.max_abs[j] <= h.abs_tol, c2 = 1.0 <= h.rel_tol;
                    if (initial[j] <= 1.0e-300 || ((h.conv_test == 1) ? c1 : (h.conv_test == 2) ? c2 : (c1 || c2)))
                    {
                        h.done[j] = 1;
                        h.converged[j] = 1;
                    }
                }
                h.all_done = 1;
                for (int j = 0; j < nrhs; ++j) h.all_done &= h.done[j];
                hip_check(hipMemcpyAsync(w.s, w.hs, sizeof(MultiScalars), hipMemcpyHostToDevice, c->stream), "scal");
                iterate<double>(c, w, rshift, g64, w.r, w.p, w.acc, static_cast<const double*>(nullptr), c->sv_pdiag,
                                w.x, CBS_PCG_FP64);
                for (int j = 0; j < nrhs; ++j)
                {
                    converged[j] = h.converged[j];
                    iters[j] = h.iters[j];
                    final_l2[j] = h.final_l2[j];
                    final_rel[j] = h.rel[j];
                    final_max[j] = h.max_abs[j];
                }
            }
            else
            {
                // Reliable-update PCG per column, as the single-system mixed
                // form: FP32 iterations on the recursive residual with the
                // iterate in FP64; once every running column has paused
                // (a decade gained, the criterion met, or the iteration
                // limit), their residuals are replaced by the true FP64
                // residuals and they continue along p (no restart).
                hipLaunchKernelGGL(k_f32_dinv, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->sv_flag, c->sv_pdiag,
                                   w.dinv, N, c->cfg.cg_preconditioner);
                sell_rowsum32(c, c->cg_p, c->cg_ap, w.pdf);  // the FP32 product's row sums
                hipLaunchKernelGGL(k_fillf, dim3(grid_for(int64_t(w.n))), dim3(kBlock), 0, c->stream, w.apf,
                                   int64_t(w.n));
                h.inner = 1;
                h.drop = (o && o->inner_rel_tol > 0.0) ? o->inner_rel_tol : kMultiReliableDrop;
                int outer_done[kMaxRhs];
                for (int j = 0; j < kMaxRhs; ++j)
                {
                    outer_done[j] = (j >= nrhs) ? 1 : 0;
                    h.done[j] = outer_done[j];
                    h.refresh[j] = 1 - outer_done[j];
                }
                h.all_done = 0;
                const int ct = c->cfg.cg_conv_test;
                hip_check(hipMemcpyAsync(w.s, w.hs, sizeof(MultiScalars), hipMemcpyHostToDevice, c->stream), "scal");
                auto upload_flags = [&] {
                    hip_check(hipMemcpyAsync(w.s->done, h.done, sizeof(h.done), hipMemcpyHostToDevice, c->stream),
                              "done");
                    hip_check(hipMemcpyAsync(w.s->refresh, h.refresh, sizeof(h.refresh), hipMemcpyHostToDevice,
                                             c->stream),
                              "refresh");
                    hip_check(hipMemcpyAsync(&w.s->all_done, &h.all_done, sizeof(int), hipMemcpyHostToDevice,
                                             c->stream),
                              "all");
                };
                for (int start = 1;; start = 0)
                {
                    // true FP64 residual of the columns in the refresh mask
                    // (at the start every unfinished column; later those that
                    // paused); the running columns are not touched
                    launch_fill(c, w.acc, int64_t(w.n), 0.0);
                    sell_spmm64(c, g64, w.x, w.acc, false, w.spart, rshift, &w.s->all_done);
                    hipLaunchKernelGGL(k_mtrue, dim3(rg), dim3(kBlock), 0, c->stream, c->sv_flag, w.b, c->sv_pdiag,
                                       w.x, w.acc, w.rf, w.pf, static_cast<const float*>(w.dinv), N, rshift, start,
                                       w.part, w.s);
                    hipLaunchKernelGGL(k_mtrue_fin, dim3(1 << rshift), dim3(kBlock), 0, c->stream, w.part, rg, rshift,
                                       start, w.s);
                    download_multi(c, w);
                    bool all = true;
                    for (int j = 0; j < kMaxRhs; ++j)
                    {
                        if (j < nrhs && h.refresh[j])
                        {
                            const double l2 = h.final_l2[j];
                            if (start) initial[j] = l2;
                            else ++refinements[j];
                            final_l2[j] = l2;
                            final_max[j] = h.max_abs[j];
                            final_rel[j] = initial[j] > 0.0 ? l2 / initial[j] : 0.0;
                            const bool c1 = final_max[j] <= c->cfg.absToler, c2 = final_rel[j] <= c->cfg.relToler;
                            if (initial[j] <= 1.0e-300 || ((ct == 1) ? c1 : (ct == 2) ? c2 : (c1 || c2)))
                            {
                                converged[j] = 1;
                                outer_done[j] = 1;
                            }
                            else if (refinements[j] >= max_ref || h.iters[j] >= c->cfg.cg_max_iter ||
                                     h.rz_old[j] <= 1.0e-300 || !std::isfinite(h.rz_old[j]))
                            {
                                outer_done[j] = 1;
                            }
                        }
                        // refreshed columns run again; finished ones stop
                        if (outer_done[j]) h.refresh[j] = 0;
                        h.done[j] = outer_done[j] ? 1 : (h.refresh[j] ? 0 : h.done[j]);
                        all = all && outer_done[j];
                    }
                    if (all) break;
                    h.all_done = 0;
                    upload_flags();
                    // p = z + beta p along the refreshed residuals
                    if (!start)
                        hipLaunchKernelGGL(k_mdirection<float>, dim3(std::min(grid_for(int64_t(w.n) / MVec<float>::kN), 4096)),
                                           dim3(kBlock), 0, c->stream, c->sv_flag, c->sv_pdiag,
                                           static_cast<const float*>(w.dinv), w.rf, w.pf, N, rshift,
                                           c->cfg.cg_preconditioner, 1, w.s);
                    iterate<float>(c, w, rshift, g32, w.rf, w.pf, w.apf, w.dinv, w.pdf, w.x, CBS_PCG_MIXED,
                                   outer_done);
                    // the next refresh: every unfinished column that paused
                    for (int j = 0; j < kMaxRhs; ++j) h.refresh[j] = (!outer_done[j] && h.done[j]) ? 1 : 0;
                    h.all_done = 0;
                    upload_flags();
                }
                for (int j = 0; j < nrhs; ++j) iters[j] = h.iters[j];
            }
            sell_scatter_columns(c, w.x, tmp, rshift);
            hipLaunchKernelGGL(k_munpack, dim3(grid_for(int64_t(w.n))), dim3(kBlock), 0, c->stream, tmp, w.acc, N, nrhs,
                               rshift);
            hip_check(hipMemcpyAsync(X, w.acc, sizeof(double) * size_t(nrhs) * N, hipMemcpyDeviceToHost, c->stream), "X");
            hip_check(hipStreamSynchronize(c->stream), "multi sync");
            const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const double s = mixed ? 4.0 : 8.0;
            for (int j = 0; j < nrhs; ++j)
            {
                cbs_cg_result& rr = res[j];
                std::memset(&rr, 0, sizeof rr);
                rr.iterations = iters[j];
                rr.refinements = refinements[j];
                rr.converged = converged[j];
                rr.initial_l2 = initial[j];
                rr.final_l2 = final_l2[j];
                rr.final_relative_l2 = final_rel[j];
                rr.final_max_abs = final_max[j];
                rr.total_seconds = secs;
                // the operator stream is shared by all columns: charge it once per sweep
                rr.algorithmic_bytes = double(iters[j]) * (double(N) * 3.0 * s + double(N) * (8.0 * s + 1.0)) +
                                       (j == 0 ? double(*std::max_element(iters, iters + nrhs)) *
                                                     double(c->sell.nnz) * (s + 2.0)
                                               : 0.0);
            }
            (void)R;
        }
        catch (...)
        {
            mfree(w);
            throw;
        }
        mfree(w);
        return rc;
    }
}
