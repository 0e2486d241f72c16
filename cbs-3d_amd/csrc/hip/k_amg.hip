// Pressure solve preconditioned by smoothed-aggregation algebraic multigrid
// (CBS_PCG_AMG): the reference's conjugate-gradient recurrence and
// convergence test (ConjugateGradient.cpp:584-869) with z = M r, M one
// V-cycle, in place of the Jacobi step z = r / pdiag.  A build-only
// preconditioner behind the solver seam (the reference dispatches its
// solvers on solver_opt, Steps.cpp:313-335); the Jacobi forms stay the
// reference's algorithm.
//
// Outer iteration (FP64, solver order): the assembled FP64 apply with its
// (p, Ap) tile partials; k_amg_update (alpha, x, r, FP32 copy of r);
// k_amg_update_fin (convergence test); the V-cycle; k_amg_dot and
// k_amg_dot_fin ((r, z), beta); k_amg_direction.
//
// V-cycle (FP32 vectors and operators): on every level one damped-Jacobi
// sweep from zero, restriction of the residual, the coarse correction, and
// one damped-Jacobi sweep after it; the coarsest level is solved with its
// dense inverse.  The fine level applies the current assembled operator
// (FP32 values) and the current pdiag; the coarser levels are the Galerkin
// products of the operator at set-up (amg_setup.hpp), scaled by theta, the
// ratio of the current to the set-up sum of pdiag over free nodes (exact
// when the time step is uniform, as the operator is then dt times a fixed
// matrix; otherwise a fixed symmetric positive definite preconditioner).
// Every sum runs in a fixed order: results are deterministic.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "amg_setup.hpp"
#include "ctx.hpp"
#include "device_common.hpp"
#include "pcg_host.hpp"

// A prolongation of the hierarchy as sliced ELLPACK: within windows of
// kAmgWindow rows the rows are sorted by length (longest first) and cut
// into slices of 64, one wavefront each, a lane per row; a slice stores its
// longest row's count of entries per lane, entry k of lane l at off[s] +
// 64k + l (padding: column 0, value 0), so every load of a product is
// coalesced and no row pointer is read.  row[64s + l] is the matrix row of
// lane l of slice s (-1 past the last row).
struct AmgSell
{
    int32_t nslices = 0;
    int64_t* off = nullptr;
    int32_t* row = nullptr;
    int32_t* col = nullptr;
    float* val = nullptr;
};

struct AmgState
{
    struct Level
    {
        int32_t n = 0;
        int32_t *a_ptr = nullptr, *a_col = nullptr;  // operator (levels >= 1)
        float* a_val = nullptr;
        float* dinv = nullptr;
        float omega = 0.0f;
        AmgSell p;                                   // prolongation from the next level
        int32_t *r_ptr = nullptr, *r_col = nullptr;  // restriction to the next level
        int a_lanes = 16, r_lanes = 16;              // lanes per row of the A and R products
        float* r_val = nullptr;
        float *b = nullptr, *x = nullptr, *t = nullptr;
    };
    std::vector<Level> lv;
    float* inv = nullptr;       // coarsest dense inverse (symmetric)
    double scale_ref = 0.0;     // sum of pdiag over free nodes at set-up
    double* scale = nullptr;    // device: current sum, then theta
    float* rowsumf = nullptr;   // row sums of the FP32 fine-level apply (sell_rowsum32; solver order)
    float* rf = nullptr;        // FP32 residual handed to the V-cycle
    float* z = nullptr;         // V-cycle output (level 0 x)
    double setup_seconds = 0.0;
    int64_t stored = 0;         // nonzeros of the coarse operators, P and R
    std::vector<void*> mem;
};

namespace cbsdev
{
    namespace
    {
        // Zero-filled on the context's stream, so the fill is ordered before
        // the uploads and kernels that follow on that stream (a fill on the
        // default stream is not: the context's stream is non-blocking).
        // Set-up only, never captured.
        template <typename T>
        T* amg_alloc(AmgState* st, hipStream_t stream, size_t n)
        {
            void* p = nullptr;
            hip_check(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc multigrid");
            st->mem.push_back(p);
            hip_check(hipMemsetAsync(p, 0, (n ? n : 1) * sizeof(T), stream), "hipMemsetAsync multigrid");
            return static_cast<T*>(p);
        }

        template <typename T, typename S>
        T* amg_upload(AmgState* st, hipStream_t stream, const std::vector<S>& h)
        {
            std::vector<T> v(h.begin(), h.end());
            T* d = amg_alloc<T>(st, stream, v.size());
            if (!v.empty())
                hip_check(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream),
                          "multigrid upload");
            hip_check(hipStreamSynchronize(stream), "multigrid upload");
            return d;
        }

        template <typename T>
        std::vector<T> download(const T* d, size_t n)
        {
            std::vector<T> h(n);
            if (n) hip_check(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "multigrid download");
            return h;
        }

    }

    // ------------------------------------------------------------------
    // V-cycle kernels (FP32).  `done`: the solve's flag; a finished solve's
    // remaining iterations return at once.
    // ------------------------------------------------------------------
    __global__ void k_amg_jacobi0(int n, const float* __restrict__ dinv, float w, const double* __restrict__ theta,
                                  const float* __restrict__ b, float* __restrict__ x, const int* __restrict__ done)
    {
        if (*done) return;
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        const float s = theta ? float(double(w) / *theta) : w;
        x[i] = s * dinv[i] * b[i];
    }

    __global__ void k_amg_jacobi(int n, const float* __restrict__ dinv, float w, const double* __restrict__ theta,
                                 const float* __restrict__ b, const float* __restrict__ t, float* __restrict__ x,
                                 const int* __restrict__ done)
    {
        if (*done) return;
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        const float s = theta ? float(double(w) / *theta) : w;
        x[i] += s * dinv[i] * (b[i] - t[i]);
    }

    template <int L>
    __device__ inline float lanes_sum(float s)
    {
#pragma unroll
        for (int m = L / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, L);
        return s;
    }

    // A lane's entries k = k0 + lane, k0 + lane + L, ... of a CSR row in
    // batches of kAmgBatch: a batch's column ids and values are loaded
    // together, then its gathers of the vector(s), then the products are
    // added in entry order (the same sum as one entry at a time, with every
    // load of a batch in flight at once).  At cavity160 batches of 4 took
    // the step from 23.9 to 22.6 ms (2: 22.9, 6: 22.6, 8: 23.3 ms).
    constexpr int kAmgBatch = 4;

    // y = theta A x, L lanes per row
    template <int L>
    __global__ void k_amg_spmv(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                               const float* __restrict__ val, const double* __restrict__ theta,
                               const float* __restrict__ x, float* __restrict__ y, const int* __restrict__ done)
    {
        if (*done) return;
        const int g = blockIdx.x * kBlock + threadIdx.x;
        const int row = g / L, lane = g % L;
        float s = 0.0f;
        if (row < n)
        {
            const int k0 = ptr[row], k1 = ptr[row + 1];
            for (int kb = k0 + lane; kb < k1; kb += kAmgBatch * L)
            {
                int c[kAmgBatch];
                float v[kAmgBatch], xv[kAmgBatch];
#pragma unroll
                for (int u = 0; u < kAmgBatch; ++u)
                {
                    const int k = kb + u * L < k1 ? kb + u * L : kb;
                    c[u] = col[k];
                    v[u] = val[k];
                }
#pragma unroll
                for (int u = 0; u < kAmgBatch; ++u) xv[u] = x[c[u]];
#pragma unroll
                for (int u = 0; u < kAmgBatch; ++u)
                    if (kb + u * L < k1) s += v[u] * xv[u];
            }
        }
        s = lanes_sum<L>(s);
        if (lane == 0 && row < n) y[row] = float(*theta) * s;
    }

    // bn = R (b - t), L lanes per row
    template <int L>
    __global__ void k_amg_restrict(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                   const float* __restrict__ val, const float* __restrict__ b,
                                   const float* __restrict__ t, float* __restrict__ bn, const int* __restrict__ done)
    {
        if (*done) return;
        const int g = blockIdx.x * kBlock + threadIdx.x;
        const int row = g / L, lane = g % L;
        float s = 0.0f;
        if (row < n)
        {
            const int k0 = ptr[row], k1 = ptr[row + 1];
            for (int kb = k0 + lane; kb < k1; kb += kAmgBatch * L)
            {
                int c[kAmgBatch];
                float v[kAmgBatch], bv[kAmgBatch], tv[kAmgBatch];
#pragma unroll
                for (int u = 0; u < kAmgBatch; ++u)
                {
                    const int k = kb + u * L < k1 ? kb + u * L : kb;
                    c[u] = col[k];
                    v[u] = val[k];
                }
#pragma unroll
                for (int u = 0; u < kAmgBatch; ++u)
                {
                    bv[u] = b[c[u]];
                    tv[u] = t[c[u]];
                }
#pragma unroll
                for (int u = 0; u < kAmgBatch; ++u)
                    if (kb + u * L < k1) s += v[u] * (bv[u] - tv[u]);
            }
        }
        s = lanes_sum<L>(s);
        if (lane == 0 && row < n) bn[row] = s;
    }

    // x += P xc on a sliced prolongation (AmgSell), a lane per row, the
    // row's entries in CSR order.  The rows of P are short (a few
    // aggregates each): as CSR with four lanes a row the fine level's
    // prolongation took 114 us of a 0.7 ms V-cycle at cavity160.  The long
    // rows of the coarse operators and of the restrictions keep the CSR
    // kernels above, several lanes a row (a lane a row made the
    // restrictions 2-5x slower).
    __global__ void k_amg_prolong(int nslices, const int64_t* __restrict__ off, const int32_t* __restrict__ rows,
                                  const int32_t* __restrict__ col, const float* __restrict__ val,
                                  const float* __restrict__ xc, float* __restrict__ x, const int* __restrict__ done)
    {
        if (*done) return;
        const int g = blockIdx.x * kBlock + threadIdx.x;
        const int s = g >> 6, l = g & 63;
        if (s >= nslices) return;  // whole wavefronts
        const int64_t o = off[s];
        const int K = int((off[s + 1] - o) >> 6);
        float acc = 0.0f;
#pragma unroll 4
        for (int k = 0; k < K; ++k)
        {
            const int64_t q = o + int64_t(k) * 64 + l;
            acc += val[q] * xc[col[q]];
        }
        const int r = rows[g];
        if (r >= 0) x[r] += acc;
    }

    // x = inv b / theta: one wavefront per row, lanes over the row (reads
    // coalesced), the wavefront butterfly for the row's sum
    __global__ void k_amg_dense(int n, const float* __restrict__ inv, const double* __restrict__ theta,
                                const float* __restrict__ b, float* __restrict__ x, const int* __restrict__ done)
    {
        if (*done) return;
        const int row = (blockIdx.x * kBlock + threadIdx.x) >> 6, lane = threadIdx.x & 63;
        float s = 0.0f;
        if (row < n)
            for (int j = lane; j < n; j += 64) s += inv[size_t(row) * n + j] * b[j];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
        if (lane == 0 && row < n) x[row] = float(double(s) / *theta);
    }

    // ------------------------------------------------------------------
    // Outer FP64 iteration
    // ------------------------------------------------------------------
    // FP32 pdiag and the smoother's inverse diagonal of the fine level; the
    // partial sums of pdiag over free nodes (for theta); the reference's
    // diagonal check (ConjugateGradient.cpp:302-369).
    __global__ __launch_bounds__(kBlock) void k_amg_prepare(BlockMap bm, const uint8_t* __restrict__ flag,
                                                             const double* __restrict__ pdiag,
                                                             float* __restrict__ dinv,
                                                             double* __restrict__ part, DevScalars* s,
                                                             const int32_t* __restrict__ gid)
    {
        __shared__ double lds[1][kBlock / 64];
        double v[1] = {0.0};
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
        {
            const double d = pdiag[i];
            float di = 0.0f;
            if (is_free(flag[i]))
            {
                if (!isfinite(d) || d <= 1.0e-30) raise_error(s, err_key(ST_CG, 1, uint32_t(gid[i] + 1), CK_BAD_DIAG));
                di = float(1.0 / d);
                v[0] += d;
            }
            dinv[i] = di;
        }
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) part[blockIdx.x * 4] = v[0];
    }

    __global__ __launch_bounds__(kBlock) void k_amg_theta(const double* __restrict__ part, int nb, double ref,
                                                           double* __restrict__ theta)
    {
        __shared__ double lds[1][kBlock / 64];
        const double sum = reduce_partials(part, nb, 4, 0, lds);
        if (threadIdx.x == 0) *theta = sum / ref;
    }

    // r = free ? b - (acc + pdiag x) : 0, rf = float(r), acc = 0; partials
    // of ||r||^2 and max|r|.
    __global__ __launch_bounds__(kBlock) void k_amg_init(BlockMap bm, const uint8_t* __restrict__ flag,
                                                          const double* __restrict__ b,
                                                          const double* __restrict__ pdiag,
                                                          const double* __restrict__ x, double* __restrict__ acc,
                                                          double* __restrict__ r, float* __restrict__ rf,
                                                          double* __restrict__ part)
    {
        __shared__ double lds[1][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        double v[1] = {0.0};
        double m = 0.0;
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
        {
            double ri = 0.0;
            if (is_free(flag[i]))
            {
                ri = b[i] - (acc[i] + pdiag[i] * x[i]);
                v[0] += ri * ri;
                const double a = fabs(ri);
                m = (m < a) ? a : m;
            }
            r[i] = ri;
            rf[i] = float(ri);
            acc[i] = 0.0;
        }
        block_sum<1>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x == 0)
        {
            part[blockIdx.x * 4 + 0] = v[0];
            part[blockIdx.x * 4 + 2] = m;
        }
    }

    __global__ __launch_bounds__(kBlock) void k_amg_init_fin(const double* __restrict__ part,
                                                              const int32_t* __restrict__ seg_ptr, DevScalars* s)
    {
        __shared__ double lds[1][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int b0 = seg_ptr[0], nb = seg_ptr[1] - b0;
        const double rr = reduce_partials(part + size_t(b0) * 4, nb, 4, 0, lds);
        const double mx = max_partials(part + size_t(b0) * 4, nb, 4, 2, ldm);
        if (threadIdx.x != 0) return;
        s->initial_l2 = sqrt(rr);
        s->final_l2 = s->initial_l2;
        s->max_abs = mx;
        s->iter = 0;
        s->rel_l2 = s->initial_l2 <= 1.0e-300 ? 0.0 : 1.0;
        const bool c1 = mx <= s->abs_tol, c2 = s->rel_l2 <= s->rel_tol;
        const bool conv = s->initial_l2 <= 1.0e-300 || ((s->conv_test == 1) ? c1 : (s->conv_test == 2) ? c2 : (c1 || c2));
        if (conv)
        {
            s->converged = 1;
            s->done = 1;
        }
    }

    // alpha from the apply's tile partials, then x += alpha p, r -= alpha Ap,
    // rf = float(r) on free nodes; partials of ||r||^2 and max|r|.
    __global__ __launch_bounds__(kBlock) void k_amg_update(BlockMap bm, const uint8_t* __restrict__ flag,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ ap, double* __restrict__ x,
                                                            double* __restrict__ r, float* __restrict__ rf,
                                                            double* __restrict__ part, DevScalars* s,
                                                            const double* __restrict__ dotpart,
                                                            const int32_t* __restrict__ seg_tile)
    {
        if (s->done) return;
        __shared__ double lds[1][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int t0 = seg_tile[0];
        const double pap = reduce_partials(dotpart + t0, seg_tile[1] - t0, 1, 0, lds);
        const bool bad = fabs(pap) <= 1.0e-300 || !isfinite(pap);
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            s->pAp = pap;
            if (bad)
            {
                raise_error(s, err_key(ST_CG, 3, uint32_t(s->iter + 1), CK_BREAKDOWN));
                s->done = 1;
            }
            else
            {
                s->alpha = s->rz_old / pap;
            }
        }
        if (bad) return;
        const double alpha = s->rz_old / pap;
        double v[1] = {0.0};
        double m = 0.0;
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
        {
            if (!is_free(flag[i])) continue;
            x[i] += alpha * p[i];
            const double ri = r[i] - alpha * ap[i];
            r[i] = ri;
            rf[i] = float(ri);
            v[0] += ri * ri;
            const double a = fabs(ri);
            m = (m < a) ? a : m;
        }
        block_sum<1>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x == 0)
        {
            part[blockIdx.x * 4 + 0] = v[0];
            part[blockIdx.x * 4 + 2] = m;
        }
    }

    // iteration count and the reference's convergence test (:796-852)
    __global__ __launch_bounds__(kBlock) void k_amg_update_fin(const double* __restrict__ part,
                                                                const int32_t* __restrict__ seg_ptr, DevScalars* s)
    {
        if (s->done) return;
        __shared__ double lds[1][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const int b0 = seg_ptr[0], nb = seg_ptr[1] - b0;
        const double rr = reduce_partials(part + size_t(b0) * 4, nb, 4, 0, lds);
        const double mx = max_partials(part + size_t(b0) * 4, nb, 4, 2, ldm);
        if (threadIdx.x != 0) return;
        s->iter += 1;
        s->final_l2 = sqrt(rr);
        s->max_abs = mx;
        s->rel_l2 = s->final_l2 / s->initial_l2;
        const bool c1 = mx <= s->abs_tol, c2 = s->rel_l2 <= s->rel_tol;
        const bool conv = (s->conv_test == 1) ? c1 : (s->conv_test == 2) ? c2 : (c1 || c2);
        if (conv)
        {
            s->converged = 1;
            s->done = 1;
        }
        else if (s->iter >= s->max_iter)
        {
            s->done = 1;
        }
    }

    // partials of (r, z)
    __global__ __launch_bounds__(kBlock) void k_amg_dot(BlockMap bm, const uint8_t* __restrict__ flag,
                                                         const double* __restrict__ r, const float* __restrict__ z,
                                                         double* __restrict__ part, const DevScalars* s)
    {
        if (s->done) return;
        __shared__ double lds[1][kBlock / 64];
        double v[1] = {0.0};
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
            if (is_free(flag[i])) v[0] += r[i] * double(z[i]);
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) part[blockIdx.x * 4 + 1] = v[0];
    }

    // beta = (r, z) / rz_old (first: 0), rz_old = (r, z)
    __global__ __launch_bounds__(kBlock) void k_amg_dot_fin(const double* __restrict__ part,
                                                             const int32_t* __restrict__ seg_ptr, int first,
                                                             DevScalars* s)
    {
        if (s->done) return;
        __shared__ double lds[1][kBlock / 64];
        const int b0 = seg_ptr[0], nb = seg_ptr[1] - b0;
        const double rz = reduce_partials(part + size_t(b0) * 4, nb, 4, 1, lds);
        if (threadIdx.x != 0) return;
        if (rz <= 1.0e-300 || !isfinite(rz))
        {
            raise_error(s, err_key(ST_CG, first ? 2 : 4, uint32_t(s->iter), first ? CK_RZ_INIT : CK_RZ_UPD));
            s->done = 1;
            return;
        }
        s->beta = first ? 0.0 : rz / s->rz_old;
        s->rz_old = rz;
    }

    // p = free ? z + beta p : 0
    __global__ __launch_bounds__(kBlock) void k_amg_direction(BlockMap bm, const uint8_t* __restrict__ flag,
                                                               const float* __restrict__ z, double* __restrict__ p,
                                                               const DevScalars* s)
    {
        if (s->done) return;
        const double beta = s->beta;
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
            p[i] = is_free(flag[i]) ? double(z[i]) + beta * p[i] : 0.0;
    }

    // ------------------------------------------------------------------
    // Host
    // ------------------------------------------------------------------
    void destroy_amg(cbs_hip_ctx* c)
    {
        if (!c->amg) return;
        for (void* p : c->amg->mem) (void)hipFree(p);
        delete c->amg;
        c->amg = nullptr;
    }

    namespace
    {
        // The current assembled operator on free rows and columns (solver
        // order), pdiag on the diagonal, from the device layout.
        cbsamg::Csr fine_operator(cbs_hip_ctx* c, std::vector<uint8_t>& active)
        {
            const cbs_hip_sell& m = c->sell;
            const int N = c->N;
            // the values and the gathered diagonal come from kernels on the
            // context's stream, which the copies below (default stream) do
            // not wait for
            hip_check(hipStreamSynchronize(c->stream), "multigrid operator");
            const auto row_base = download(m.row_base, size_t(m.ntiles));
            const auto row_cnt = download(m.row_cnt, size_t(m.ntiles));
            const auto halo_ptr = download(m.halo_ptr, size_t(m.ntiles) + 1);
            const auto halo = download(m.halo, size_t(halo_ptr.back()));
            const auto slice_ptr = download(m.slice_ptr, size_t(m.ntiles) + 1);
            const auto slice_off = download(m.slice_off, size_t(slice_ptr.back()) + 1);
            const auto cols = download(m.cols, size_t(m.entries / 2));
            const auto val = download(m.v64, size_t(m.entries));
            const auto pdiag = download(c->sv_pdiag, size_t(N));
            const auto flag = download(c->sv_flag, size_t(N));
            active.assign(size_t(N), 0);
            for (int i = 0; i < N; ++i) active[size_t(i)] = (flag[size_t(i)] & 1u) && !(flag[size_t(i)] & 4u);
            std::vector<std::vector<std::pair<int32_t, double>>> rows(static_cast<size_t>(N));
            for (int t = 0; t < m.ntiles; ++t)
            {
                const int rb = row_base[size_t(t)], nr = row_cnt[size_t(t)], hb = halo_ptr[size_t(t)];
                for (int s = slice_ptr[size_t(t)]; s < slice_ptr[size_t(t) + 1]; ++s)
                {
                    const int64_t off = slice_off[size_t(s)];
                    const int np = int((slice_off[size_t(s) + 1] - off) >> 7);
                    for (int l = 0; l < 64; ++l)
                    {
                        const int row = (s - slice_ptr[size_t(t)]) * 64 + l;
                        if (row >= nr) continue;
                        const int i = rb + row;
                        if (!active[size_t(i)]) continue;
                        for (int k = 0; k < 2 * np; ++k)
                        {
                            const int64_t q = off + int64_t(k >> 1) * 128 + 2 * l + (k & 1);
                            const double v = val[size_t(q)];
                            if (v == 0.0) continue;  // slice padding or a zero coupling
                            const uint32_t w = cols[size_t(q >> 1)];
                            const int lc = int((k & 1) ? (w >> 16) : (w & 0xffffu));
                            const int j = lc < nr ? rb + lc : halo[size_t(hb + lc - nr)];
                            if (active[size_t(j)]) rows[size_t(i)].push_back({j, v});
                        }
                    }
                }
            }
            cbsamg::Csr A;
            A.n = A.m = N;
            for (int i = 0; i < N; ++i)
            {
                auto& r = rows[size_t(i)];
                if (active[size_t(i)]) r.push_back({i, pdiag[size_t(i)]});
                std::sort(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
                for (const auto& e : r)
                {
                    A.col.push_back(e.first);
                    A.val.push_back(e.second);
                }
                A.ptr.push_back(int64_t(A.col.size()));
                r.clear();
                r.shrink_to_fit();
            }
            return A;
        }

        // lanes per row of a CSR product: 64 on levels whose rows hold more
        // than 96 entries on average (few long rows: a 16-lane row walk is
        // latency-bound there, 54 -> 16 us for the last restriction at
        // cavity160), otherwise 16 (32 lanes measured slower on the 70-entry
        // rows of level 1)
        int lanes_for(int64_t nnz, int n)
        {
            const double mean = n > 0 ? double(nnz) / n : 0.0;
            return mean > 96 ? 64 : 16;
        }

        void upload_csr(AmgState* st, hipStream_t stream, const cbsamg::Csr& A, int32_t*& ptr, int32_t*& col,
                        float*& val)
        {
            if (A.nnz() >= (int64_t(1) << 31)) throw std::runtime_error("multigrid operator too large");
            ptr = amg_upload<int32_t>(st, stream, A.ptr);
            col = amg_upload<int32_t>(st, stream, A.col);
            val = amg_upload<float>(st, stream, A.val);
            st->stored += A.nnz();
        }

        // rows sorted by length within windows of this many (AmgSell)
        constexpr int kAmgWindow = 1024;

        // A CSR matrix of the hierarchy as AmgSell (values rounded to FP32).
        AmgSell upload_sell(AmgState* st, hipStream_t stream, const cbsamg::Csr& A)
        {
            if (A.nnz() >= (int64_t(1) << 31)) throw std::runtime_error("multigrid operator too large");
            const int n = A.n;
            const int nslices = (n + 63) / 64;
            std::vector<int32_t> order(size_t(nslices) * 64, -1);
            for (int w0 = 0; w0 < n; w0 += kAmgWindow)
            {
                const int w1 = std::min(n, w0 + kAmgWindow);
                std::vector<int32_t> rows(size_t(w1 - w0));
                for (int i = w0; i < w1; ++i) rows[size_t(i - w0)] = i;
                std::stable_sort(rows.begin(), rows.end(), [&](int32_t a, int32_t b) {
                    return A.ptr[size_t(a) + 1] - A.ptr[size_t(a)] > A.ptr[size_t(b) + 1] - A.ptr[size_t(b)];
                });
                for (size_t k = 0; k < rows.size(); ++k) order[size_t(w0) + k] = rows[k];
            }
            std::vector<int64_t> off(size_t(nslices) + 1, 0);
            for (int s = 0; s < nslices; ++s)
            {
                int64_t width = 0;
                for (int l = 0; l < 64; ++l)
                {
                    const int r = order[size_t(s) * 64 + size_t(l)];
                    if (r >= 0) width = std::max(width, A.ptr[size_t(r) + 1] - A.ptr[size_t(r)]);
                }
                off[size_t(s) + 1] = off[size_t(s)] + 64 * width;
            }
            std::vector<int32_t> col(size_t(off.back()), 0);
            std::vector<float> val(size_t(off.back()), 0.0f);
            for (int s = 0; s < nslices; ++s)
                for (int l = 0; l < 64; ++l)
                {
                    const int r = order[size_t(s) * 64 + size_t(l)];
                    if (r < 0) continue;
                    for (int64_t k = 0; k < A.ptr[size_t(r) + 1] - A.ptr[size_t(r)]; ++k)
                    {
                        const size_t q = size_t(off[size_t(s)] + k * 64 + l), e = size_t(A.ptr[size_t(r)] + k);
                        col[q] = A.col[e];
                        val[q] = float(A.val[e]);
                    }
                }
            AmgSell m;
            m.nslices = nslices;
            m.off = amg_upload<int64_t>(st, stream, off);
            m.row = amg_upload<int32_t>(st, stream, order);
            m.col = amg_upload<int32_t>(st, stream, col);
            m.val = amg_upload<float>(st, stream, val);
            st->stored += A.nnz();
            return m;
        }

        void build(cbs_hip_ctx* c)
        {
            const auto t0 = std::chrono::steady_clock::now();
            std::vector<uint8_t> active;
            const cbsamg::Csr A0 = fine_operator(c, active);
            cbsamg::Hierarchy H;
            std::string err;
            if (!cbsamg::setup(A0, active, H, err)) throw std::runtime_error(err);
            auto* st = new AmgState;
            c->amg = st;
            for (int i = 0; i < A0.n; ++i)
                if (active[size_t(i)])
                    for (int64_t k = A0.ptr[i]; k < A0.ptr[i + 1]; ++k)
                        if (A0.col[size_t(k)] == i) st->scale_ref += A0.val[size_t(k)];
            const size_t L = H.lv.size();
            st->lv.resize(L);
            for (size_t l = 0; l < L; ++l)
            {
                const cbsamg::Level& h = H.lv[l];
                AmgState::Level& d = st->lv[l];
                d.n = h.A.n;
                d.omega = float(h.omega);
                if (l > 0)
                {
                    upload_csr(st, c->stream, h.A, d.a_ptr, d.a_col, d.a_val);
                    d.a_lanes = lanes_for(h.A.nnz(), h.A.n);
                    d.dinv = amg_upload<float>(st, c->stream, h.dinv);
                    d.b = amg_alloc<float>(st, c->stream, size_t(d.n));
                    d.x = amg_alloc<float>(st, c->stream, size_t(d.n));
                    d.t = amg_alloc<float>(st, c->stream, size_t(d.n));
                }
                if (l + 1 < L)
                {
                    d.p = upload_sell(st, c->stream, h.P);
                    upload_csr(st, c->stream, h.R, d.r_ptr, d.r_col, d.r_val);
                    d.r_lanes = lanes_for(h.R.nnz(), h.R.n);
                }
            }
            st->inv = amg_upload<float>(st, c->stream, H.inv);
            AmgState::Level& f = st->lv[0];
            const size_t N = size_t(c->N);
            f.dinv = amg_alloc<float>(st, c->stream, N);
            f.b = amg_alloc<float>(st, c->stream, N);  // the FP32 residual
            f.x = amg_alloc<float>(st, c->stream, N);  // z
            f.t = amg_alloc<float>(st, c->stream, N);
            st->rf = f.b;
            st->z = f.x;
            st->rowsumf = amg_alloc<float>(st, c->stream, N);
            st->scale = amg_alloc<double>(st, c->stream, 1);
            // every fill and upload above complete before the context's
            // stream uses the hierarchy
            hip_check(hipDeviceSynchronize(), "multigrid set-up");
            st->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }

        int grid_lanes(int n, int lanes) { return int((int64_t(n) * lanes + kBlock - 1) / kBlock); }

        template <typename... A>
        void launch_spmv(int lanes, hipStream_t s, int n, A... a)
        {
            const dim3 g(grid_lanes(n, lanes)), b(kBlock);
            switch (lanes)
            {
            case 4: hipLaunchKernelGGL(k_amg_spmv<4>, g, b, 0, s, n, a...); break;
            case 8: hipLaunchKernelGGL(k_amg_spmv<8>, g, b, 0, s, n, a...); break;
            case 16: hipLaunchKernelGGL(k_amg_spmv<16>, g, b, 0, s, n, a...); break;
            case 32: hipLaunchKernelGGL(k_amg_spmv<32>, g, b, 0, s, n, a...); break;
            default: hipLaunchKernelGGL(k_amg_spmv<64>, g, b, 0, s, n, a...); break;
            }
        }

        template <typename... A>
        void launch_restrict(int lanes, hipStream_t s, int n, A... a)
        {
            const dim3 g(grid_lanes(n, lanes)), b(kBlock);
            switch (lanes)
            {
            case 4: hipLaunchKernelGGL(k_amg_restrict<4>, g, b, 0, s, n, a...); break;
            case 8: hipLaunchKernelGGL(k_amg_restrict<8>, g, b, 0, s, n, a...); break;
            case 16: hipLaunchKernelGGL(k_amg_restrict<16>, g, b, 0, s, n, a...); break;
            case 32: hipLaunchKernelGGL(k_amg_restrict<32>, g, b, 0, s, n, a...); break;
            default: hipLaunchKernelGGL(k_amg_restrict<64>, g, b, 0, s, n, a...); break;
            }
        }

        void launch_prolong(hipStream_t s, const AmgSell& p, const float* xc, float* x, const int* done)
        {
            if (p.nslices == 0) return;
            hipLaunchKernelGGL(k_amg_prolong, dim3(grid_for(int64_t(p.nslices) * 64)), dim3(kBlock), 0, s,
                               p.nslices, p.off, p.row, p.col, p.val, xc, x, done);
        }

        // z = M rf (st->lv[0].b -> st->lv[0].x)
        void vcycle(cbs_hip_ctx* c, AmgState* st, const int* done)
        {
            const size_t L = st->lv.size();
            hipStream_t s = c->stream;
            const double* theta = st->scale;
            for (size_t l = 0; l + 1 < L; ++l)
            {
                AmgState::Level& v = st->lv[l];
                AmgState::Level& w = st->lv[l + 1];
                const double* th = l == 0 ? nullptr : theta;
                hipLaunchKernelGGL(k_amg_jacobi0, dim3(grid_for(v.n)), dim3(kBlock), 0, s, v.n, v.dinv, v.omega, th,
                                   v.b, v.x, done);
                if (l == 0) sell_apply32(c, v.x, v.t, st->rowsumf, done);
                else
                    launch_spmv(v.a_lanes, s, v.n, v.a_ptr, v.a_col, v.a_val, theta, v.x, v.t, done);
                launch_restrict(v.r_lanes, s, w.n, v.r_ptr, v.r_col, v.r_val, v.b, v.t, w.b, done);
            }
            AmgState::Level& cl = st->lv[L - 1];
            hipLaunchKernelGGL(k_amg_dense, dim3(grid_for(int64_t(cl.n) * 64)), dim3(kBlock), 0, s, cl.n, st->inv, theta,
                               cl.b, cl.x,
                               done);
            for (size_t l = L - 1; l-- > 0;)
            {
                AmgState::Level& v = st->lv[l];
                AmgState::Level& w = st->lv[l + 1];
                const double* th = l == 0 ? nullptr : theta;
                launch_prolong(s, v.p, w.x, v.x, done);
                if (l == 0) sell_apply32(c, v.x, v.t, st->rowsumf, done);
                else
                    launch_spmv(v.a_lanes, s, v.n, v.a_ptr, v.a_col, v.a_val, theta, v.x, v.t, done);
                hipLaunchKernelGGL(k_amg_jacobi, dim3(grid_for(v.n)), dim3(kBlock), 0, s, v.n, v.dinv, v.omega, th,
                                   v.b, v.t, v.x, done);
            }
        }
    }

    double amg_setup_seconds(const cbs_hip_ctx* c) { return c->amg ? c->amg->setup_seconds : 0.0; }
    int amg_levels(const cbs_hip_ctx* c) { return c->amg ? int(c->amg->lv.size()) : 0; }
    int64_t amg_stored(const cbs_hip_ctx* c) { return c->amg ? c->amg->stored : 0; }

    int solve_pressure_amg(cbs_hip_ctx* c, cbs_cg_result* res)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const int N = c->N;
        phase_reset(c);
        reset_cg_scalars(c);
        upload_cg_scalars(c);
        timed_group(c, PH_CONSTRAINT, [&] { launch_constraints(c); });
        sol_gather(c, c->pres, c->sv_x);
        sol_gather(c, c->rhs1, c->sv_b);
        sol_gather(c, c->pdiag, c->sv_pdiag);
        sell_refresh_values(c);
        if (!c->amg) build(c);
        AmgState* st = c->amg;
        const BlockMap nb = c->nblk.map();
        const int32_t* gid = c->sell.gid;
        hipLaunchKernelGGL(k_amg_prepare, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag, c->sv_pdiag,
                           st->lv[0].dinv, c->partial, c->scal, gid);
        sell_rowsum32(c, c->cg_p, c->cg_ap, st->rowsumf);
        hipLaunchKernelGGL(k_amg_theta, dim3(1), dim3(kBlock), 0, c->stream, c->partial, c->nblk.n, st->scale_ref,
                           st->scale);
        sell_autotune64(c, c->sv_x, c->cg_ap);
        sell_autotune32(c, st->lv[0].x, st->lv[0].t, st->rowsumf);
        launch_fill(c, c->cg_ap, N, 0.0);
        launch_apply_fast(c, c->sv_x, c->cg_ap);
        hipLaunchKernelGGL(k_amg_init, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag, c->sv_b,
                           c->sv_pdiag, c->sv_x, c->cg_ap, c->cg_r, st->rf, c->partial);
        hipLaunchKernelGGL(k_amg_init_fin, dim3(1), dim3(kBlock), 0, c->stream, c->partial, c->nblk.seg_ptr, c->scal);
        const int* done = &c->scal->done;
        // z = M r, p = z
        vcycle(c, st, done);
        hipLaunchKernelGGL(k_amg_dot, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag, c->cg_r, st->z,
                           c->partial, c->scal);
        hipLaunchKernelGGL(k_amg_dot_fin, dim3(1), dim3(kBlock), 0, c->stream, c->partial, c->nblk.seg_ptr, 1, c->scal);
        hipLaunchKernelGGL(k_amg_direction, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag, st->z,
                           c->cg_p, c->scal);
        hip_check(hipGetLastError(), "multigrid CG start");
        download_scalars(c);
        const double setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        int chunk = 4;
        int64_t seq = 0;
        while (!c->h_scal[0].done && c->h_scal[0].err == ~0ull)
        {
            const int iter_before = c->h_scal[0].iter;
            launch_iterations(
                c, 2, CBS_PCG_AMG, chunk, false, seq, [&] { sell_apply64(c, c->cg_p, c->cg_ap, true, done); },
                [&](const PhaseMark& mark) {
                    hipLaunchKernelGGL(k_amg_update, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       c->cg_p, c->cg_ap, c->sv_x, c->cg_r, st->rf, c->partial, c->scal, c->dotpart,
                                       c->sell.seg_tile_ptr);
                    mark(PH_VECTOR);
                    hipLaunchKernelGGL(k_amg_update_fin, dim3(1), dim3(kBlock), 0, c->stream, c->partial,
                                       c->nblk.seg_ptr, c->scal);
                    mark(PH_DOT);
                    vcycle(c, st, done);
                    mark(PH_PRECOND);
                    hipLaunchKernelGGL(k_amg_dot, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       c->cg_r, st->z, c->partial, c->scal);
                    hipLaunchKernelGGL(k_amg_dot_fin, dim3(1), dim3(kBlock), 0, c->stream, c->partial,
                                       c->nblk.seg_ptr, 0, c->scal);
                    mark(PH_DOT);
                    hipLaunchKernelGGL(k_amg_direction, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, nb, c->sv_flag,
                                       st->z, c->cg_p, c->scal);
                    mark(PH_VECTOR);
                });
            hip_check(hipGetLastError(), "multigrid CG launch");
            download_scalars(c, iter_before);
            const DevScalars& h = c->h_scal[0];
            if (h.done || h.err != ~0ull) break;
            int next = 8;
            if (h.iter > 0 && h.rel_l2 > 0.0 && h.rel_l2 < 1.0 && c->cfg.cg_conv_test == 2)
            {
                const double rate = std::log(h.rel_l2) / h.iter;
                next = int(std::ceil(std::log(c->cfg.relToler / h.rel_l2) / rate * 0.9));
            }
            chunk = std::max(1, std::min(next, std::min(64, c->cfg.cg_max_iter - h.iter)));
        }
        sol_scatter(c, c->sv_x, c->pres);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        collect_results(c, res, secs, 0.0);
        phase_fill(c, res, setup_s);
        res->preconditioner_calls = res->iterations + 1;
        res->dot_calls = 2 * res->iterations + 2;
        c->seg_result.assign(1, *res);
        return check_device_error(c);
    }
}
