// Matrix-free pressure operator and Jacobi-PCG on the GPU.
//
// FP64 path: MatrixVectorCalc::pressureMultiply (src/linalg/
// MatrixVectorCalc.cpp:125-205) and ConjugateGradient::solvePressure
// (src/linalg/ConjugateGradient.cpp:584-869).
//
// The apply is one launch per colour over the fluid elements of that colour
// in the reference's Coloring order.  No two elements of a colour share a
// node, so each node receives at most one contribution per launch, and the
// launches run in colour order: every node therefore accumulates exactly the
// reference's sequence of additions (y = 0, + colour 0 term, + colour 1 term,
// ...), and the FP64 apply is bit-identical to the CPU.  Each thread streams
// one element: an int4 of node ids and six coefficient planes, all coalesced.
//
// The CG vector work is fused into three kernels per iteration that also
// produce the block partials of the dot products; the last block to finish
// reduces the partials in a fixed order (deterministic) and updates the
// device-resident scalars (alpha, beta, norms, convergence flag).  The host
// only reads the convergence flag between chunks of iterations.  The dot
// products are tree sums, so they differ from the reference's ascending
// serial sums by rounding only.
//
// Mixed form: the same PCG in FP32 (apply with FP32 coefficients and
// vectors, Jacobi in FP32, every reduction accumulated in FP64) solves the
// correction equation A d = r to a loose relative tolerance inside an FP64
// iterative-refinement loop whose true residual b - A x is recomputed with
// the FP64 apply.  Convergence is tested on that true residual with the
// reference's criterion.
#include <chrono>
#include <cmath>

#include "device_common.hpp"

namespace cbsdev
{
    // ------------------------------------------------------------------
    // FP64 apply
    // ------------------------------------------------------------------
    __global__ __launch_bounds__(kBlock) void k_apply_color(const int4* __restrict__ conn,
                                                            const double* __restrict__ g, int64_t F,
                                                            const double* __restrict__ x,
                                                            double* __restrict__ y, int beg, int end,
                                                            const int* __restrict__ done)
    {
        if (*done) return;
        const int k = beg + int(blockIdx.x) * kBlock + int(threadIdx.x);
        if (k >= end) return;
        const int4 n = conn[k];
        const double x1 = x[n.x], x2 = x[n.y], x3 = x[n.z], x4 = x[n.w];
        const double a12 = g[k], a13 = g[F + k], a14 = g[2 * F + k];
        const double a23 = g[3 * F + k], a24 = g[4 * F + k], a34 = g[5 * F + k];
        y[n.x] += a12 * x2 + a13 * x3 + a14 * x4;
        y[n.y] += a12 * x1 + a23 * x3 + a24 * x4;
        y[n.z] += a13 * x1 + a23 * x2 + a34 * x4;
        y[n.w] += a14 * x1 + a24 * x2 + a34 * x3;
    }

    // Event pairs around apply sweeps; collected after each synchronisation.
    void timer_begin(cbs_hip_ctx* c, int precision)
    {
        if (!c->profile) return;
        if (size_t(c->events_used + 2) > c->events.size())
        {
            const size_t old = c->events.size();
            c->events.resize(old + 512);
            for (size_t i = old; i < c->events.size(); ++i) hip_check(hipEventCreate(&c->events[i]), "hipEventCreate");
        }
        hip_check(hipEventRecord(c->events[size_t(c->events_used)], c->stream), "hipEventRecord");
        c->event_prec.push_back(precision);
    }

    void timer_end(cbs_hip_ctx* c)
    {
        if (!c->profile) return;
        hip_check(hipEventRecord(c->events[size_t(c->events_used) + 1], c->stream), "hipEventRecord");
        c->events_used += 2;
    }

    void timer_collect(cbs_hip_ctx* c)
    {
        for (int i = 0; i + 1 < c->events_used; i += 2)
        {
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, c->events[size_t(i)], c->events[size_t(i) + 1]), "hipEventElapsedTime");
            const int p = c->event_prec[size_t(i / 2)] == CBS_PCG_MIXED ? 1 : 0;
            c->apply_ms[p] += ms;
            c->apply_n[p] += 1;
        }
        c->events_used = 0;
        c->event_prec.clear();
    }

    void launch_apply(cbs_hip_ctx* c, const double* x, double* acc, const int* done)
    {
        const int64_t F = c->F;
        for (int col = 0; col < c->ncolor; ++col)
        {
            const int beg = c->h_fcolor_ptr[col], end = c->h_fcolor_ptr[col + 1];
            if (end <= beg) continue;
            hipLaunchKernelGGL(k_apply_color, dim3(grid_for(end - beg)), dim3(kBlock), 0, c->stream, c->slot_conn,
                               c->gstif_s, F, x, acc, beg, end, done);
        }
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

    // gstif = gstifE * delte, in slot order (TimeStep.cpp:873-901), plus the
    // FP32 copy used by the mixed-precision apply.
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
            gf[p * F + k] = float(v);
        }
    }

    void launch_gstif(cbs_hip_ctx* c)
    {
        if (c->F == 0) return;
        hipLaunchKernelGGL(k_gstif, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->slot_elem, c->gstifE_s,
                           c->delte, c->gstif_s, c->gstif_f, int64_t(c->F));
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
            pres[ip] = val[i];
            rhs1[ip] = val[i];
            pdiag[ip] = 1.0;
        }
    }

    void launch_constraints(cbs_hip_ctx* c)
    {
        if (c->nnonfree > 0)
            hipLaunchKernelGGL(k_constraints, dim3(grid_for(c->nnonfree)), dim3(kBlock), 0, c->stream, c->nf_node,
                               c->nf_mode, c->nf_val, c->nnonfree, c->pres, c->rhs1, c->pdiag);
    }

    __device__ inline bool is_free(uint8_t f) { return (f & 1u) && !(f & 4u); }

    enum CgCheck : uint32_t
    {
        CK_BAD_DIAG = 1, CK_BAD_R = 2, CK_RZ_INIT = 3, CK_BREAKDOWN = 4, CK_RZ_UPD = 5, CK_NOT_FREE = 6
    };

    // Preconditioner checks of apply_preconditioner (ConjugateGradient.cpp:302-369).
    __device__ inline void jacobi_check(DevScalars* s, int precond, double diag, double r, int ip,
                                        unsigned long long* pend)
    {
        if (precond == 1 && (!isfinite(diag) || diag <= 1.0e-30))
            atomicMin(pend, err_key(ST_CG, 1, uint32_t(ip + 1), CK_BAD_DIAG));
        else if (!isfinite(r))
            atomicMin(pend, err_key(ST_CG, 1, uint32_t(ip + 1), CK_BAD_R));
        (void)s;
    }

    // Initial residual, preconditioned residual and first direction
    // (ConjugateGradient.cpp:634-723):
    //   r = free ? rhs1 - (acc + pdiag*pres) : 0,  z = r/pdiag,  p = z.
    __global__ __launch_bounds__(kBlock) void k_cg_init(const uint8_t* __restrict__ nflag,
                                                         const double* __restrict__ rhs1,
                                                         const double* __restrict__ pdiag,
                                                         const double* __restrict__ pres, double* __restrict__ acc,
                                                         double* __restrict__ r, double* __restrict__ p, int n,
                                                         int precond, double* __restrict__ part, DevScalars* s)
    {
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        {
            double ri = 0.0, pi = 0.0;
            if (is_free(nflag[i]))
            {
                const double ap = acc[i] + pdiag[i] * pres[i];
                ri = rhs1[i] - ap;
                const double d = pdiag[i];
                jacobi_check(s, precond, d, ri, i, &s->pending_err);
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
        if (!last_block_done(&s->counter[0])) return;
        const double rr = reduce_partials(part, gridDim.x, 4, 0, lds);
        const double rz = reduce_partials(part, gridDim.x, 4, 1, lds);
        const double mx = max_partials(part, gridDim.x, 4, 2, ldm);
        if (threadIdx.x == 0)
        {
            s->counter[0] = 0;
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
                raise_error(s, s->pending_err);
                s->done = 1;
                return;
            }
            if (rz <= 1.0e-300 || !isfinite(rz))
            {
                raise_error(s, err_key(ST_CG, 2, 0, CK_RZ_INIT));
                s->done = 1;
                return;
            }
            s->rz_old = rz;
        }
    }

    // Ap = free ? acc + pdiag*p : 0 (in place), pAp, alpha = rz_old / pAp
    // (ConjugateGradient.cpp:727-758).
    __global__ __launch_bounds__(kBlock) void k_cg_ap(const uint8_t* __restrict__ nflag,
                                                       const double* __restrict__ pdiag,
                                                       const double* __restrict__ p, double* __restrict__ ap, int n,
                                                       double* __restrict__ part, DevScalars* s)
    {
        if (s->done) return;
        __shared__ double lds[1][kBlock / 64];
        double v[1] = {0.0};
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        {
            double a = 0.0;
            if (is_free(nflag[i]))
            {
                a = ap[i] + pdiag[i] * p[i];
                v[0] += p[i] * a;
            }
            ap[i] = a;
        }
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) part[blockIdx.x * 4 + 0] = v[0];
        if (!last_block_done(&s->counter[1])) return;
        const double pap = reduce_partials(part, gridDim.x, 4, 0, lds);
        if (threadIdx.x == 0)
        {
            s->counter[1] = 0;
            s->pAp = pap;
            if (fabs(pap) <= 1.0e-300 || !isfinite(pap))
            {
                raise_error(s, err_key(ST_CG, 3, uint32_t(s->iter + 1), CK_BREAKDOWN));
                s->done = 1;
                return;
            }
            s->alpha = s->rz_old / pap;
        }
    }

    // pres += alpha p, r -= alpha Ap on free nodes; ||r||2, max|r|; z = r/pdiag,
    // (r,z); convergence test; beta (ConjugateGradient.cpp:760-852).
    __global__ __launch_bounds__(kBlock) void k_cg_update(const uint8_t* __restrict__ nflag,
                                                           const double* __restrict__ pdiag,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ ap,
                                                           double* __restrict__ pres, double* __restrict__ r, int n,
                                                           int precond, double* __restrict__ part, DevScalars* s)
    {
        if (s->done) return;
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        const double alpha = s->alpha;
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        {
            if (!is_free(nflag[i])) continue;
            pres[i] += alpha * p[i];
            const double ri = r[i] - alpha * ap[i];
            r[i] = ri;
            const double d = pdiag[i];
            jacobi_check(s, precond, d, ri, i, &s->pending_err);
            const double zi = (precond == 1) ? ri / d : ri;
            v[0] += ri * ri;
            v[1] += ri * zi;
            const double a = fabs(ri);
            m = (m < a) ? a : m;
        }
        block_sum<2>(v, lds);
        m = block_max(m, ldm);
        if (threadIdx.x == 0)
        {
            part[blockIdx.x * 4 + 0] = v[0];
            part[blockIdx.x * 4 + 1] = v[1];
            part[blockIdx.x * 4 + 2] = m;
        }
        if (!last_block_done(&s->counter[2])) return;
        const double rr = reduce_partials(part, gridDim.x, 4, 0, lds);
        const double rz = reduce_partials(part, gridDim.x, 4, 1, lds);
        const double mx = max_partials(part, gridDim.x, 4, 2, ldm);
        if (threadIdx.x == 0)
        {
            s->counter[2] = 0;
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
                return;
            }
            if (s->pending_err != ~0ull)
            {
                raise_error(s, s->pending_err);
                s->done = 1;
                return;
            }
            if (rz <= 1.0e-300 || !isfinite(rz))
            {
                raise_error(s, err_key(ST_CG, 4, uint32_t(s->iter), CK_RZ_UPD));
                s->done = 1;
                return;
            }
            s->beta = rz / s->rz_old;
            s->rz_old = rz;
            if (s->iter >= s->max_iter) s->done = 1;
        }
    }

    // p = free ? r/pdiag + beta p : 0, and clear the apply accumulator.
    __global__ void k_cg_direction(const uint8_t* __restrict__ nflag, const double* __restrict__ pdiag,
                                   const double* __restrict__ r, double* __restrict__ p, double* __restrict__ acc,
                                   int n, int precond, const DevScalars* s)
    {
        if (s->done) return;
        const double beta = s->beta;
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        double pi = 0.0;
        if (is_free(nflag[i]))
        {
            const double zi = (precond == 1) ? r[i] / pdiag[i] : r[i];
            pi = zi + beta * p[i];
        }
        p[i] = pi;
        acc[i] = 0.0;
    }

    static void reset_cg_scalars(cbs_hip_ctx* c)
    {
        DevScalars& h = *c->h_scal;
        h = DevScalars{};
        h.err = ~0ull;
        h.pending_err = ~0ull;
        h.max_iter = c->cfg.cg_max_iter;
        h.conv_test = c->cfg.cg_conv_test;
        h.rel_tol = c->cfg.relToler;
        h.abs_tol = c->cfg.absToler;
    }

    // Pushes only the CG part of the scalars (the rest stays device-owned).
    static void upload_cg_scalars(cbs_hip_ctx* c)
    {
        const size_t off = offsetof(DevScalars, rz_old);
        const size_t len = offsetof(DevScalars, counter) - off;
        hip_check(hipMemcpyAsync(reinterpret_cast<char*>(c->scal) + off, reinterpret_cast<char*>(c->h_scal) + off,
                                 len, hipMemcpyHostToDevice, c->stream),
                  "upload CG scalars");
    }

    static void download_scalars(cbs_hip_ctx* c)
    {
        hip_check(hipMemcpyAsync(c->h_scal, c->scal, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream),
                  "download scalars");
        hip_check(hipStreamSynchronize(c->stream), "stream sync");
        timer_collect(c);
    }

    static double bytes_fp64_iter(const cbs_hip_ctx* c)
    {
        // B_apply = F*(6*8 + 16) + N*3*8 ; B_vec = N*(8*8 + 1)
        return double(c->F) * 64.0 + double(c->N) * 24.0 + double(c->N) * 65.0;
    }

    static double bytes_fp32_iter(const cbs_hip_ctx* c)
    {
        return double(c->F) * 40.0 + double(c->N) * 12.0 + double(c->N) * 33.0;
    }

    int solve_pressure_fp64(cbs_hip_ctx* c, cbs_cg_result* res)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const int N = c->N;
        const int pre = c->cfg.cg_preconditioner;
        reset_cg_scalars(c);
        upload_cg_scalars(c);
        // pending_err and err are outside the uploaded window: clear them.
        hip_check(hipMemcpyAsync(&c->scal->pending_err, &c->h_scal->pending_err, sizeof(unsigned long long),
                                 hipMemcpyHostToDevice, c->stream),
                  "clear pending");
        if (c->nnonfree > 0)
        {
            hipLaunchKernelGGL(k_constraints, dim3(grid_for(c->nnonfree)), dim3(kBlock), 0, c->stream, c->nf_node,
                               c->nf_mode, c->nf_val, c->nnonfree, c->pres, c->rhs1, c->pdiag);
        }
        launch_fill(c, c->cg_ap, N, 0.0);
        launch_apply(c, c->pres, c->cg_ap, &c->scal->pad0);
        const int rg = reduce_grid(N);
        hipLaunchKernelGGL(k_cg_init, dim3(rg), dim3(kBlock), 0, c->stream, c->nflag, c->rhs1, c->pdiag, c->pres,
                           c->cg_ap, c->cg_r, c->cg_p, N, pre, c->partial, c->scal);
        download_scalars(c);
        int iters_launched = 0;
        int chunk = 4;
        while (!c->h_scal->done && c->h_scal->err == ~0ull)
        {
            for (int k = 0; k < chunk; ++k)
            {
                timer_begin(c, CBS_PCG_FP64);
                launch_apply(c, c->cg_p, c->cg_ap, &c->scal->done);
                timer_end(c);
                hipLaunchKernelGGL(k_cg_ap, dim3(rg), dim3(kBlock), 0, c->stream, c->nflag, c->pdiag, c->cg_p,
                                   c->cg_ap, N, c->partial, c->scal);
                hipLaunchKernelGGL(k_cg_update, dim3(rg), dim3(kBlock), 0, c->stream, c->nflag, c->pdiag, c->cg_p,
                                   c->cg_ap, c->pres, c->cg_r, N, pre, c->partial, c->scal);
                hipLaunchKernelGGL(k_cg_direction, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->nflag,
                                   c->pdiag, c->cg_r, c->cg_p, c->cg_ap, N, pre, c->scal);
            }
            iters_launched += chunk;
            hip_check(hipGetLastError(), "CG launch");
            download_scalars(c);
            // Choose the next chunk from the observed convergence rate.
            const DevScalars& h = *c->h_scal;
            if (h.done || h.err != ~0ull) break;
            int next = 64;
            if (h.iter > 0 && h.rel_l2 > 0.0 && h.rel_l2 < 1.0 && c->cfg.cg_conv_test == 2)
            {
                const double rate = std::log(h.rel_l2) / h.iter;
                const double need = std::log(c->cfg.relToler / h.rel_l2) / rate;
                next = int(std::ceil(need * 0.9));
            }
            chunk = std::max(1, std::min(next, std::min(256, c->cfg.cg_max_iter - h.iter)));
        }
        const DevScalars& h = *c->h_scal;
        res->iterations = h.iter;
        res->converged = h.converged;
        res->initial_l2 = h.initial_l2;
        res->final_l2 = h.final_l2;
        res->final_relative_l2 = h.rel_l2;
        res->final_max_abs = h.max_abs;
        res->matvec_calls = 1 + h.iter;
        res->dot_calls = 2 + 3 * h.iter;
        res->preconditioner_calls = 1 + h.iter;
        res->refinements = 0;
        res->algorithmic_bytes = bytes_fp64_iter(c) * h.iter;
        res->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        (void)iters_launched;
        return check_device_error(c);
    }

    // ------------------------------------------------------------------
    // Mixed precision
    // ------------------------------------------------------------------
    __global__ __launch_bounds__(kBlock) void k_apply_color_f32(const int4* __restrict__ conn,
                                                                const float* __restrict__ g, int64_t F,
                                                                const float* __restrict__ x, float* __restrict__ y,
                                                                int beg, int end, const int* __restrict__ done)
    {
        if (*done) return;
        const int k = beg + int(blockIdx.x) * kBlock + int(threadIdx.x);
        if (k >= end) return;
        const int4 n = conn[k];
        const float x1 = x[n.x], x2 = x[n.y], x3 = x[n.z], x4 = x[n.w];
        const float a12 = g[k], a13 = g[F + k], a14 = g[2 * F + k];
        const float a23 = g[3 * F + k], a24 = g[4 * F + k], a34 = g[5 * F + k];
        y[n.x] += a12 * x2 + a13 * x3 + a14 * x4;
        y[n.y] += a12 * x1 + a23 * x3 + a24 * x4;
        y[n.z] += a13 * x1 + a23 * x2 + a34 * x4;
        y[n.w] += a14 * x1 + a24 * x2 + a34 * x3;
    }

    static void launch_apply_f32(cbs_hip_ctx* c, const float* x, float* acc, const int* done)
    {
        const int64_t F = c->F;
        for (int col = 0; col < c->ncolor; ++col)
        {
            const int beg = c->h_fcolor_ptr[col], end = c->h_fcolor_ptr[col + 1];
            if (end <= beg) continue;
            hipLaunchKernelGGL(k_apply_color_f32, dim3(grid_for(end - beg)), dim3(kBlock), 0, c->stream,
                               c->slot_conn, c->gstif_f, F, x, acc, beg, end, done);
        }
    }

    // True residual in FP64: r = free ? rhs1 - (acc + pdiag*x) : 0, its norms,
    // and the FP32 start of the inner solve: rf = r, d = 0, z = dinv*rf, p = z.
    __global__ __launch_bounds__(kBlock) void k_mx_true_residual(
        const uint8_t* __restrict__ nflag, const double* __restrict__ rhs1, const double* __restrict__ pdiag,
        const double* __restrict__ x, double* __restrict__ acc, float* __restrict__ rf, float* __restrict__ pf,
        float* __restrict__ apf, double* __restrict__ dx, const float* __restrict__ dinv, int n,
        double* __restrict__ part, DevScalars* s)
    {
        __shared__ double lds[2][kBlock / 64];
        __shared__ double ldm[kBlock / 64];
        double v[2] = {0.0, 0.0};
        double m = 0.0;
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        {
            double ri = 0.0;
            float zf = 0.0f, rff = 0.0f;
            if (is_free(nflag[i]))
            {
                ri = rhs1[i] - (acc[i] + pdiag[i] * x[i]);
                rff = float(ri);
                zf = rff * dinv[i];
                v[0] += ri * ri;
                v[1] += double(rff) * double(zf);
                const double a = fabs(ri);
                m = (m < a) ? a : m;
            }
            rf[i] = rff;
            pf[i] = zf;
            apf[i] = 0.0f;
            dx[i] = 0.0;
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
        if (!last_block_done(&s->counter[3])) return;
        const double rr = reduce_partials(part, gridDim.x, 4, 0, lds);
        const double rz = reduce_partials(part, gridDim.x, 4, 1, lds);
        const double mx = max_partials(part, gridDim.x, 4, 2, ldm);
        if (threadIdx.x == 0)
        {
            s->counter[3] = 0;
            s->true_l2 = sqrt(rr);
            s->true_max = mx;
            s->f_rz_old = rz;
            s->f_initial = sqrt(rr);
            s->f_iter = 0;
            s->f_done = 0;
        }
    }

    __global__ __launch_bounds__(kBlock) void k_mx_ap(const uint8_t* __restrict__ nflag,
                                                       const float* __restrict__ pdiagf, const float* __restrict__ p,
                                                       float* __restrict__ ap, int n, double* __restrict__ part,
                                                       DevScalars* s)
    {
        if (s->f_done) return;
        __shared__ double lds[1][kBlock / 64];
        double v[1] = {0.0};
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        {
            float a = 0.0f;
            if (is_free(nflag[i]))
            {
                a = ap[i] + pdiagf[i] * p[i];
                v[0] += double(p[i]) * double(a);
            }
            ap[i] = a;
        }
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) part[blockIdx.x * 4 + 0] = v[0];
        if (!last_block_done(&s->counter[4])) return;
        const double pap = reduce_partials(part, gridDim.x, 4, 0, lds);
        if (threadIdx.x == 0)
        {
            s->counter[4] = 0;
            if (fabs(pap) <= 1.0e-300 || !isfinite(pap))
            {
                s->f_done = 1;  // stagnated inner solve: let the outer FP64 loop judge
                s->f_alpha = 0.0;
                return;
            }
            s->f_alpha = s->f_rz_old / pap;
        }
    }

    __global__ __launch_bounds__(kBlock) void k_mx_update(const uint8_t* __restrict__ nflag,
                                                           const float* __restrict__ dinv,
                                                           const float* __restrict__ p, const float* __restrict__ ap,
                                                           double* __restrict__ dx, float* __restrict__ r, int n,
                                                           double* __restrict__ part, DevScalars* s)
    {
        if (s->f_done) return;
        __shared__ double lds[2][kBlock / 64];
        const double alpha = s->f_alpha;
        const float af = float(alpha);
        double v[2] = {0.0, 0.0};
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        {
            if (!is_free(nflag[i])) continue;
            dx[i] += alpha * double(p[i]);
            const float ri = r[i] - af * ap[i];
            r[i] = ri;
            const float zi = ri * dinv[i];
            v[0] += double(ri) * double(ri);
            v[1] += double(ri) * double(zi);
        }
        block_sum<2>(v, lds);
        if (threadIdx.x == 0)
        {
            part[blockIdx.x * 4 + 0] = v[0];
            part[blockIdx.x * 4 + 1] = v[1];
        }
        if (!last_block_done(&s->counter[5])) return;
        const double rr = reduce_partials(part, gridDim.x, 4, 0, lds);
        const double rz = reduce_partials(part, gridDim.x, 4, 1, lds);
        if (threadIdx.x == 0)
        {
            s->counter[5] = 0;
            s->f_iter += 1;
            s->iter += 1;
            s->f_final = sqrt(rr);
            s->f_rel = s->f_final / s->f_initial;
            if (s->f_rel <= s->f_tol || s->f_iter >= s->f_max_iter || s->iter >= s->max_iter ||
                rz <= 1.0e-300 || !isfinite(rz))
            {
                s->f_done = 1;
                return;
            }
            s->f_beta = rz / s->f_rz_old;
            s->f_rz_old = rz;
        }
    }

    __global__ void k_mx_direction(const uint8_t* __restrict__ nflag, const float* __restrict__ dinv,
                                   const float* __restrict__ r, float* __restrict__ p, float* __restrict__ acc,
                                   int n, const DevScalars* s)
    {
        if (s->f_done) return;
        const float beta = float(s->f_beta);
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        float pi = 0.0f;
        if (is_free(nflag[i])) pi = r[i] * dinv[i] + beta * p[i];
        p[i] = pi;
        acc[i] = 0.0f;
    }

    // x += d on free nodes.
    __global__ void k_mx_correct(const uint8_t* __restrict__ nflag, const double* __restrict__ dx,
                                 double* __restrict__ x, int n)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n && is_free(nflag[i])) x[i] += dx[i];
    }

    // FP32 Jacobi inverse diagonal and FP32 pdiag.
    __global__ void k_mx_prepare(const uint8_t* __restrict__ nflag, const double* __restrict__ pdiag,
                                 float* __restrict__ dinv, float* __restrict__ pdiagf, int n, int precond,
                                 DevScalars* s)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        const double d = pdiag[i];
        float di = 0.0f;
        if (is_free(nflag[i]))
        {
            if (precond == 1 && (!isfinite(d) || d <= 1.0e-30))
                raise_error(s, err_key(ST_CG, 1, uint32_t(i + 1), CK_BAD_DIAG));
            di = (precond == 1) ? float(1.0 / d) : 1.0f;
        }
        dinv[i] = di;
        pdiagf[i] = float(d);
    }

    int solve_pressure_mixed(cbs_hip_ctx* c, const cbs_cg_options* o, cbs_cg_result* res)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const int N = c->N;
        const int rg = reduce_grid(N);
        const double inner_tol = (o && o->inner_rel_tol > 0.0) ? o->inner_rel_tol : 1.0e-4;
        const int max_ref = (o && o->max_refinements > 0) ? o->max_refinements : 50;
        reset_cg_scalars(c);
        c->h_scal->f_tol = inner_tol;
        c->h_scal->f_max_iter = c->cfg.cg_max_iter;
        upload_cg_scalars(c);
        hip_check(hipMemcpyAsync(&c->scal->pending_err, &c->h_scal->pending_err, sizeof(unsigned long long),
                                 hipMemcpyHostToDevice, c->stream),
                  "clear pending");
        if (c->nnonfree > 0)
        {
            hipLaunchKernelGGL(k_constraints, dim3(grid_for(c->nnonfree)), dim3(kBlock), 0, c->stream, c->nf_node,
                               c->nf_mode, c->nf_val, c->nnonfree, c->pres, c->rhs1, c->pdiag);
        }
        float* pdiagf = c->f_z;  // FP32 pdiag shares the z buffer (z itself is never stored)
        hipLaunchKernelGGL(k_mx_prepare, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->nflag, c->pdiag,
                           c->f_dinv, pdiagf, N, c->cfg.cg_preconditioner, c->scal);
        double initial = -1.0;
        int refinements = 0, total_iter = 0, matvecs = 0;
        bool converged = false;
        double final_l2 = 0.0, final_rel = 0.0, final_max = 0.0;
        for (;;)
        {
            // true residual of the current iterate
            launch_fill(c, c->cg_ap, N, 0.0);
            launch_apply(c, c->pres, c->cg_ap, &c->scal->pad0);
            ++matvecs;
            hipLaunchKernelGGL(k_mx_true_residual, dim3(rg), dim3(kBlock), 0, c->stream, c->nflag, c->rhs1,
                               c->pdiag, c->pres, c->cg_ap, c->f_r, c->f_p, c->f_ap, c->f_x, c->f_dinv, N,
                               c->partial, c->scal);
            download_scalars(c);
            if (c->h_scal->err != ~0ull) break;
            const double l2 = c->h_scal->true_l2, mx = c->h_scal->true_max;
            if (initial < 0.0)
            {
                initial = l2;
                if (initial <= 1.0e-300)
                {
                    converged = true;
                    final_l2 = l2;
                    final_rel = 0.0;
                    final_max = mx;
                    break;
                }
            }
            final_l2 = l2;
            final_rel = l2 / initial;
            final_max = mx;
            const bool c1 = mx <= c->cfg.absToler, c2 = final_rel <= c->cfg.relToler;
            const int ct = c->cfg.cg_conv_test;
            if ((ct == 1) ? c1 : (ct == 2) ? c2 : (c1 || c2))
            {
                converged = true;
                break;
            }
            total_iter = c->h_scal->iter;
            if (refinements >= max_ref || total_iter >= c->cfg.cg_max_iter) break;
            ++refinements;
            // inner FP32 PCG on A d = r
            int chunk = 8;
            for (;;)
            {
                for (int k = 0; k < chunk; ++k)
                {
                    timer_begin(c, CBS_PCG_MIXED);
                    launch_apply_f32(c, c->f_p, c->f_ap, &c->scal->f_done);
                    timer_end(c);
                    hipLaunchKernelGGL(k_mx_ap, dim3(rg), dim3(kBlock), 0, c->stream, c->nflag, pdiagf, c->f_p,
                                       c->f_ap, N, c->partial, c->scal);
                    hipLaunchKernelGGL(k_mx_update, dim3(rg), dim3(kBlock), 0, c->stream, c->nflag, c->f_dinv,
                                       c->f_p, c->f_ap, c->f_x, c->f_r, N, c->partial, c->scal);
                    hipLaunchKernelGGL(k_mx_direction, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->nflag,
                                       c->f_dinv, c->f_r, c->f_p, c->f_ap, N, c->scal);
                }
                hip_check(hipGetLastError(), "mixed CG launch");
                download_scalars(c);
                const DevScalars& h = *c->h_scal;
                if (h.f_done || h.err != ~0ull) break;
                int next = 64;
                if (h.f_iter > 0 && h.f_rel > 0.0 && h.f_rel < 1.0)
                {
                    const double rate = std::log(h.f_rel) / h.f_iter;
                    next = int(std::ceil(std::log(inner_tol / h.f_rel) / rate * 0.9));
                }
                chunk = std::max(1, std::min(next, 256));
            }
            matvecs += c->h_scal->f_iter;
            hipLaunchKernelGGL(k_mx_correct, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, c->nflag, c->f_x,
                               c->pres, N);
        }
        res->iterations = c->h_scal->iter;
        res->converged = converged ? 1 : 0;
        res->initial_l2 = initial < 0.0 ? 0.0 : initial;
        res->final_l2 = final_l2;
        res->final_relative_l2 = final_rel;
        res->final_max_abs = final_max;
        res->matvec_calls = matvecs;
        res->dot_calls = 3 * c->h_scal->iter + 2 * (refinements + 1);
        res->preconditioner_calls = c->h_scal->iter + refinements;
        res->refinements = refinements;
        res->algorithmic_bytes = bytes_fp32_iter(c) * c->h_scal->iter + bytes_fp64_iter(c) * (refinements + 1);
        res->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return check_device_error(c);
    }
}
