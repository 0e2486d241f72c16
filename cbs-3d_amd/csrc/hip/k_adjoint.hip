// Discrete adjoint (reverse mode) of the CBS step for a fixed time step.
//
// Forward-model extensions (absent from the reference; with both off the
// step is the reference's, bit for bit, because the extension code paths
// are not taken):
//  - body force f (per node and component, an acceleration): the Step 1
//    predictor becomes u* = u1 + (rhs + m f) elcoe2 with m the node's lumped
//    fluid mass, so that u* gains dt f;
//  - inflow control: the BC-508 profile 6y(1-y) (Boundary.cpp:180-183)
//    plus sum_q c_q phi_q(y, z), phi_(j,k) = sin(j pi yh) sin(k pi zh),
//    j, k = 1..4 (q = 4(j-1) + k-1), yh, zh the inlet's normalised
//    coordinates (bounding box of the 508 nodes).
// Objective: Phi = sum over the last `objective_steps` steps n of
//   phi(u^n) = sum_e nu_e V_e sum_ij (du_i/dx_j)_e^2 over fluid elements.
//
// With dt fixed (dtfixed = 1), the time step, the LHS diagonals and the
// pressure operator do not depend on the state, and one step is
//   u*pre = u^n + (R1(u^n) + m f) w,      u* = BC(u*pre)
//   p = A_ff^-1 (R2(u*, u^n) - A_fc p_c) on free nodes
//   u**pre = u* + R3(p) w,                 u^{n+1} = BC(u**pre)
// (w = elcoe2; BC = symmetry projections, prescribed values, outlet
// backflow clip, solid zeroing; R1 the Step 1 element terms, cubic in u;
// R2 and R3 linear).  The reverse pass applies the transposes in reverse
// order: BC^T (branches of the clip re-evaluated from the stored u**pre /
// u*pre, prescribed entries captured for the inflow gradient), R3^T, the
// adjoint pressure solve A_ff lambda = pbar (the same SPD operator, the
// same PCG), R2^T, the predictor's transpose and R1^T (hand-derived element
// vector-Jacobian products, one launch per colour).  The pressure solve is
// differentiated as an exact solve (implicit function), so the forward and
// adjoint solves use the run's tight relative tolerance.  States are
// checkpointed every K steps and recomputed segment by segment.
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "element_common.hpp"

struct AdjointState
{
    int nsteps = 0, obj_steps = 0, K = 10;
    double rel_tol = 1e-12;
    bool have_forward = false;
    double objective = 0.0;
    std::vector<double*> ck_u, ck_p;       // checkpoint before step k*K+1
    double *u_raw = nullptr;               // initial field before the initial BCs
    double *u_fin = nullptr, *p_fin = nullptr;
    double *t_un = nullptr, *t_us = nullptr, *t_uss = nullptr;  // K slots of 3N each
    double *ubar = nullptr, *usbar = nullptr, *rbar = nullptr, *fbar = nullptr, *u_end = nullptr;
    double *pbar = nullptr, *lam = nullptr, *bc_ubar = nullptr;
    double* part = nullptr;
    double* obj_acc = nullptr;
    std::vector<void*> mem;
};

namespace cbsdev
{
    // ------------------------------------------------------------------
    // Forward-model extensions
    // ------------------------------------------------------------------
    __global__ void k_node_mass(int N, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                                const int32_t* __restrict__ mat, const double* __restrict__ elcoe,
                                double* __restrict__ mass)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double m = 0.0;
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k)
        {
            const int e = inc[k] >> 2, a = inc[k] & 3;
            if (mat[e] == 0) m += elcoe[size_t(e) * 4 + a];
        }
        mass[i] = m;
    }

    // u* = u1 + (rhs + m f) w
    __global__ void k_step1_update_force(int N, const double* __restrict__ unkn1, const double* __restrict__ rhs,
                                         const double* __restrict__ elcoe2, const double* __restrict__ mass,
                                         const double* __restrict__ f, double* __restrict__ unkno)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= N) return;
        const double w = elcoe2[i], m = mass[i];
        for (int d = 0; d < 3; ++d)
        {
            const size_t k = 3 * size_t(i) + d;
            unkno[k] = unkn1[k] + (rhs[k] + m * f[k]) * w;
        }
    }

    void launch_step1_update_forced(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_step1_update_force, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->unkn1,
                           c->rhs, c->elcoe2, c->node_mass, c->bforce, c->unkno);
    }

    // ------------------------------------------------------------------
    // Objective phi(u) = sum_e nu V sum_ij G_ij^2 and its gradient
    // ------------------------------------------------------------------
    struct ViscParams
    {
        double ani;
        int dim_mat;
    };

    __device__ inline double elem_nu(const ViscParams& P, const double* rho, const double* mu, int e)
    {
        return P.dim_mat ? mu[e] / rho[e] : P.ani;
    }

    __device__ inline void elem_grad(const int4& n, const double* g, const double* u, double (&G)[3][3])
    {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                double t = 0.0;
                for (int a = 1; a <= 4; ++a) t += g[j * 4 + a - 1] * u[3 * size_t(conn_at(n, a)) + i];
                G[i][j] = t;
            }
    }

    __global__ __launch_bounds__(kBlock) void k_obj_elem(BlockMap bm, ViscParams P, const int32_t* __restrict__ mat,
                                                          const int4* __restrict__ conn, const double* __restrict__ dN,
                                                          const double* __restrict__ detJ,
                                                          const double* __restrict__ rho, const double* __restrict__ mu,
                                                          const double* __restrict__ u, double* __restrict__ part)
    {
        __shared__ double lds[1][kBlock / 64];
        double v[1] = {0.0};
        for (int e = bm.beg[blockIdx.x] + int(threadIdx.x); e < bm.end[blockIdx.x]; e += kBlock)
        {
            if (mat[e] != 0) continue;
            double G[3][3];
            elem_grad(conn[e], dN + size_t(e) * 12, u, G);
            double s = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) s += G[i][j] * G[i][j];
            v[0] += elem_nu(P, rho, mu, e) * (detJ[e] / 6.0) * s;
        }
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = v[0];
    }

    __global__ __launch_bounds__(kBlock) void k_obj_fin(const double* __restrict__ part, int nb, double* acc)
    {
        __shared__ double lds[1][kBlock / 64];
        const double s = reduce_partials(part, nb, 1, 0, lds);
        if (threadIdx.x == 0) acc[0] += s;
    }

    // ubar += d phi / d u  (per-node gather over the incident fluid elements)
    __global__ void k_obj_grad(int N, ViscParams P, const int32_t* __restrict__ inc_ptr,
                               const int32_t* __restrict__ inc, const int32_t* __restrict__ mat,
                               const int4* __restrict__ conn, const double* __restrict__ dN,
                               const double* __restrict__ detJ, const double* __restrict__ rho,
                               const double* __restrict__ mu, const double* __restrict__ u, double* __restrict__ ubar)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double acc[3] = {0.0, 0.0, 0.0};
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k)
        {
            const int e = inc[k] >> 2, b = (inc[k] & 3) + 1;
            if (mat[e] != 0) continue;
            const double* g = dN + size_t(e) * 12;
            double G[3][3];
            elem_grad(conn[e], g, u, G);
            const double c2 = 2.0 * elem_nu(P, rho, mu, e) * (detJ[e] / 6.0);
            for (int d = 0; d < 3; ++d)
            {
                double t = 0.0;
                for (int j = 0; j < 3; ++j) t += G[d][j] * g[j * 4 + b - 1];
                acc[d] += c2 * t;
            }
        }
        for (int d = 0; d < 3; ++d) ubar[3 * size_t(i) + d] += acc[d];
    }

    // ------------------------------------------------------------------
    // Transposes of the step's pieces
    // ------------------------------------------------------------------

    // BC^T for one application of the velocity programme (which 0: the
    // package of Steps.cpp:132-176, which 1: applyVelocity alone).  The
    // outlet clip's branches are re-evaluated from the pre-BC state u_pre;
    // the adjoint reaching a prescribed entry is added to bc_ubar (the
    // inflow gradient) and cleared.
    __global__ void k_bc_velocity_vjp(int m, int which, const int32_t* __restrict__ node,
                                      const int32_t* __restrict__ flags, const double* __restrict__ val,
                                      const int32_t* __restrict__ sym_ptr, const double* __restrict__ sym_n,
                                      const int32_t* __restrict__ out_ptr, const double* __restrict__ out_n,
                                      const double* __restrict__ u_pre, double* __restrict__ ubar,
                                      double* __restrict__ bc_ubar)
    {
        const int t = blockIdx.x * kBlock + threadIdx.x;
        if (t >= m) return;
        const size_t ip = size_t(node[t]);
        const int f = flags[t];
        // forward replay: branch bits of the outlet clip
        uint32_t active = 0;
        if (which == 0)
        {
            double u = u_pre[3 * ip], v = u_pre[3 * ip + 1], w = u_pre[3 * ip + 2];
            for (int k = sym_ptr[t]; k < sym_ptr[t + 1]; ++k)
            {
                const double nx = sym_n[3 * k], ny = sym_n[3 * k + 1], nz = sym_n[3 * k + 2];
                const double un = u * nx + v * ny + w * nz;
                u -= un * nx;
                v -= un * ny;
                w -= un * nz;
            }
            if (f & BCV_PRESCRIBED)
            {
                u = val[3 * t];
                v = val[3 * t + 1];
                w = val[3 * t + 2];
            }
            for (int k = out_ptr[t], j = 0; k < out_ptr[t + 1]; ++k, ++j)
            {
                if (f & BCV_SOLID)
                {
                    u = v = w = 0.0;
                    continue;
                }
                const double nx = out_n[3 * k], ny = out_n[3 * k + 1], nz = out_n[3 * k + 2];
                const double un = u * nx + v * ny + w * nz;
                if (un < 0.0)
                {
                    active |= 1u << (j & 31);
                    u -= un * nx;
                    v -= un * ny;
                    w -= un * nz;
                }
            }
        }
        // reverse
        double bu = ubar[3 * ip], bv = ubar[3 * ip + 1], bw = ubar[3 * ip + 2];
        if (which == 0)
        {
            if (f & BCV_SOLID) bu = bv = bw = 0.0;
            const int n_out = out_ptr[t + 1] - out_ptr[t];
            for (int j = n_out - 1; j >= 0; --j)
            {
                const int k = out_ptr[t] + j;
                if (f & BCV_SOLID)
                {
                    bu = bv = bw = 0.0;
                    continue;
                }
                if (active & (1u << (j & 31)))
                {
                    const double nx = out_n[3 * k], ny = out_n[3 * k + 1], nz = out_n[3 * k + 2];
                    const double bn = bu * nx + bv * ny + bw * nz;
                    bu -= bn * nx;
                    bv -= bn * ny;
                    bw -= bn * nz;
                }
            }
        }
        if (f & BCV_PRESCRIBED)
        {
            bc_ubar[3 * size_t(t)] += bu;
            bc_ubar[3 * size_t(t) + 1] += bv;
            bc_ubar[3 * size_t(t) + 2] += bw;
            bu = bv = bw = 0.0;
        }
        if (which == 0)
            for (int k = sym_ptr[t + 1] - 1; k >= sym_ptr[t]; --k)
            {
                const double nx = sym_n[3 * k], ny = sym_n[3 * k + 1], nz = sym_n[3 * k + 2];
                const double bn = bu * nx + bv * ny + bw * nz;
                bu -= bn * nx;
                bv -= bn * ny;
                bw -= bn * nz;
            }
        ubar[3 * ip] = bu;
        ubar[3 * ip + 1] = bv;
        ubar[3 * ip + 2] = bw;
    }

    // Step 3 transposed: pbar(a) = -sum_{e ∋ a} vol4 sum_d dN_d(a) S_d(e),
    // S_d(e) = sum over e's nodes of ubar_d * elcoe2 (Steps.cpp:377-428).
    __global__ void k_step3_vjp(int N, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                                const int32_t* __restrict__ mat, const int4* __restrict__ conn,
                                const double* __restrict__ dN, const double* __restrict__ detJ,
                                const double* __restrict__ elcoe2, const double* __restrict__ ubar,
                                const uint8_t* __restrict__ nflag, double* __restrict__ pbar)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double acc = 0.0;
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k)
        {
            const int e = inc[k] >> 2, a = (inc[k] & 3) + 1;
            if (mat[e] != 0) continue;
            const int4 n = conn[e];
            const double* g = dN + size_t(e) * 12;
            double S[3] = {0.0, 0.0, 0.0};
            for (int b = 1; b <= 4; ++b)
            {
                const int ip = conn_at(n, b);
                const double w = elcoe2[ip];
                for (int d = 0; d < 3; ++d) S[d] += ubar[3 * size_t(ip) + d] * w;
            }
            const double vol4 = detJ[e] * FCON1;
            acc -= vol4 * (g[a - 1] * S[0] + g[4 + a - 1] * S[1] + g[8 + a - 1] * S[2]);
        }
        // pressure is a constant on non-free nodes (fixed values, inactive nodes)
        const uint8_t fl = nflag[i];
        pbar[i] = ((fl & 1u) && !(fl & 4u)) ? acc : 0.0;
    }

    // Step 2 RHS transposed (PressureAssembly.cpp:236-335):
    //   usbar_d(i) += sum_{e ∋ i} vol4 sum_a bbar(a) dN_d(a)
    //   unbar_d(i) -= sum_{e ∋ i} sum_{faces f ∋ i, fedge 1} nf_d(f) (sum_{a in f} bbar(a) + bbar(i))
    __global__ void k_step2_vjp(int N, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                                const int32_t* __restrict__ mat, const int4* __restrict__ conn,
                                const double* __restrict__ dN, const double* __restrict__ detJ,
                                const uint32_t* __restrict__ fedge, const double* __restrict__ annxf,
                                const double* __restrict__ bbar, double* __restrict__ usbar,
                                double* __restrict__ unbar)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double vs[3] = {0.0, 0.0, 0.0}, vn[3] = {0.0, 0.0, 0.0};
        const double bi = bbar[i];
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k)
        {
            const int e = inc[k] >> 2, ai = (inc[k] & 3) + 1;
            if (mat[e] != 0) continue;
            const int4 n = conn[e];
            const double* g = dN + size_t(e) * 12;
            const double vol4 = detJ[e] * FCON1;
            for (int d = 0; d < 3; ++d)
            {
                double t = 0.0;
                for (int a = 1; a <= 4; ++a) t += bbar[conn_at(n, a)] * g[d * 4 + a - 1];
                vs[d] += vol4 * t;
            }
            const uint32_t fe = fedge[e];
            for (int is = 1; is <= 4; ++is)
            {
                if (fedge_at(fe, is) != 1 || is == ai) continue;
                double bs = 0.0;
                for (int tt = 1; tt <= 3; ++tt) bs += bbar[conn_at(n, ippn1(is, tt))];
                for (int d = 0; d < 3; ++d)
                {
                    const double nf = annxf[(size_t(e) * 4 + is - 1) * 4 + d] * FCON2;
                    vn[d] -= nf * (bs + bi);
                }
            }
        }
        for (int d = 0; d < 3; ++d)
        {
            usbar[3 * size_t(i) + d] += vs[d];
            unbar[3 * size_t(i) + d] += vn[d];
        }
    }

    // Predictor transposed: rbar = usbar w, unbar += usbar, fbar += rbar m.
    __global__ void k_step1_update_vjp(int N, const double* __restrict__ usbar, const double* __restrict__ elcoe2,
                                       const double* __restrict__ mass, double* __restrict__ rbar,
                                       double* __restrict__ unbar, double* __restrict__ fbar)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= N) return;
        const double w = elcoe2[i], m = mass[i];
        for (int d = 0; d < 3; ++d)
        {
            const size_t k = 3 * size_t(i) + d;
            const double r = usbar[k] * w;
            rbar[k] = r;
            unbar[k] += usbar[k];
            if (fbar) fbar[k] += r * m;
        }
    }

    // Step 1 element terms transposed (MomentumAssembly.cpp:144-432): for
    // each element, dU = J_e^T W with W the element's rbar, added to unbar;
    // one launch per colour, so no two threads add to one node.
    __global__ __launch_bounds__(kBlock) void k_step1_vjp_color(int beg, int end, ViscParams P,
                                                                 const int32_t* __restrict__ slot_elem,
                                                                 const int4* __restrict__ conn,
                                                                 const double* __restrict__ dNg,
                                                                 const double* __restrict__ detJ,
                                                                 const double* __restrict__ delte,
                                                                 const uint32_t* __restrict__ fedge,
                                                                 const double* __restrict__ annxf,
                                                                 const double* __restrict__ rho,
                                                                 const double* __restrict__ mu,
                                                                 const double* __restrict__ un,
                                                                 const double* __restrict__ rbar,
                                                                 double* __restrict__ unbar)
    {
        const int k = beg + int(blockIdx.x) * kBlock + int(threadIdx.x);
        if (k >= end) return;
        const int e = slot_elem[k];
        const int4 n = conn[e];
        const double* g = dNg + size_t(e) * 12;
#define DN(d, a) g[((d) - 1) * 4 + (a) - 1]
        const double* an = annxf + size_t(e) * 16;
#define ANN(cc, is) an[((is) - 1) * 4 + (cc) - 1]
        const uint32_t fe = fedge[e];
        const double dj = detJ[e], dt = delte[e];
        const double nu = P.dim_mat ? mu[e] / rho[e] : P.ani;
        double U[4][5], W[4][5], dU[4][5];
#pragma unroll
        for (int a = 1; a <= 4; ++a)
        {
            const size_t ip = size_t(conn_at(n, a));
#pragma unroll
            for (int i = 1; i <= 3; ++i)
            {
                U[i][a] = un[3 * ip + i - 1];
                W[i][a] = rbar[3 * ip + i - 1];
                dU[i][a] = 0.0;
            }
        }
        // A[i][j] = sum_a W[i][a] dN_j(a); B[i][j] = sum_faces ANN(j,f) Wf[i]
        double A[4][4], B[4][4];
#pragma unroll
        for (int i = 1; i <= 3; ++i)
#pragma unroll
            for (int j = 1; j <= 3; ++j)
            {
                double t = 0.0;
#pragma unroll
                for (int a = 1; a <= 4; ++a) t += W[i][a] * DN(j, a);
                A[i][j] = t;
                B[i][j] = 0.0;
            }
        const double vf1 = dj * FCON1;
#pragma unroll
        for (int is = 1; is <= 4; ++is)
        {
            if (fedge_at(fe, is) == 0) continue;
            double Wf[4];
#pragma unroll
            for (int i = 1; i <= 3; ++i)
                Wf[i] = W[i][ippn1(is, 1)] + W[i][ippn1(is, 2)] + W[i][ippn1(is, 3)];
#pragma unroll
            for (int i = 1; i <= 3; ++i)
#pragma unroll
                for (int j = 1; j <= 3; ++j) B[i][j] += ANN(j, is) * Wf[i];
            // Galerkin face term: -sum_{b in f} sum_ij c_j (Wf_i + W_ib) U_jb U_ib
#pragma unroll
            for (int fn = 1; fn <= 3; ++fn)
            {
                const int b = ippn1(is, fn);
                double QU = 0.0, cU = 0.0;
#pragma unroll
                for (int i = 1; i <= 3; ++i)
                {
                    QU += (Wf[i] + W[i][b]) * U[i][b];
                    cU += ANN(i, is) * FCON2 * U[i][b];
                }
#pragma unroll
                for (int m = 1; m <= 3; ++m)
                    dU[m][b] -= ANN(m, is) * FCON2 * QU + (Wf[m] + W[m][b]) * cU;
            }
        }
        // Galerkin volume term
#pragma unroll
        for (int b = 1; b <= 4; ++b)
#pragma unroll
            for (int m = 1; m <= 3; ++m)
            {
                double t = 0.0;
#pragma unroll
                for (int i = 1; i <= 3; ++i) t += A[i][m] * U[i][b] + A[m][i] * U[i][b];
                dU[m][b] += vf1 * t;
            }
        // characteristic correction: C2 = -ldelte A + ldelte2 B
        {
            const double ldelte = 0.5 * dt * dj * FCON3, ldelte2 = 0.5 * dt * FCON4;
            double umean[4], div[5], sumj[4], D[4], E[4];
#pragma unroll
            for (int kk = 1; kk <= 3; ++kk) umean[kk] = (U[kk][1] + U[kk][2] + U[kk][3] + U[kk][4]) / 4.0;
#pragma unroll
            for (int a = 1; a <= 4; ++a) div[a] = DN(1, a) * U[1][a] + DN(2, a) * U[2][a] + DN(3, a) * U[3][a];
#pragma unroll
            for (int i = 1; i <= 3; ++i)
            {
                double t = 0.0;
#pragma unroll
                for (int a = 1; a <= 4; ++a) t += U[i][a] * div[a];
                sumj[i] = t;
            }
#pragma unroll
            for (int i = 1; i <= 3; ++i)
            {
                double t = 0.0;
#pragma unroll
                for (int kk = 1; kk <= 3; ++kk) t += (-ldelte * A[i][kk] + ldelte2 * B[i][kk]) * umean[kk];
                D[i] = t;
            }
#pragma unroll
            for (int kk = 1; kk <= 3; ++kk)
            {
                double t = 0.0;
#pragma unroll
                for (int i = 1; i <= 3; ++i) t += (-ldelte * A[i][kk] + ldelte2 * B[i][kk]) * sumj[i];
                E[kk] = t;
            }
#pragma unroll
            for (int b = 1; b <= 4; ++b)
            {
                double DU = 0.0;
#pragma unroll
                for (int i = 1; i <= 3; ++i) DU += D[i] * U[i][b];
#pragma unroll
                for (int m = 1; m <= 3; ++m) dU[m][b] += E[m] / 4.0 + DN(m, b) * DU + D[m] * div[b];
            }
        }
        // diffusion: C3 = -nu dj FDIF1 A + nu FDIF2 B
        {
            const double vfd = nu * dj * FDIF1, fdf = nu * FDIF2;
#pragma unroll
            for (int b = 1; b <= 4; ++b)
#pragma unroll
                for (int m = 1; m <= 3; ++m)
                {
                    double t = 0.0;
#pragma unroll
                    for (int j = 1; j <= 3; ++j) t += (-vfd * A[m][j] + fdf * B[m][j]) * DN(j, b);
                    dU[m][b] += t;
                }
        }
#undef DN
#undef ANN
#pragma unroll
        for (int a = 1; a <= 4; ++a)
        {
            const size_t ip = size_t(conn_at(n, a));
#pragma unroll
            for (int i = 1; i <= 3; ++i) unbar[3 * ip + i - 1] += dU[i][a];
        }
    }

    // ------------------------------------------------------------------
    // Host driver
    // ------------------------------------------------------------------
    namespace
    {
        double* adj_alloc(AdjointState* a, size_t n)
        {
            void* p = nullptr;
            hip_check(hipMalloc(&p, (n ? n : 1) * sizeof(double)), "hipMalloc adjoint");
            a->mem.push_back(p);
            return static_cast<double*>(p);
        }

        void dcopy(cbs_hip_ctx* c, double* dst, const double* src, size_t n)
        {
            hip_check(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream), "d2d");
        }

        ViscParams visc(const cbs_hip_ctx* c)
        {
            return {c->cfg.ani, (c->cfg.dimensional_mode > 0 && c->cfg.material_properties_enabled > 0) ? 1 : 0};
        }

        // Runs the pressure solve of the current rhs1/pres at the run's
        // tolerance; throws when it does not converge.
        void tight_solve(cbs_hip_ctx* c, double rel_tol)
        {
            const double saved = c->cfg.relToler;
            const int saved_ct = c->cfg.cg_conv_test;
            c->cfg.relToler = rel_tol;
            c->cfg.cg_conv_test = 2;
            cbs_cg_result r{};
            int rc = CBS_OK;
            try
            {
                rc = solve_pressure_fp64(c, &r);
            }
            catch (...)
            {
                c->cfg.relToler = saved;
                c->cfg.cg_conv_test = saved_ct;
                throw;
            }
            c->cfg.relToler = saved;
            c->cfg.cg_conv_test = saved_ct;
            if (rc != CBS_OK) throw std::runtime_error(c->err);
            if (!r.converged)
                throw std::runtime_error("adjoint run: pressure solve did not reach the relative tolerance " +
                                         std::to_string(rel_tol));
        }

        void objective_add(cbs_hip_ctx* c, AdjointState* a, const double* u)
        {
            hipLaunchKernelGGL(k_obj_elem, dim3(c->eblk.n), dim3(kBlock), 0, c->stream, c->eblk.map(), visc(c),
                               c->mat, c->conn, c->dNkdx, c->detJ, c->rho, c->mu, u, a->part);
            hipLaunchKernelGGL(k_obj_fin, dim3(1), dim3(kBlock), 0, c->stream, a->part, c->eblk.n, a->obj_acc);
        }

        void bc_vjp(cbs_hip_ctx* c, AdjointState* a, int which, const double* u_pre, double* ubar)
        {
            if (c->nbcv > 0)
                hipLaunchKernelGGL(k_bc_velocity_vjp, dim3(grid_for(c->nbcv)), dim3(kBlock), 0, c->stream, c->nbcv,
                                   which, c->bcv_node, c->bcv_flags, c->bcv_val, c->bcv_sym_ptr, c->bcv_sym_n,
                                   c->bcv_out_ptr, c->bcv_out_n, u_pre, ubar, a->bc_ubar);
        }

        // One CBS step of the adjoint workload; with slot >= 0 the step's
        // input, predictor and corrector states are stored in that tape slot.
        void forward_step(cbs_hip_ctx* c, AdjointState* a, int iitime, int slot)
        {
            const size_t N = size_t(c->N), n3 = 3 * N;
            dcopy(c, c->unkn1, c->unkno, n3);
            dcopy(c, c->pres1, c->pres, N);
            dcopy(c, c->temp1, c->temp, N);
            if (slot >= 0) dcopy(c, a->t_un + size_t(slot) * n3, c->unkno, n3);
            launch_timestep(c, iitime);
            launch_lhs(c);
            launch_step1_rhs(c);
            launch_step1_update(c);
            if (slot >= 0) dcopy(c, a->t_us + size_t(slot) * n3, c->unkno, n3);
            launch_bc(c, 0);
            launch_step2_rhs(c);
            tight_solve(c, a->rel_tol);
            launch_bc(c, 3);
            launch_step3(c);
            if (slot >= 0) dcopy(c, a->t_uss + size_t(slot) * n3, c->unkno, n3);
            launch_bc(c, 0);
            if (c->cfg.temp_calc >= 1)
            {
                launch_step4_rhs(c);
                launch_step4_update(c);
                launch_bc(c, 2);
            }
            launch_velocity_magnitude(c);
        }
    }

    void destroy_adjoint(cbs_hip_ctx* c)
    {
        if (!c->adj) return;
        for (void* p : c->adj->mem) (void)hipFree(p);
        delete c->adj;
        c->adj = nullptr;
    }
}

using namespace cbsdev;

namespace
{
    struct Fail : std::runtime_error
    {
        int status;
        Fail(int s, const std::string& m) : std::runtime_error(m), status(s) {}
    };

    template <typename F>
    int guarded_adj(cbs_hip_ctx* c, F&& fn)
    {
        if (!c) return CBS_ERR_BAD_ARGUMENT;
        try
        {
            return fn();
        }
        catch (const Fail& e)
        {
            c->err = e.what();
            return e.status;
        }
        catch (const std::exception& e)
        {
            c->err = e.what();
            return CBS_ERR_HIP;
        }
    }

    void ensure_mass(cbs_hip_ctx* c)
    {
        if (c->node_mass) return;
        void* p = nullptr;
        hip_check(hipMalloc(&p, sizeof(double) * size_t(c->N)), "hipMalloc node mass");
        c->allocations.push_back(p);
        c->node_mass = static_cast<double*>(p);
        hipLaunchKernelGGL(k_node_mass, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->inc_ptr, c->inc,
                           c->mat, c->elcoe, c->node_mass);
    }

    // phi_q at every inflow slot (16 per slot), on the inlet's bounding box
    std::vector<double> inflow_basis(const cbs_hip_ctx* c)
    {
        const size_t ns = c->h_inflow_slot.size();
        std::vector<double> phi(ns * 16, 0.0);
        if (ns == 0) return phi;
        double y0 = 1e300, y1 = -1e300, z0 = 1e300, z1 = -1e300;
        for (size_t s = 0; s < ns; ++s)
        {
            y0 = std::min(y0, c->h_inflow_yz[2 * s]);
            y1 = std::max(y1, c->h_inflow_yz[2 * s]);
            z0 = std::min(z0, c->h_inflow_yz[2 * s + 1]);
            z1 = std::max(z1, c->h_inflow_yz[2 * s + 1]);
        }
        const double pi = 3.14159265358979323846;
        for (size_t s = 0; s < ns; ++s)
        {
            const double yh = (y1 > y0) ? (c->h_inflow_yz[2 * s] - y0) / (y1 - y0) : 0.0;
            const double zh = (z1 > z0) ? (c->h_inflow_yz[2 * s + 1] - z0) / (z1 - z0) : 0.0;
            for (int j = 1; j <= 4; ++j)
                for (int k = 1; k <= 4; ++k) phi[s * 16 + size_t((j - 1) * 4 + k - 1)] = std::sin(j * pi * yh) * std::sin(k * pi * zh);
        }
        return phi;
    }
}

extern "C" {

int cbs_hip_set_body_force(cbs_hip_ctx* c, const double* f)
{
    return guarded_adj(c, [&] {
        if (c->nseg > 1) throw Fail(CBS_ERR_UNSUPPORTED, "body force: single-case contexts only");
        if (!f)
        {
            c->bforce = nullptr;
            return int(CBS_OK);
        }
        ensure_mass(c);
        if (!c->bforce_buf)
        {
            void* p = nullptr;
            hip_check(hipMalloc(&p, sizeof(double) * 3 * size_t(c->N)), "hipMalloc body force");
            c->allocations.push_back(p);
            c->bforce_buf = static_cast<double*>(p);
        }
        hip_check(hipMemcpyAsync(c->bforce_buf, f, sizeof(double) * 3 * size_t(c->N), hipMemcpyHostToDevice,
                                 c->stream),
                  "force");
        hip_check(hipStreamSynchronize(c->stream), "force sync");
        c->bforce = c->bforce_buf;
        return int(CBS_OK);
    });
}

int cbs_hip_set_inflow_coefficients(cbs_hip_ctx* c, const double* coef, int32_t n)
{
    return guarded_adj(c, [&] {
        if (c->nseg > 1) throw Fail(CBS_ERR_UNSUPPORTED, "inflow coefficients: single-case contexts only");
        if (n != 0 && (n != 16 || !coef))
            throw Fail(CBS_ERR_BAD_ARGUMENT, "cbs_hip_set_inflow_coefficients - pass 16 coefficients, or 0 to clear");
        std::vector<double> val = c->h_bcv_val;
        if (n == 16)
        {
            if (c->h_inflow_slot.empty())
                throw Fail(CBS_ERR_BAD_ARGUMENT, "cbs_hip_set_inflow_coefficients - the case has no BC-508 inflow");
            const std::vector<double> phi = inflow_basis(c);
            for (size_t s = 0; s < c->h_inflow_slot.size(); ++s)
            {
                double add = 0.0;
                for (int q = 0; q < 16; ++q) add += coef[q] * phi[s * 16 + size_t(q)];
                val[3 * size_t(c->h_inflow_slot[s])] += add;
            }
            c->h_inflow_coef.assign(coef, coef + 16);
        }
        else
        {
            c->h_inflow_coef.clear();
        }
        if (c->nbcv > 0)
        {
            hip_check(hipMemcpyAsync(c->bcv_val, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice,
                                     c->stream),
                      "inflow values");
            hip_check(hipStreamSynchronize(c->stream), "inflow sync");
        }
        return int(CBS_OK);
    });
}

int cbs_hip_inflow_basis(cbs_hip_ctx* c, int32_t* nslots, int32_t* node, double* phi)
{
    return guarded_adj(c, [&] {
        if (!nslots) throw Fail(CBS_ERR_BAD_ARGUMENT, "null output");
        *nslots = int32_t(c->h_inflow_slot.size());
        if (node || phi)
        {
            std::vector<int32_t> bn(size_t(c->nbcv));
            if (c->nbcv > 0)
                hip_check(hipMemcpy(bn.data(), c->bcv_node, sizeof(int32_t) * bn.size(), hipMemcpyDeviceToHost),
                          "bc nodes");
            const std::vector<double> ph = inflow_basis(c);
            for (size_t s = 0; s < c->h_inflow_slot.size(); ++s)
            {
                if (node) node[s] = bn[size_t(c->h_inflow_slot[s])] + 1;
                if (phi)
                    for (int q = 0; q < 16; ++q) phi[s * 16 + size_t(q)] = ph[s * 16 + size_t(q)];
            }
        }
        return int(CBS_OK);
    });
}

int cbs_hip_adjoint_forward(cbs_hip_ctx* c, const cbs_adjoint_options* o, double* objective)
{
    return guarded_adj(c, [&] {
        if (!o || !objective) throw Fail(CBS_ERR_BAD_ARGUMENT, "cbs_hip_adjoint_forward - null argument");
        if (c->nseg > 1) throw Fail(CBS_ERR_UNSUPPORTED, "adjoint: single-case contexts only");
        const cbs_config& g = c->cfg;
        if (!(g.dtfixed > 0 || g.ilots <= -1))
            throw Fail(CBS_ERR_UNSUPPORTED,
                       "adjoint: needs a fixed time step (dtfixed = 1); the minimum over elements of the stability "
                       "time step is not differentiable");
        if (g.cbs_scheme != 1) throw Fail(CBS_ERR_UNSUPPORTED, "adjoint: semi-implicit scheme only");
        if (o->nsteps < 1 || o->objective_steps < 0 || o->objective_steps > o->nsteps || o->checkpoint_every < 0 ||
            o->rel_tol < 0.0)
            throw Fail(CBS_ERR_BAD_ARGUMENT, "cbs_hip_adjoint_forward - invalid options");
        ensure_mass(c);
        destroy_adjoint(c);
        c->adj = new AdjointState();
        AdjointState* a = c->adj;
        a->nsteps = o->nsteps;
        a->obj_steps = o->objective_steps;
        a->K = o->checkpoint_every > 0 ? o->checkpoint_every : 10;
        a->rel_tol = o->rel_tol > 0.0 ? o->rel_tol : 1.0e-12;
        const size_t N = size_t(c->N), n3 = 3 * N;
        const int nck = (a->nsteps + a->K - 1) / a->K;
        for (int k = 0; k < nck; ++k)
        {
            a->ck_u.push_back(adj_alloc(a, n3));
            a->ck_p.push_back(adj_alloc(a, N));
        }
        a->u_raw = adj_alloc(a, n3);
        a->u_fin = adj_alloc(a, n3);
        a->p_fin = adj_alloc(a, N);
        a->t_un = adj_alloc(a, size_t(a->K) * n3);
        a->t_us = adj_alloc(a, size_t(a->K) * n3);
        a->t_uss = adj_alloc(a, size_t(a->K) * n3);
        for (double** p : {&a->ubar, &a->usbar, &a->rbar, &a->fbar, &a->u_end}) *p = adj_alloc(a, n3);
        a->pbar = adj_alloc(a, N);
        a->lam = adj_alloc(a, N);
        a->bc_ubar = adj_alloc(a, 3 * size_t(std::max(c->nbcv, 1)));
        a->part = adj_alloc(a, size_t(std::max(c->eblk.n, 1)));
        a->obj_acc = adj_alloc(a, 1);
        hip_check(hipMemsetAsync(a->obj_acc, 0, sizeof(double), c->stream), "objective");
        // Solver::initialise's boundary application (Solver.cpp:296-300):
        // the inflow controls set the initial inlet values as well
        dcopy(c, a->u_raw, c->unkno, n3);
        launch_bc(c, 2);
        launch_bc(c, 1);
        launch_bc(c, 3);
        launch_velocity_magnitude(c);
        for (int it = 1; it <= a->nsteps; ++it)
        {
            if ((it - 1) % a->K == 0)
            {
                dcopy(c, a->ck_u[size_t((it - 1) / a->K)], c->unkno, n3);
                dcopy(c, a->ck_p[size_t((it - 1) / a->K)], c->pres, N);
            }
            forward_step(c, a, it, -1);
            if (it > a->nsteps - a->obj_steps) objective_add(c, a, c->unkno);
        }
        dcopy(c, a->u_fin, c->unkno, n3);
        dcopy(c, a->p_fin, c->pres, N);
        hip_check(hipMemcpyAsync(&a->objective, a->obj_acc, sizeof(double), hipMemcpyDeviceToHost, c->stream), "obj");
        hip_check(hipStreamSynchronize(c->stream), "forward sync");
        const int rc = check_device_error(c);
        if (rc != CBS_OK) return rc;
        a->have_forward = true;
        *objective = a->objective;
        return int(CBS_OK);
    });
}

int cbs_hip_adjoint_backward(cbs_hip_ctx* c, double seed, double* grad_force, double* grad_inflow)
{
    return guarded_adj(c, [&] {
        AdjointState* a = c->adj;
        if (!a || !a->have_forward)
            throw Fail(CBS_ERR_BAD_ARGUMENT, "cbs_hip_adjoint_backward - run cbs_hip_adjoint_forward first");
        const size_t N = size_t(c->N), n3 = 3 * N;
        const ViscParams P = visc(c);
        for (double* p : {a->ubar, a->fbar})
            hip_check(hipMemsetAsync(p, 0, n3 * sizeof(double), c->stream), "clear");
        hip_check(hipMemsetAsync(a->bc_ubar, 0, 3 * size_t(std::max(c->nbcv, 1)) * sizeof(double), c->stream),
                  "clear");
        const int nck = int(a->ck_u.size());
        for (int seg = nck - 1; seg >= 0; --seg)
        {
            const int n0 = seg * a->K + 1, n1 = std::min(a->nsteps, (seg + 1) * a->K);
            dcopy(c, c->unkno, a->ck_u[size_t(seg)], n3);
            dcopy(c, c->pres, a->ck_p[size_t(seg)], N);
            for (int it = n0; it <= n1; ++it) forward_step(c, a, it, it - n0);
            dcopy(c, a->u_end, c->unkno, n3);
            for (int it = n1; it >= n0; --it)
            {
                const int slot = it - n0;
                const double* u_after = (it == n1) ? a->u_end : a->t_un + size_t(slot + 1) * n3;
                if (it > a->nsteps - a->obj_steps)
                    hipLaunchKernelGGL(k_obj_grad, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, P,
                                       c->inc_ptr, c->inc, c->mat, c->conn, c->dNkdx, c->detJ, c->rho, c->mu, u_after,
                                       a->ubar);
                // u^{n+1} = BC(u**pre): transpose of the second velocity package
                bc_vjp(c, a, 0, a->t_uss + size_t(slot) * n3, a->ubar);
                // u**pre = u* + R3(p) w
                hipLaunchKernelGGL(k_step3_vjp, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->inc_ptr,
                                   c->inc, c->mat, c->conn, c->dNkdx, c->detJ, c->elcoe2, a->ubar, c->nflag, a->pbar);
                // p = A_ff^-1 b: the adjoint solve A_ff lambda = pbar
                hip_check(hipMemsetAsync(a->lam, 0, N * sizeof(double), c->stream), "lambda");
                std::swap(c->pres, a->lam);
                std::swap(c->rhs1, a->pbar);
                c->constraint_zero = true;
                try
                {
                    tight_solve(c, a->rel_tol);
                }
                catch (...)
                {
                    c->constraint_zero = false;
                    std::swap(c->pres, a->lam);
                    std::swap(c->rhs1, a->pbar);
                    throw;
                }
                c->constraint_zero = false;
                std::swap(c->pres, a->lam);
                std::swap(c->rhs1, a->pbar);
                // b = R2(u*, u^n)
                dcopy(c, a->usbar, a->ubar, n3);
                hip_check(hipMemsetAsync(a->ubar, 0, n3 * sizeof(double), c->stream), "clear");
                hipLaunchKernelGGL(k_step2_vjp, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->inc_ptr,
                                   c->inc, c->mat, c->conn, c->dNkdx, c->detJ, c->fedge, c->annxf, a->lam, a->usbar,
                                   a->ubar);
                // u* = BC(u*pre)
                bc_vjp(c, a, 0, a->t_us + size_t(slot) * n3, a->usbar);
                // u*pre = u^n + (R1(u^n) + m f) w
                hipLaunchKernelGGL(k_step1_update_vjp, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N,
                                   a->usbar, c->elcoe2, c->node_mass, a->rbar, a->ubar, a->fbar);
                for (int col = 0; col < c->ncolor; ++col)
                {
                    const int beg = c->h_fcolor_ptr[size_t(col)], end = c->h_fcolor_ptr[size_t(col) + 1];
                    if (end <= beg) continue;
                    hipLaunchKernelGGL(k_step1_vjp_color, dim3(grid_for(end - beg)), dim3(kBlock), 0, c->stream, beg,
                                       end, P, c->slot_elem, c->conn, c->dNkdx, c->detJ, c->delte, c->fedge, c->annxf,
                                       c->rho, c->mu, a->t_un + size_t(slot) * n3, a->rbar, a->ubar);
                }
            }
        }
        // u^0 = applyVelocity(u_raw)
        bc_vjp(c, a, 1, a->u_raw, a->ubar);
        std::vector<double> fb(n3), bcu(3 * size_t(std::max(c->nbcv, 1)));
        hip_check(hipMemcpyAsync(fb.data(), a->fbar, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "fbar");
        hip_check(hipMemcpyAsync(bcu.data(), a->bc_ubar, bcu.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                  "bc_ubar");
        // leave the context at the end of the forward run
        dcopy(c, c->unkno, a->u_fin, n3);
        dcopy(c, c->pres, a->p_fin, N);
        hip_check(hipStreamSynchronize(c->stream), "backward sync");
        const int rc = check_device_error(c);
        if (rc != CBS_OK) return rc;
        if (grad_force)
            for (size_t k = 0; k < n3; ++k) grad_force[k] = seed * fb[k];
        if (grad_inflow)
        {
            const std::vector<double> phi = inflow_basis(c);
            for (int q = 0; q < 16; ++q)
            {
                double s = 0.0;
                for (size_t sl = 0; sl < c->h_inflow_slot.size(); ++sl)
                    s += phi[sl * 16 + size_t(q)] * bcu[3 * size_t(c->h_inflow_slot[sl])];
                grad_inflow[q] = seed * s;
            }
        }
        return int(CBS_OK);
    });
}

}  // extern "C"
