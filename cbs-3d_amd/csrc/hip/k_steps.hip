// CBS step kernels: element pressure terms, time step, LHS diagonals,
// Steps 1-4, boundary programmes, velocity magnitude and convergence norms.
//
// Ordering rules that make every stage bit-identical to the reference:
//  - Step 1 is coloured in the reference (MomentumAssembly.cpp:476-543), so
//    it runs as one launch per colour over the same colour classes: every
//    node receives its element terms in colour order.
//  - The serial element loops (Step 2 RHS, Step 3, Step 4, the LHS sums)
//    run as per-node gathers over incidence lists sorted by element, so every
//    node sums its element terms in ascending element order, exactly like
//    the reference's loop; each gather recomputes the element quantities it
//    needs with the reference's operation order.
//  - Boundary conditions are per-node programmes built once on the host in
//    the reference's face order (last writer wins, projections in sequence).
// Every kernel is compiled with -ffp-contract=off; no expression is fused.
#include <cfloat>
#include <cmath>

#include "element_common.hpp"

namespace cbsdev
{
    // ------------------------------------------------------------------
    // PressureAssembly::buildElementPressureTerms (PressureAssembly.cpp:136-178)
    // ------------------------------------------------------------------
    __global__ void k_build_pterms(int E, int64_t F, const int32_t* __restrict__ mat,
                                   const double* __restrict__ dN, const double* __restrict__ detJ,
                                   const int32_t* __restrict__ elem_slot, const int32_t* __restrict__ elem_tslot,
                                   double* __restrict__ pdiagE, double* __restrict__ gstifE,
                                   double* __restrict__ gE_s, double* __restrict__ gE_t, DevScalars* s)
    {
        const int e = blockIdx.x * kBlock + threadIdx.x;
        if (e >= E) return;
        double* pd = pdiagE + size_t(e) * 4;
        double* gs = gstifE + size_t(e) * 6;
        if (mat[e] != 0)
        {
            for (int i = 0; i < 4; ++i) pd[i] = 0.0;
            for (int i = 0; i < 6; ++i) gs[i] = 0.0;
            return;
        }
        const double dj = detJ[e];
        if (dj <= 0.0 || !isfinite(dj))
        {
            raise_error(s, err_key(ST_PTERMS, 1, uint32_t(e + 1), 1));
            return;
        }
        const double* g = dN + size_t(e) * 12;
        auto gd = [&](int a, int b) { return g[a] * g[b] + g[4 + a] * g[4 + b] + g[8 + a] * g[8 + b]; };
        const double volume = dj / 6.0;
        for (int a = 0; a < 4; ++a) pd[a] = volume * gd(a, a);
        const double v[6] = {volume * gd(0, 1), volume * gd(0, 2), volume * gd(0, 3),
                             volume * gd(1, 2), volume * gd(1, 3), volume * gd(2, 3)};
        const int slot = elem_slot[e], tslot = elem_tslot[e];
        for (int p = 0; p < 6; ++p)
        {
            gs[p] = v[p];
            gE_s[p * F + slot] = v[p];
            gE_t[p * F + tslot] = v[p];
        }
    }

    void launch_build_pressure_terms(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_build_pterms, dim3(grid_for(c->E)), dim3(kBlock), 0, c->stream, c->E, int64_t(c->F),
                           c->mat, c->dNkdx, c->detJ, c->elem_slot, c->elem_tslot, c->pdiagE, c->gstifE,
                           c->gstifE_s, c->gstifE_t, c->scal);
        c->sell.base_stale = true;
    }

    // ------------------------------------------------------------------
    // Nodal lengths (TimeStep.cpp:319-355).  Geometry is static, so the
    // minimum over incident elements is computed once; min is exact.
    // ------------------------------------------------------------------
    __global__ void k_static_alen(int N, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                                  const double* __restrict__ detJ, const double* __restrict__ annxf,
                                  double* __restrict__ alen)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double h = DBL_MAX;
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k)
        {
            const int e = inc[k] >> 2, a = inc[k] & 3;
            const double volum = 0.5 * detJ[e];
            const double hite = volum / annxf[(size_t(e) * 4 + a) * 4 + 3];
            h = dmin(h, hite);
        }
        alen[i] = h;
    }

    void launch_static_alen(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_static_alen, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->inc_ptr,
                           c->inc, c->detJ, c->annxf, c->alen_static);
    }

    // ------------------------------------------------------------------
    // TimeStep::computeTimeStep (TimeStep.cpp:473-656)
    // ------------------------------------------------------------------
    struct TsParams
    {
        double csafm, ani, deltr, dtfix;
        int ilots, htype, beta_opt, rem_deltp, global, dim_mat;
    };

    __device__ inline double node_speed(const double* u, int i)
    {
        const double a = u[3 * size_t(i)], b = u[3 * size_t(i) + 1], c = u[3 * size_t(i) + 2];
        return sqrt(a * a + b * b + c * c);
    }

    // seg_of_node / seg_ani: per-member ani of a batch context (null otherwise).
    __global__ void k_ts_nodal(int N, TsParams P, const double* __restrict__ unkno,
                               const double* __restrict__ alen_static, double* __restrict__ alen,
                               double* __restrict__ velcp, double* __restrict__ unit_vec, double* __restrict__ deltp,
                               double* __restrict__ deltp2, double* __restrict__ beta1, double* __restrict__ vvis,
                               const uint16_t* __restrict__ seg_of_node, const double* __restrict__ seg_ani,
                               DevScalars* s)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= N) return;
        const double ani = seg_ani ? seg_ani[seg_of_node[i]] : P.ani;
        // compute_velocity_helpers (:364-384)
        const double speed = node_speed(unkno, i);
        velcp[i] = speed;
        for (int d = 0; d < 3; ++d) unit_vec[3 * size_t(i) + d] = (speed > 1.0e-14) ? unkno[3 * size_t(i) + d] / speed : 0.0;
        // compute_nodal_lengths_from_elements (:516; geometry only, so cached)
        const double al = alen_static[i];
        alen[i] = al;
        // computeNodalLocalTimeStep (:600-656)
        const double h = (al > 0.0 && isfinite(al)) ? al : 0.0;
        const double sp = dmax(speed, 1.0e-14);
        if (h <= 0.0)
        {
            deltp[i] = P.dtfix;
            deltp2[i] = P.dtfix;
            beta1[i] = 1.0;
            vvis[i] = 0.0;
            return;
        }
        const double adv = h / sp;
        const double dif = h * h * (0.5 / ani);
        vvis[i] = dif;
        double dt = P.csafm * dmin(adv, dif);
        if (P.ilots == 2 && P.deltr > 0.0 && isfinite(P.deltr)) dt = dmin(dt, P.deltr);
        if (dt <= 0.0 || !isfinite(dt)) raise_error(s, err_key(ST_TIMESTEP, 1, uint32_t(i + 1), 1));
        deltp[i] = dt;
        deltp2[i] = dt;
        if (P.beta_opt > 0)
        {
            const double denom = adv + dif + 1.0e-30;
            const double b = dif / denom;
            beta1[i] = (b < 0.0) ? 0.0 : ((1.0 < b) ? 1.0 : b);
        }
        else
        {
            beta1[i] = 1.0;
        }
    }

    __global__ __launch_bounds__(kBlock) void k_ts_elem(BlockMap bm, TsParams P, const int4* __restrict__ conn,
                                                         const int32_t* __restrict__ mat,
                                                         const double* __restrict__ alen_e,
                                                         const double* __restrict__ alen,
                                                         const double* __restrict__ unkno,
                                                         const double* __restrict__ rho,
                                                         const double* __restrict__ mu,
                                                         const double* __restrict__ alpha,
                                                         double* __restrict__ delte, double* __restrict__ part,
                                                         const double* __restrict__ seg_ani, DevScalars* s)
    {
        __shared__ double ldm[kBlock / 64];
        double m = DBL_MAX;
        const int e0 = bm.beg[blockIdx.x], e1 = bm.end[blockIdx.x];
        const double seg_diff = seg_ani ? seg_ani[bm.seg[blockIdx.x]] : P.ani;
        for (int e = e0 + int(threadIdx.x); e < e1; e += kBlock)
        {
            const int4 n = conn[e];
            const bool fluid = mat[e] == 0;
            // timestep_length (:127-147)
            double h = alen_e[e];
            if (h <= 0.0 || !isfinite(h))
            {
                raise_error(s, err_key(ST_TIMESTEP, 2, uint32_t(e + 1), 1));
                continue;
            }
            if (P.htype == 2)
            {
                double ha = 0.0;
                bool fb = false;
                for (int a = 1; a <= 4; ++a)
                {
                    const double v = alen[conn_at(n, a)];
                    if (v <= 0.0 || !isfinite(v))
                    {
                        fb = true;
                        break;
                    }
                    ha += v;
                }
                ha = fb ? h : ha / 4.0;
                h = dmin(h, ha);
            }
            // timestep_speed (:149-159)
            double speed;
            if (P.htype == 3)
            {
                double u = 0.0, v = 0.0, w = 0.0;
                for (int a = 1; a <= 4; ++a)
                {
                    const int ip = conn_at(n, a);
                    u += unkno[3 * size_t(ip)];
                    v += unkno[3 * size_t(ip) + 1];
                    w += unkno[3 * size_t(ip) + 2];
                }
                u /= 4.0;
                v /= 4.0;
                w /= 4.0;
                speed = sqrt(u * u + v * v + w * w);
            }
            else
            {
                speed = 0.0;
                for (int a = 1; a <= 4; ++a) speed = dmax(speed, node_speed(unkno, conn_at(n, a)));
            }
            // element_diffusivity_for_timestep (:170-215)
            double diff = seg_diff;
            if (P.dim_mat)
            {
                const double al = alpha[e];
                if (al <= 0.0 || !isfinite(al))
                {
                    raise_error(s, err_key(ST_TIMESTEP, 2, uint32_t(e + 1), 2));
                    continue;
                }
                diff = al;
                if (fluid)
                {
                    if (rho[e] <= 0.0 || mu[e] < 0.0)
                    {
                        raise_error(s, err_key(ST_TIMESTEP, 2, uint32_t(e + 1), 3));
                        continue;
                    }
                    diff = dmax(diff, mu[e] / rho[e]);
                }
            }
            if (diff <= 0.0 || !isfinite(diff))
            {
                raise_error(s, err_key(ST_TIMESTEP, 2, uint32_t(e + 1), 4));
                continue;
            }
            // local_stability_dt (:217-260)
            const double adv = fluid ? h / dmax(speed, 1.0e-14) : DBL_MAX;
            const double dif = h * h * (0.5 / diff);
            double dt = P.csafm * dmin(adv, dif);
            if (P.ilots == 2 && P.deltr > 0.0 && isfinite(P.deltr)) dt = dmin(dt, P.deltr);
            if (dt <= 0.0 || !isfinite(dt))
            {
                raise_error(s, err_key(ST_TIMESTEP, 2, uint32_t(e + 1), 5));
                continue;
            }
            delte[e] = dt;
            m = dmin(m, dt);
        }
        // block minimum (order-independent, exact)
        {
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = dmin(m, __shfl_xor(m, o));
            if (lane == 0) ldm[wave] = m;
            __syncthreads();
            m = ldm[0];
            for (int w = 1; w < kBlock / 64; ++w) m = dmin(m, ldm[w]);
        }
        if (threadIdx.x == 0) part[blockIdx.x * 4] = m;
    }

    // Minimum of the block minima of one segment (exact in any order); one
    // block per segment.
    __global__ __launch_bounds__(kBlock) void k_ts_elem_fin(const double* __restrict__ part,
                                                             const int32_t* __restrict__ seg_ptr, DevScalars* s)
    {
        __shared__ double ldm[kBlock / 64];
        const int sg = blockIdx.x;
        double mm = DBL_MAX;
        for (int b = seg_ptr[sg] + int(threadIdx.x); b < seg_ptr[sg + 1]; b += kBlock) mm = dmin(mm, part[size_t(b) * 4]);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mm = dmin(mm, __shfl_xor(mm, o));
        if (lane == 0) ldm[wave] = mm;
        __syncthreads();
        if (threadIdx.x != 0) return;
        mm = ldm[0];
        for (int w = 1; w < kBlock / 64; ++w) mm = dmin(mm, ldm[w]);
        if (mm <= 0.0 || !isfinite(mm)) raise_error(s, err_key(ST_TIMESTEP, 3, 0, 1));
        s[sg].dt_min = mm;
        s[sg].dtreal = mm;
    }

    __global__ void k_ts_global(int N, int E, TsParams P, double* __restrict__ delte, double* __restrict__ deltp,
                                double* __restrict__ deltp1, double* __restrict__ deltp2,
                                const uint16_t* __restrict__ seg_of_node, const uint16_t* __restrict__ seg_of_elem,
                                const DevScalars* s)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (P.global)
        {
            if (i < E) delte[i] = s[seg_of_elem ? seg_of_elem[i] : 0].dt_min;
            if (i < N)
            {
                const double dtm = s[seg_of_node ? seg_of_node[i] : 0].dt_min;
                deltp[i] = dtm;
                deltp2[i] = dtm;
            }
        }
        if (i < N)
        {
            if (P.rem_deltp > 0)
            {
                if (deltp1[i] > 0.0)
                {
                    deltp[i] = dmin(deltp[i], deltp1[i]);
                    deltp2[i] = dmin(deltp2[i], deltp1[i]);
                }
            }
            deltp1[i] = deltp[i];
        }
    }

    __global__ void k_ts_fixed(int N, int E, int nseg, double dtfix, double* __restrict__ delte,
                               double* __restrict__ deltp, double* __restrict__ deltp2, double* __restrict__ beta1,
                               DevScalars* s)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < E) delte[i] = dtfix;
        if (i < N)
        {
            deltp[i] = dtfix;
            deltp2[i] = dtfix;
            beta1[i] = 1.0;
        }
        if (i < nseg) s[i].dtreal = dtfix;
    }

    void launch_timestep(cbs_hip_ctx* c, int iitime)
    {
        const cbs_config& g = c->cfg;
        TsParams P{g.csafm, g.ani, g.deltr, g.dtfix, g.ilots, g.htype, g.beta_opt, g.rem_deltp, 0,
                   (g.dimensional_mode > 0 && g.material_properties_enabled > 0) ? 1 : 0};
        const int NE = c->N > c->E ? c->N : c->E;
        if (g.ilots <= -1 || g.dtfixed > 0 || (g.dtfix_end > 0 && iitime <= g.dtfix_end))
        {
            const int n = std::max(NE, c->nseg);
            hipLaunchKernelGGL(k_ts_fixed, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->N, c->E, c->nseg,
                               g.dtfix, c->delte, c->deltp, c->deltp2, c->beta1, c->scal);
            return;
        }
        P.global = (g.transient_on > 0 || g.solver_opt > 1 || g.dtfixed == -1 || (g.ilots != 1 && g.ilots != 2)) ? 1 : 0;
        hipLaunchKernelGGL(k_ts_nodal, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, P, c->unkno,
                           c->alen_static, c->alen, c->velcp, c->unit_vec, c->deltp, c->deltp2, c->beta1, c->vvis,
                           c->seg_of_node, c->seg_ani,
                           c->scal);
        hipLaunchKernelGGL(k_ts_elem, dim3(c->eblk.n), dim3(kBlock), 0, c->stream, c->eblk.map(), P, c->conn, c->mat,
                           c->alen_e, c->alen_static, c->unkno, c->rho, c->mu, c->alpha, c->delte, c->partial,
                           c->seg_ani,
                           c->scal);
        hipLaunchKernelGGL(k_ts_elem_fin, dim3(c->nseg), dim3(kBlock), 0, c->stream, c->partial, c->eblk.seg_ptr,
                           c->scal);
        hipLaunchKernelGGL(k_ts_global, dim3(grid_for(NE)), dim3(kBlock), 0, c->stream, c->N, c->E, P, c->delte,
                           c->deltp, c->deltp1, c->deltp2, c->seg_of_node, c->seg_of_elem, c->scal);
    }

    // ------------------------------------------------------------------
    // Node sums of the two-pass assemblies (cbs_hip_ctx::NodeSlots): the
    // element passes write each (element, vertex) term to its slot; a node
    // pass thread reads its node's slots in order, coalesced across the
    // 64 nodes of a slice.
    // ------------------------------------------------------------------
    struct SlotView
    {
        const int64_t* off;
        const uint32_t* pos;
        double* q;
        int64_t n;  // slots per plane
        const int32_t* grp_ptr;
        const uint32_t* grp_line;
        const uint16_t* grp_src;
    };

    inline SlotView slot_view(const cbs_hip_ctx::NodeSlots& ns)
    {
        return {ns.off, ns.pos, ns.q, ns.nslots, ns.grp_ptr, ns.grp_line, ns.grp_src};
    }

    inline int slot_window_grid(int E) { return (E + kSlotWindow - 1) / kSlotWindow; }

    // The block's staged terms sq[p][local element * 4 + vertex] stored as
    // whole lines (SlotView's line groups): 16 consecutive lanes write one
    // 128-byte line of slots, so a line the block completes is written by
    // one store instruction instead of piecemeal by several wavefronts
    // (each partial write made the L2 fetch the line first).  The LHS and
    // Step 3 element passes (slot order slots_e, three planes) stage their
    // terms this way.  Called by every thread of the block after the terms
    // are staged.
    template <int P>
    __device__ inline void store_line_groups(const SlotView& sv, const double (*sq)[kSlotWindow * 4])
    {
        const int g0 = sv.grp_ptr[blockIdx.x], g1 = sv.grp_ptr[blockIdx.x + 1];
        for (int idx = g0 * 16 + int(threadIdx.x); idx < g1 * 16; idx += kSlotWindow)
        {
            const uint16_t src = sv.grp_src[idx];
            if (src == 0xFFFFu) continue;
            const size_t p = size_t(sv.grp_line[idx >> 4]) * 16 + size_t(idx & 15);
#pragma unroll
            for (int pl = 0; pl < P; ++pl) sv.q[size_t(pl) * size_t(sv.n) + p] = sq[pl][src];
        }
    }

    // Element e's terms qv[plane][vertex] into their slots: through the
    // block's line groups (kLines: every thread of the block calls it), or
    // one 8-byte store each (meshes too small for the staging to pay).
    template <int P, bool kLines>
    __device__ inline void put_terms(const SlotView& sv, int e, int E, const double (&qv)[P][4])
    {
        if constexpr (kLines)
        {
            __shared__ double sq[P][kSlotWindow * 4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int pl = 0; pl < P; ++pl) sq[pl][threadIdx.x * 4 + a] = qv[pl][a];
            __syncthreads();
            store_line_groups<P>(sv, sq);
        }
        else if (e < E)
        {
#pragma unroll
            for (int a = 0; a < 4; ++a)
            {
                const size_t p = sv.pos[size_t(e) * 4 + a];
#pragma unroll
                for (int pl = 0; pl < P; ++pl) sv.q[size_t(pl) * size_t(sv.n) + p] = qv[pl][a];
            }
        }
    }

    // the element passes' launch: line groups where they were built
    template <typename K0, typename K1, typename... A>
    void launch_elem_pass(cbs_hip_ctx* c, const cbs_hip_ctx::NodeSlots& ns, K0 plain, K1 lines, A... a)
    {
        if (ns.grp_ptr)
            hipLaunchKernelGGL(lines, dim3(slot_window_grid(c->E)), dim3(kSlotWindow), 0, c->stream, a...);
        else
            hipLaunchKernelGGL(plain, dim3(grid_for(c->E)), dim3(kBlock), 0, c->stream, a...);
    }

    // node i's slots: o + k*64 for k < K.  The node passes read each slot
    // once, with non-temporal loads (0.18 ms less per step at cavity160
    // across the four passes than ordinary loads).
    __device__ inline void node_slots(const int64_t* __restrict__ off, int i, int64_t& o, int64_t& K)
    {
        const int s = i >> 6;
        o = off[s] + (i & 63);
        K = (off[s + 1] - off[s]) >> 6;
    }

    // ------------------------------------------------------------------
    // TimeStep::updateLhsDiagonal (TimeStep.cpp:778-954), in two passes.
    // Element pass (element order, reads coalesced, stores as whole lines):
    // the three terms of every (element, vertex) -- mass/dt (fluid), rho_cp
    // mass/dt, pdiagE dt (fluid) -- with +0.0 for a term the reference
    // skips.  Node pass: each node adds
    // its terms in element order, the reference's sequence of additions.  A
    // sum that starts at +0.0 can never become -0.0, so the +0.0 terms and
    // the padding slots change no bit.
    // ------------------------------------------------------------------
    template <bool kLines>
    __global__ __launch_bounds__(kSlotWindow) void k_lhs_elem(int E, const int32_t* __restrict__ mat,
                                                              const double* __restrict__ delte,
                                                              const double* __restrict__ elcoe,
                                                              const double* __restrict__ rcp,
                                                              const double* __restrict__ pdiagE, SlotView sv,
                                                              DevScalars* s)
    {
        const int e = int(blockIdx.x * blockDim.x + threadIdx.x);
        double qv[3][4] = {};
        if (e < E)
        {
            const double dt = delte[e];
            const bool dt_ok = !(dt <= 0.0 || !isfinite(dt));
            if (!dt_ok) raise_error(s, err_key(ST_LHS, 1, uint32_t(e + 1), 1));
            const bool fluid = mat[e] == 0;
            const double rc = rcp[e];
            for (int a = 0; a < 4; ++a)
            {
                double q1 = 0.0, q2 = 0.0, q3 = 0.0;
                const double mass = elcoe[size_t(e) * 4 + a];
                if (dt_ok)
                {
                    if (mass <= 0.0 || !isfinite(mass))
                    {
                        raise_error(s, err_key(ST_LHS, 1, uint32_t(e + 1), 2));
                    }
                    else
                    {
                        if (fluid) q1 = mass / dt;
                        q2 = rc * mass / dt;
                        if (fluid) q3 = pdiagE[size_t(e) * 4 + a] * dt;
                    }
                }
                qv[0][a] = q1;
                qv[1][a] = q2;
                qv[2][a] = q3;
            }
        }
        put_terms<3, kLines>(sv, e, E, qv);
    }

    __global__ void k_lhs(int N, SlotView sv, const uint8_t* __restrict__ nflag, double* __restrict__ elcoe2,
                          double* __restrict__ elcoe2p, double* __restrict__ pdiag, double* __restrict__ pres,
                          double* __restrict__ rhs1, DevScalars* s)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double m2 = 0.0, m2p = 0.0, pd = 0.0;
        int64_t o, K;
        node_slots(sv.off, i, o, K);
        for (int64_t k = 0; k < K; ++k)
        {
            const size_t p = size_t(o + k * 64);
            m2 += __builtin_nontemporal_load(&sv.q[p]);
            m2p += __builtin_nontemporal_load(&sv.q[sv.n + p]);
            pd += __builtin_nontemporal_load(&sv.q[2 * sv.n + p]);
        }
        const bool tf = nflag[i] & 1u, ts = nflag[i] & 2u;
        if (!tf && !ts) raise_error(s, err_key(ST_LHS, 2, uint32_t(i + 1), 1));
        if (tf)
        {
            if (m2 <= 0.0 || !isfinite(m2)) raise_error(s, err_key(ST_LHS, 2, uint32_t(i + 1), 2));
            elcoe2[i] = 1.0 / m2;
        }
        else
        {
            elcoe2[i] = 0.0;
        }
        if (m2p <= 0.0 || !isfinite(m2p)) raise_error(s, err_key(ST_LHS, 2, uint32_t(i + 1), 3));
        elcoe2p[i] = 1.0 / m2p;
        if (tf)
        {
            if (fabs(pd) <= 1.0e-14 || !isfinite(pd)) raise_error(s, err_key(ST_LHS, 3, uint32_t(i + 1), 1));
        }
        else if (ts)
        {
            pd = 0.0;
            pres[i] = 0.0;
            rhs1[i] = 0.0;
        }
        pdiag[i] = pd;
    }

    __global__ void k_lhs_fixed(const int32_t* __restrict__ node, const double* __restrict__ val, int m,
                                double* __restrict__ pdiag, double* __restrict__ pres, double* __restrict__ rhs1)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= m) return;
        pdiag[node[i]] = 1.0;
        pres[node[i]] = val[i];
        rhs1[node[i]] = val[i];
    }

    void launch_lhs(cbs_hip_ctx* c)
    {
        launch_elem_pass(c, c->slots_e, k_lhs_elem<false>, k_lhs_elem<true>, c->E, c->mat, c->delte, c->elcoe,
                         c->rcp, c->pdiagE, slot_view(c->slots_e), c->scal);
        hipLaunchKernelGGL(k_lhs, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, slot_view(c->slots_e),
                           c->nflag, c->elcoe2, c->elcoe2p, c->pdiag, c->pres, c->rhs1, c->scal);
        if (c->npres_bc > 0)
        {
            hipLaunchKernelGGL(k_lhs_fixed, dim3(grid_for(c->npres_bc)), dim3(kBlock), 0, c->stream, c->pbc_node,
                               c->pbc_val, c->npres_bc, c->pdiag, c->pres, c->rhs1);
        }
        launch_gstif(c);
    }

    // ------------------------------------------------------------------
    // Step 1: MomentumAssembly::assembleStep1Rhs (MomentumAssembly.cpp:144-554),
    // in two passes.  The reference's colour loop adds each fluid element's
    // 3 x 4 terms to its nodes colour by colour (one element per node and
    // colour), so every node sums its elements in (colour, element) order.
    // Element pass (element order, coalesced): the element's terms, +0.0 for
    // an element the loop rejects.  Node pass: each node adds its terms in
    // (colour, element) order through the colour-sorted incidence list.
    // ------------------------------------------------------------------
    struct S1Params
    {
        double ani;
        int dim_mat;
    };

    __global__ __launch_bounds__(kBlock) void k_step1_elem(int F, S1Params P,
                                                            const int32_t* __restrict__ slot_elem,
                                                            const int4* __restrict__ conn,
                                                             const double* __restrict__ dNg,
                                                             const double* __restrict__ detJ,
                                                             const double* __restrict__ delte,
                                                             const uint32_t* __restrict__ fedge,
                                                             const double* __restrict__ annxf,
                                                             const double* __restrict__ rho,
                                                             const double* __restrict__ mu,
                                                             const double* __restrict__ unkn1, SlotView sv,
                                                             const uint16_t* __restrict__ seg_of_elem,
                                                             const double* __restrict__ seg_ani, DevScalars* s)
    {
        const int k = int(blockIdx.x) * kBlock + int(threadIdx.x);
        if (k >= F) return;
        const int e = slot_elem[k];
        size_t p[4];
        #pragma unroll
        for (int a = 0; a < 4; ++a) p[a] = sv.pos[size_t(e) * 4 + a];
        // an element that fails a check leaves zero terms (the step reports
        // the failure)
        auto fail = [&](uint64_t key) {
            raise_error(s, key);
            #pragma unroll
            for (int a = 0; a < 4; ++a) sv.q[p[a]] = sv.q[sv.n + p[a]] = sv.q[2 * sv.n + p[a]] = 0.0;
        };
        const double dj = detJ[e];
        if (dj <= 0.0 || !isfinite(dj))
        {
            fail(err_key(ST_STEP1, 2, uint32_t(e + 1), 1));
            return;
        }
        const int4 n = conn[e];
        const double* g = dNg + size_t(e) * 12;
#define DN(d, a) g[((d) - 1) * 4 + (a) - 1]
        const uint32_t fe = fedge[e];
        const double* an = annxf + size_t(e) * 16;
#define ANN(cc, is) an[((is) - 1) * 4 + (cc) - 1]
        double lunkno[4][5];
        #pragma unroll
        for (int a = 1; a <= 4; ++a)
        {
            const int ip = conn_at(n, a);
            #pragma unroll
            for (int i = 1; i <= 3; ++i) lunkno[i][a] = unkn1[3 * size_t(ip) + i - 1];
        }
        // lunk(i, j, a) = u_j u_i at vertex a (the reference's lunk array),
        // formed at each use rather than held: the same product, so the same
        // bits, in fewer registers
#define LUNK(i, j, a) (lunkno[j][a] * lunkno[i][a])
        double lrhs[4][5];
        #pragma unroll
        for (int i = 0; i < 4; ++i)
            #pragma unroll
            for (int a = 0; a < 5; ++a) lrhs[i][a] = 0.0;

        // step1_convective_galerkin
        {
            double lunksum[4][4];
            const double volume_factor = dj * FCON1;
            #pragma unroll
            for (int i = 1; i <= 3; ++i)
                #pragma unroll
                for (int j = 1; j <= 3; ++j)
                {
                    double t = 0.0;
                    #pragma unroll
                    for (int in = 1; in <= 4; ++in) t += LUNK(i, j, in);
                    lunksum[i][j] = t;
                }
            #pragma unroll
            for (int a = 1; a <= 4; ++a)
                #pragma unroll
                for (int i = 1; i <= 3; ++i)
                    #pragma unroll
                    for (int j = 1; j <= 3; ++j) lrhs[i][a] += DN(j, a) * volume_factor * lunksum[i][j];
            #pragma unroll
            for (int is = 1; is <= 4; ++is)
            {
                if (fedge_at(fe, is) == 0) continue;
                double lsf[4][4];
                #pragma unroll
                for (int i = 1; i <= 3; ++i)
                    #pragma unroll
                    for (int j = 1; j <= 3; ++j) lsf[i][j] = lunksum[i][j] - LUNK(i, j, is);
                #pragma unroll
                for (int j = 1; j <= 3; ++j)
                {
                    const double face_factor = ANN(j, is) * FCON2;
                    #pragma unroll
                    for (int fn = 1; fn <= 3; ++fn)
                    {
                        const int a = ippn1(is, fn);
                        #pragma unroll
                        for (int i = 1; i <= 3; ++i) lrhs[i][a] -= (lsf[i][j] + LUNK(i, j, a)) * face_factor;
                    }
                }
            }
        }
        // step1_characteristic_correction
        {
            double lunksum[4][4], lunksumk[4][4], umean[4];
            #pragma unroll
            for (int i = 1; i <= 3; ++i)
                #pragma unroll
                for (int j = 1; j <= 3; ++j)
                {
                    double t = 0.0;
                    #pragma unroll
                    for (int a = 1; a <= 4; ++a) t += DN(j, a) * LUNK(i, j, a);
                    lunksum[i][j] = t;
                }
            #pragma unroll
            for (int i = 1; i <= 3; ++i)
            {
                double t = 0.0;
                #pragma unroll
                for (int a = 1; a <= 4; ++a) t += lunkno[i][a];
                umean[i] = t / 4.0;
            }
            #pragma unroll
            for (int i = 1; i <= 3; ++i)
            {
                double sum_j = 0.0;
                #pragma unroll
                for (int j = 1; j <= 3; ++j) sum_j += lunksum[i][j];
                #pragma unroll
                for (int kk = 1; kk <= 3; ++kk) lunksumk[i][kk] = umean[kk] * sum_j;
            }
            const double dt = delte[e];
            if (dt <= 0.0 || !isfinite(dt))
            {
                fail(err_key(ST_STEP1, 1, uint32_t(e + 1), 2));
                return;
            }
            const double ldelte = 0.5 * dt * dj * FCON3;
            const double ldelte2 = 0.5 * dt * FCON4;
            #pragma unroll
            for (int a = 1; a <= 4; ++a)
                #pragma unroll
                for (int kk = 1; kk <= 3; ++kk)
                    #pragma unroll
                    for (int i = 1; i <= 3; ++i) lrhs[i][a] -= DN(kk, a) * lunksumk[i][kk] * ldelte;
            #pragma unroll
            for (int is = 1; is <= 4; ++is)
            {
                if (fedge_at(fe, is) == 0) continue;
                #pragma unroll
                for (int kk = 1; kk <= 3; ++kk)
                {
                    const double face_factor = ANN(kk, is) * ldelte2;
                    #pragma unroll
                    for (int fn = 1; fn <= 3; ++fn)
                    {
                        const int a = ippn1(is, fn);
                        #pragma unroll
                        for (int i = 1; i <= 3; ++i) lrhs[i][a] += lunksumk[i][kk] * face_factor;
                    }
                }
            }
        }
        // step1_diffusion
        {
            double ani = seg_ani ? seg_ani[seg_of_elem[e]] : P.ani;
            if (P.dim_mat)
            {
                if (rho[e] <= 0.0 || mu[e] < 0.0)
                {
                    fail(err_key(ST_STEP1, 1, uint32_t(e + 1), 3));
                    return;
                }
                ani = mu[e] / rho[e];
            }
            double dNu[4][4];
            #pragma unroll
            for (int i = 1; i <= 3; ++i)
                #pragma unroll
                for (int j = 1; j <= 3; ++j) dNu[i][j] = 0.0;
            #pragma unroll
            for (int i = 1; i <= 3; ++i)
                #pragma unroll
                for (int a = 1; a <= 4; ++a)
                    #pragma unroll
                    for (int j = 1; j <= 3; ++j) dNu[i][j] += DN(j, a) * lunkno[i][a];
            const double volume_factor = ani * dj * FDIF1;
            #pragma unroll
            for (int i = 1; i <= 3; ++i)
                #pragma unroll
                for (int a = 1; a <= 4; ++a)
                    #pragma unroll
                    for (int j = 1; j <= 3; ++j) lrhs[i][a] -= DN(j, a) * dNu[i][j] * volume_factor;
            const double fdf = ani * FDIF2;
            #pragma unroll
            for (int is = 1; is <= 4; ++is)
            {
                if (fedge_at(fe, is) == 0) continue;
                #pragma unroll
                for (int j = 1; j <= 3; ++j)
                {
                    const double face_factor = ANN(j, is) * fdf;
                    #pragma unroll
                    for (int i = 1; i <= 3; ++i)
                        #pragma unroll
                        for (int fn = 1; fn <= 3; ++fn)
                        {
                            const int a = ippn1(is, fn);
                            lrhs[i][a] += dNu[i][j] * face_factor;
                        }
                }
            }
        }
#undef DN
#undef ANN
#undef LUNK
        #pragma unroll
        for (int a = 1; a <= 4; ++a)
            #pragma unroll
            for (int i = 1; i <= 3; ++i) sv.q[size_t(i - 1) * sv.n + p[a - 1]] = lrhs[i][a];
    }

    // rhs(node) = its elements' terms summed in (colour, element) order
    __global__ void k_step1_node(int N, SlotView sv, double* __restrict__ rhs)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double r1 = 0.0, r2 = 0.0, r3 = 0.0;
        int64_t o, K;
        node_slots(sv.off, i, o, K);
        for (int64_t k = 0; k < K; ++k)
        {
            const size_t p = size_t(o + k * 64);
            r1 += __builtin_nontemporal_load(&sv.q[p]);
            r2 += __builtin_nontemporal_load(&sv.q[sv.n + p]);
            r3 += __builtin_nontemporal_load(&sv.q[2 * sv.n + p]);
        }
        rhs[3 * size_t(i)] = r1;
        rhs[3 * size_t(i) + 1] = r2;
        rhs[3 * size_t(i) + 2] = r3;
    }

    __global__ void k_fill(double* __restrict__ p, int64_t n, double v)
    {
        const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
        if (i < n) p[i] = v;
    }

    __global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, int64_t n)
    {
        const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
        if (i < n) d[i] = s[i];
    }

    void launch_fill(cbs_hip_ctx* c, double* p, int64_t n, double v)
    {
        if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, p, n, v);
    }

    void launch_copy(cbs_hip_ctx* c, double* dst, const double* src, int64_t n)
    {
        if (n > 0) hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, dst, src, n);
    }

    __global__ void k_check_dtreal(DevScalars* s, int nseg, uint32_t stage)
    {
        for (int k = 0; k < nseg; ++k)
        {
            const double dt = s[k].dtreal;
            if (dt <= 0.0 || !isfinite(dt)) raise_error(s, err_key(stage, 0, 0, 1));
        }
    }

    void launch_step1_rhs(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_check_dtreal, dim3(1), dim3(1), 0, c->stream, c->scal, c->nseg, uint32_t(ST_STEP1));
        const S1Params P{c->cfg.ani, (c->cfg.dimensional_mode > 0 && c->cfg.material_properties_enabled > 0) ? 1 : 0};
        if (c->F > 0)
            hipLaunchKernelGGL(k_step1_elem, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->F, P, c->elem_fluid,
                               c->conn, c->dNkdx, c->detJ, c->delte, c->fedge, c->annxf, c->rho, c->mu, c->unkn1,
                               slot_view(c->slots_c), c->seg_of_elem, c->seg_ani, c->scal);
        hipLaunchKernelGGL(k_step1_node, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N,
                           slot_view(c->slots_c), c->rhs);
    }

    // u = unkn1 + rhs * elcoe2 (Steps.cpp:132-142)
    __global__ void k_step1_update(int N, const double* __restrict__ unkn1, const double* __restrict__ rhs,
                                   const double* __restrict__ elcoe2, double* __restrict__ unkno)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= N) return;
        const double w = elcoe2[i];
        for (int d = 0; d < 3; ++d) unkno[3 * size_t(i) + d] = unkn1[3 * size_t(i) + d] + rhs[3 * size_t(i) + d] * w;
    }

    void launch_step1_update(cbs_hip_ctx* c)
    {
        if (c->bforce)
        {
            launch_step1_update_forced(c);  // body force of the adjoint workload
            return;
        }
        hipLaunchKernelGGL(k_step1_update, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->unkn1, c->rhs,
                           c->elcoe2, c->unkno);
    }

    // ------------------------------------------------------------------
    // Step 2 RHS: PressureAssembly::assembleStep2Rhs (PressureAssembly.cpp:236-335),
    // in two passes.  Element pass: the term of every (element, vertex) --
    // the volume term, then each boundary face's term subtracted in face
    // order, exactly as the element loop forms it -- and +0.0 for solid or
    // rejected elements.  Node pass: each node adds its terms in element order.
    // ------------------------------------------------------------------
    // The four terms of element e (+0.0 for a solid or rejected element).
    __device__ inline void step2_terms(int e, const int32_t* __restrict__ mat, const int4* __restrict__ conn,
                                       const double* __restrict__ dNg, const double* __restrict__ detJ,
                                       const uint32_t* __restrict__ fedge, const double* __restrict__ annxf,
                                       const double* __restrict__ unkno, const double* __restrict__ unkn1,
                                       DevScalars* s, double* out)
    {
        if (mat[e] != 0) return;
        const double dj = detJ[e];
        if (dj <= 0.0 || !isfinite(dj))
        {
            raise_error(s, err_key(ST_STEP2RHS, 1, uint32_t(e + 1), 1));
            return;
        }
        const double vol4 = dj * FCON1;
        const int4 n = conn[e];
        double us[4] = {0.0, 0.0, 0.0, 0.0};
        for (int in = 1; in <= 4; ++in)
        {
            const int ip = conn_at(n, in);
            us[1] += unkno[3 * size_t(ip)];
            us[2] += unkno[3 * size_t(ip) + 1];
            us[3] += unkno[3 * size_t(ip) + 2];
        }
        us[1] *= vol4;
        us[2] *= vol4;
        us[3] *= vol4;
        const double* g = dNg + size_t(e) * 12;
        const uint32_t fe = fedge[e];
        // face sums of the old velocity, shared by the vertices
        double fus[5][4];
        for (int is = 1; is <= 4; ++is)
        {
            fus[is][1] = fus[is][2] = fus[is][3] = 0.0;
            if (fedge_at(fe, is) != 1) continue;
            for (int t = 1; t <= 3; ++t)
            {
                const int ip = conn_at(n, ippn1(is, t));
                fus[is][1] += unkn1[3 * size_t(ip)];
                fus[is][2] += unkn1[3 * size_t(ip) + 1];
                fus[is][3] += unkn1[3 * size_t(ip) + 2];
            }
        }
        for (int a = 1; a <= 4; ++a)
        {
            double l = g[a - 1] * us[1] + g[4 + a - 1] * us[2] + g[8 + a - 1] * us[3];
            const int ipa = conn_at(n, a);
            for (int is = 1; is <= 4; ++is)
            {
                if (fedge_at(fe, is) != 1 || is == a) continue;
                for (int idim = 1; idim <= 3; ++idim)
                {
                    const double nf = annxf[(size_t(e) * 4 + is - 1) * 4 + idim - 1] * FCON2;
                    l -= (fus[is][idim] + unkn1[3 * size_t(ipa) + idim - 1]) * nf;
                }
            }
            out[a - 1] = l;
        }
    }


    template <bool kLines>
    __global__ __launch_bounds__(kSlotWindow) void k_step2_elem(int E, const int32_t* __restrict__ mat,
                                                                const int4* __restrict__ conn,
                                                                const double* __restrict__ dNg,
                                                                const double* __restrict__ detJ,
                                                                const uint32_t* __restrict__ fedge,
                                                                const double* __restrict__ annxf,
                                                                const double* __restrict__ unkno,
                                                                const double* __restrict__ unkn1, SlotView sv,
                                                                DevScalars* s)
    {
        const int e = int(blockIdx.x * blockDim.x + threadIdx.x);
        double qv[1][4] = {};
        if (e < E) step2_terms(e, mat, conn, dNg, detJ, fedge, annxf, unkno, unkn1, s, qv[0]);
        put_terms<1, kLines>(sv, e, E, qv);
    }

    __global__ void k_step2_rhs(int N, SlotView sv, double* __restrict__ rhs1)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double acc = 0.0;
        int64_t o, K;
        node_slots(sv.off, i, o, K);
        for (int64_t k = 0; k < K; ++k) acc += __builtin_nontemporal_load(&sv.q[size_t(o + k * 64)]);
        rhs1[i] = acc;
    }

    void launch_step2_rhs(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_check_dtreal, dim3(1), dim3(1), 0, c->stream, c->scal, c->nseg, uint32_t(ST_STEP2RHS));
        // one plane: its stores one at a time (through the line groups it
        // measured 1.19 against 1.16 ms at cavity160; LHS and Step 3 gain)
        hipLaunchKernelGGL(k_step2_elem<false>, dim3(grid_for(c->E)), dim3(kBlock), 0, c->stream, c->E, c->mat,
                           c->conn, c->dNkdx, c->detJ, c->fedge, c->annxf, c->unkno, c->unkn1, slot_view(c->slots_e),
                           c->scal);
        hipLaunchKernelGGL(k_step2_rhs, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N,
                           slot_view(c->slots_e), c->rhs1);
    }

    // ------------------------------------------------------------------
    // Step 3: Steps::step3SemiImplicit (Steps.cpp:377-428), fused with
    // u += rhs * elcoe2, in two passes.  Element pass: the element's term
    // vol/4 * grad p (the same for its four vertices), +0.0 for solid or
    // rejected elements.  Node pass: each node subtracts its elements' terms
    // in element order.
    // ------------------------------------------------------------------
    template <bool kLines>
    __global__ __launch_bounds__(kSlotWindow) void k_step3_elem(int E, const int32_t* __restrict__ mat,
                                                                const int4* __restrict__ conn,
                                                                const double* __restrict__ dNg,
                                                                const double* __restrict__ detJ,
                                                                const double* __restrict__ pres, SlotView sv,
                                                                DevScalars* s)
    {
        const int e = int(blockIdx.x * blockDim.x + threadIdx.x);
        double qe[3] = {0.0, 0.0, 0.0};
        if (e < E)
        {
            const double dj = detJ[e];
            if (mat[e] == 0 && (dj <= 0.0 || !isfinite(dj))) raise_error(s, err_key(ST_STEP3, 1, uint32_t(e + 1), 1));
            else if (mat[e] == 0)
            {
                const int4 n = conn[e];
                const double* g = dNg + size_t(e) * 12;
                double gp[3] = {0.0, 0.0, 0.0};
                for (int d = 0; d < 3; ++d)
                    for (int a = 1; a <= 4; ++a) gp[d] += g[d * 4 + a - 1] * pres[conn_at(n, a)];
                const double vol4 = dj * FCON1;
                qe[0] = vol4 * gp[0];
                qe[1] = vol4 * gp[1];
                qe[2] = vol4 * gp[2];
            }
        }
        double qv[3][4];
        for (int a = 0; a < 4; ++a)
            for (int pl = 0; pl < 3; ++pl) qv[pl][a] = qe[pl];
        put_terms<3, kLines>(sv, e, E, qv);
    }

    __global__ void k_step3(int N, SlotView sv, const double* __restrict__ elcoe2, double* __restrict__ rhs,
                            double* __restrict__ unkno)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double r1 = 0.0, r2 = 0.0, r3 = 0.0;
        int64_t o, K;
        node_slots(sv.off, i, o, K);
        for (int64_t k = 0; k < K; ++k)
        {
            const size_t p = size_t(o + k * 64);
            r1 -= __builtin_nontemporal_load(&sv.q[p]);
            r2 -= __builtin_nontemporal_load(&sv.q[sv.n + p]);
            r3 -= __builtin_nontemporal_load(&sv.q[2 * sv.n + p]);
        }
        rhs[3 * size_t(i)] = r1;
        rhs[3 * size_t(i) + 1] = r2;
        rhs[3 * size_t(i) + 2] = r3;
        const double w = elcoe2[i];
        unkno[3 * size_t(i)] += r1 * w;
        unkno[3 * size_t(i) + 1] += r2 * w;
        unkno[3 * size_t(i) + 2] += r3 * w;
    }

    void launch_step3(cbs_hip_ctx* c)
    {
        launch_elem_pass(c, c->slots_e, k_step3_elem<false>, k_step3_elem<true>, c->E, c->mat, c->conn, c->dNkdx,
                         c->detJ, c->pres, slot_view(c->slots_e), c->scal);
        hipLaunchKernelGGL(k_step3, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, slot_view(c->slots_e),
                           c->elcoe2, c->rhs, c->unkno);
    }

    // ------------------------------------------------------------------
    // Step 4: EnergyAssembly::assembleStep4Rhs (EnergyAssembly.cpp:118-611)
    // ------------------------------------------------------------------
    __global__ void k_step4_rhs(int N, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                                const int32_t* __restrict__ mat, const int4* __restrict__ conn,
                                const double* __restrict__ dNg, const double* __restrict__ detJ,
                                const double* __restrict__ delte, const double* __restrict__ rcp,
                                const double* __restrict__ kcond, const double* __restrict__ qvol,
                                const double* __restrict__ unkno, const double* __restrict__ temp1,
                                const int32_t* __restrict__ hf_ptr, const double* __restrict__ hf_val,
                                const int32_t* __restrict__ hf_slot, double* __restrict__ rhs1, DevScalars* s)
    {
        const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
        if (i >= N) return;
        double acc = 0.0;
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k)
        {
            const int e = inc[k] >> 2, a = (inc[k] & 3) + 1;
            const double dj = detJ[e];
            if (dj <= 0.0 || !isfinite(dj))
            {
                raise_error(s, err_key(ST_STEP4, 1, uint32_t(e + 1), 1));
                continue;
            }
            const double rc = rcp[e], kk = kcond[e], q = qvol[e];
            if (rc <= 0.0 || !isfinite(rc))
            {
                raise_error(s, err_key(ST_STEP4, 1, uint32_t(e + 1), 2));
                continue;
            }
            if (kk <= 0.0 || !isfinite(kk))
            {
                raise_error(s, err_key(ST_STEP4, 1, uint32_t(e + 1), 3));
                continue;
            }
            if (!isfinite(q))
            {
                raise_error(s, err_key(ST_STEP4, 1, uint32_t(e + 1), 4));
                continue;
            }
            const int4 n = conn[e];
            const double* g = dNg + size_t(e) * 12;
            double dTdx = 0.0, dTdy = 0.0, dTdz = 0.0;
            for (int b = 1; b <= 4; ++b)
            {
                const double T = temp1[conn_at(n, b)];
                dTdx += g[b - 1] * T;
                dTdy += g[4 + b - 1] * T;
                dTdz += g[8 + b - 1] * T;
            }
            double l = 0.0;
            if (mat[e] == 0)
            {
                const double adv_factor = rc * dj / 120.0;
                double us = 0.0, vs = 0.0, ws = 0.0;
                for (int b = 1; b <= 4; ++b)
                {
                    const int ip = conn_at(n, b);
                    us += unkno[3 * size_t(ip)];
                    vs += unkno[3 * size_t(ip) + 1];
                    ws += unkno[3 * size_t(ip) + 2];
                }
                const int ipa = conn_at(n, a);
                const double uw = us + unkno[3 * size_t(ipa)];
                const double vw = vs + unkno[3 * size_t(ipa) + 1];
                const double ww = ws + unkno[3 * size_t(ipa) + 2];
                const double advT = uw * dTdx + vw * dTdy + ww * dTdz;
                l -= adv_factor * advT;
                const double dt = delte[e];
                if (dt <= 0.0 || !isfinite(dt))
                {
                    raise_error(s, err_key(ST_STEP4, 1, uint32_t(e + 1), 5));
                    continue;
                }
                double ub = 0.0, vb = 0.0, wb = 0.0;
                for (int b = 1; b <= 4; ++b)
                {
                    const int ip = conn_at(n, b);
                    ub += unkno[3 * size_t(ip)];
                    vb += unkno[3 * size_t(ip) + 1];
                    wb += unkno[3 * size_t(ip) + 2];
                }
                ub /= 4.0;
                vb /= 4.0;
                wb /= 4.0;
                const double advT2 = ub * dTdx + vb * dTdy + wb * dTdz;
                const double volume = dj / 6.0;
                const double tau = 0.5 * dt * rc * volume;
                const double ug = ub * g[a - 1] + vb * g[4 + a - 1] + wb * g[8 + a - 1];
                l += tau * ug * advT2;
            }
            {
                const double volume = dj / 6.0;
                const double gg = g[a - 1] * dTdx + g[4 + a - 1] * dTdy + g[8 + a - 1] * dTdz;
                l -= kk * volume * gg;
            }
            if (q != 0.0)
            {
                const double src = q * dj * MASS_FACTOR;
                l += src;
            }
            acc += l;
        }
        // BC-532 heat flux, after all element terms, in face order
        const int h = hf_slot[i];
        if (h >= 0)
            for (int k = hf_ptr[h]; k < hf_ptr[h + 1]; ++k) acc += hf_val[k];
        rhs1[i] = acc;
    }

    void launch_step4_rhs(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_step4_rhs, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->inc_ptr, c->inc,
                           c->mat, c->conn, c->dNkdx, c->detJ, c->delte, c->rcp, c->kcond, c->qvol, c->unkno, c->temp1,
                           c->hf_ptr, c->hf_val, c->hf_slot, c->rhs1, c->scal);
    }

    // T = T1 + rhs1 * elcoe2p (Steps.cpp:459-464)
    __global__ void k_step4_update(int N, const double* __restrict__ temp1, const double* __restrict__ rhs1,
                                   const double* __restrict__ elcoe2p, double* __restrict__ temp)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < N) temp[i] = temp1[i] + rhs1[i] * elcoe2p[i];
    }

    void launch_step4_update(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_step4_update, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->temp1, c->rhs1,
                           c->elcoe2p, c->temp);
    }

    // ------------------------------------------------------------------
    // Boundary programmes (Boundary.cpp:275-723, Steps.cpp:102-176)
    // ------------------------------------------------------------------

    __global__ void k_bc_velocity(int m, int which, const int32_t* __restrict__ node,
                                  const int32_t* __restrict__ flags, const double* __restrict__ val,
                                  const int32_t* __restrict__ sym_ptr, const double* __restrict__ sym_n,
                                  const int32_t* __restrict__ out_ptr, const double* __restrict__ out_n,
                                  double* __restrict__ unkno)
    {
        const int t = blockIdx.x * kBlock + threadIdx.x;
        if (t >= m) return;
        const size_t ip = size_t(node[t]);
        double u = unkno[3 * ip], v = unkno[3 * ip + 1], w = unkno[3 * ip + 2];
        const int f = flags[t];
        if (which == 0)
        {
            // applySymmetry: u -= (u.n) n per symmetry face, in face order
            for (int k = sym_ptr[t]; k < sym_ptr[t + 1]; ++k)
            {
                const double nx = sym_n[3 * k], ny = sym_n[3 * k + 1], nz = sym_n[3 * k + 2];
                const double un = u * nx + v * ny + w * nz;
                u -= un * nx;
                v -= un * ny;
                w -= un * nz;
            }
        }
        // applyVelocity: the value left by its last writer
        if (f & BCV_PRESCRIBED)
        {
            u = val[3 * t];
            v = val[3 * t + 1];
            w = val[3 * t + 2];
        }
        if (which == 0)
        {
            // applyOutletBackflowControl, then enforce_zero_velocity_in_solid
            for (int k = out_ptr[t]; k < out_ptr[t + 1]; ++k)
            {
                if (f & BCV_SOLID)
                {
                    u = 0.0;
                    v = 0.0;
                    w = 0.0;
                    continue;
                }
                const double nx = out_n[3 * k], ny = out_n[3 * k + 1], nz = out_n[3 * k + 2];
                const double un = u * nx + v * ny + w * nz;
                if (un < 0.0)
                {
                    u -= un * nx;
                    v -= un * ny;
                    w -= un * nz;
                }
            }
            if (f & BCV_SOLID)
            {
                u = 0.0;
                v = 0.0;
                w = 0.0;
            }
        }
        unkno[3 * ip] = u;
        unkno[3 * ip + 1] = v;
        unkno[3 * ip + 2] = w;
    }

    __global__ void k_bc_scalar(int m, const int32_t* __restrict__ node, const double* __restrict__ val,
                                double* __restrict__ x)
    {
        const int t = blockIdx.x * kBlock + threadIdx.x;
        if (t < m) x[node[t]] = val[t];
    }

    void launch_bc(cbs_hip_ctx* c, int which)
    {
        if (which == 0 || which == 1)
        {
            if (c->nbcv > 0)
                hipLaunchKernelGGL(k_bc_velocity, dim3(grid_for(c->nbcv)), dim3(kBlock), 0, c->stream, c->nbcv, which,
                                   c->bcv_node, c->bcv_flags, c->bcv_val, c->bcv_sym_ptr, c->bcv_sym_n, c->bcv_out_ptr,
                                   c->bcv_out_n, c->unkno);
        }
        else if (which == 2)
        {
            if (c->ntbc > 0)
                hipLaunchKernelGGL(k_bc_scalar, dim3(grid_for(c->ntbc)), dim3(kBlock), 0, c->stream, c->ntbc,
                                   c->tbc_node, c->tbc_val, c->temp);
        }
        else if (which == 3)
        {
            if (c->npres_bc > 0)
                hipLaunchKernelGGL(k_bc_scalar, dim3(grid_for(c->npres_bc)), dim3(kBlock), 0, c->stream, c->npres_bc,
                                   c->pbc_node, c->pbc_val, c->pres);
        }
    }

    // Solver::updateVelocityMagnitude (Solver.cpp:689-699)
    __global__ void k_velmag(int N, const double* __restrict__ unkno, double* __restrict__ vel)
    {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= N) return;
        const double u = unkno[3 * size_t(i)], v = unkno[3 * size_t(i) + 1], w = unkno[3 * size_t(i) + 2];
        vel[i] = sqrt(u * u + v * v + w * w + 1.0e-16);
    }

    void launch_velocity_magnitude(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_velmag, dim3(grid_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->unkno, c->velocity);
    }

    // ------------------------------------------------------------------
    // Convergence::evaluate (Convergence.cpp:78-349)
    // ------------------------------------------------------------------
    __device__ inline double safe_inverse_weight(double x)
    {
        if (fabs(x) <= 1.0e-300 || !isfinite(x)) return 1.0;
        return 1.0 / x;
    }

    __global__ __launch_bounds__(kBlock) void k_conv(BlockMap bm, int temp_calc, const uint8_t* __restrict__ nflag,
                                                      const double* __restrict__ unkno,
                                                      const double* __restrict__ unkn1,
                                                      const double* __restrict__ pres,
                                                      const double* __restrict__ pres1,
                                                      const double* __restrict__ temp,
                                                      const double* __restrict__ temp1,
                                                      const double* __restrict__ elcoe2,
                                                      const double* __restrict__ elcoe2p,
                                                      double* __restrict__ part, DevScalars* s)
    {
        __shared__ double lds[15][kBlock / 64];
        double v[15];
        for (int q = 0; q < 15; ++q) v[q] = 0.0;
        for (int i = bm.beg[blockIdx.x] + int(threadIdx.x); i < bm.end[blockIdx.x]; i += kBlock)
        {
            const bool tf = nflag[i] & 1u, ts = nflag[i] & 2u;
            if (!tf && !ts) raise_error(s, err_key(ST_CONV, 1, uint32_t(i + 1), 1));
            if (tf && !ts)
            {
                const double wgt = safe_inverse_weight(elcoe2[i]);
                for (int d = 0; d < 3; ++d)
                {
                    const double nv = unkno[3 * size_t(i) + d];
                    const double delta = nv - unkn1[3 * size_t(i) + d];
                    v[3 * d] += delta * delta;
                    v[3 * d + 1] += nv * nv;
                    const double rs = delta * wgt;
                    v[3 * d + 2] += rs * rs;
                }
            }
            if (tf)
            {
                const double delta = pres[i] - pres1[i];
                v[9] += delta * delta;
                v[10] += pres[i] * pres[i];
                v[11] += delta * delta;
            }
            if (temp_calc > 0 && (tf || ts))
            {
                const double delta = temp[i] - temp1[i];
                v[12] += delta * delta;
                v[13] += temp[i] * temp[i];
                const double rs = delta * safe_inverse_weight(elcoe2p[i]);
                v[14] += rs * rs;
            }
        }
        block_sum<15>(v, lds);
        if (threadIdx.x == 0)
            for (int q = 0; q < 15; ++q) part[blockIdx.x * 16 + q] = v[q];
    }

    // Convergence::evaluate's norms from the block partials of one segment
    // (Convergence.cpp:276-349); one block per segment.
    __global__ __launch_bounds__(kBlock) void k_conv_fin(const double* __restrict__ part,
                                                          const int32_t* __restrict__ seg_ptr, int temp_calc,
                                                          DevScalars* sbase)
    {
        __shared__ double lds[1][kBlock / 64];
        const int b0 = seg_ptr[blockIdx.x], nb = seg_ptr[blockIdx.x + 1] - b0;
        DevScalars* s = sbase + blockIdx.x;
        double hb[15];
        for (int q = 0; q < 15; ++q) hb[q] = reduce_partials(part + size_t(b0) * 16, nb, 16, q, lds);
        if (threadIdx.x == 0)
        {
            const double eps = 1.0e-30;
            for (int b = 0; b <= 6; b += 3)
            {
                s->hb[b] = sqrt(hb[b] / (hb[b + 1] + eps));
                s->hb[b + 1] = sqrt(hb[b + 1]);
                s->hb[b + 2] = sqrt(hb[b + 2]);
            }
            s->hb[9] = sqrt(hb[9] / (hb[10] + eps));
            s->hb[10] = sqrt(hb[10]);
            s->hb[11] = sqrt(hb[11]);
            if (temp_calc > 0)
            {
                s->hb[12] = sqrt(hb[12] / (hb[13] + eps));
                s->hb[13] = sqrt(hb[13]);
                s->hb[14] = sqrt(hb[14]);
            }
            else
            {
                s->hb[12] = s->hb[13] = s->hb[14] = 0.0;
            }
        }
    }

    void launch_convergence(cbs_hip_ctx* c)
    {
        hipLaunchKernelGGL(k_conv, dim3(c->nblk.n), dim3(kBlock), 0, c->stream, c->nblk.map(), c->cfg.temp_calc,
                           c->nflag, c->unkno, c->unkn1, c->pres, c->pres1, c->temp, c->temp1, c->elcoe2, c->elcoe2p,
                           c->partial, c->scal);
        hipLaunchKernelGGL(k_conv_fin, dim3(c->nseg), dim3(kBlock), 0, c->stream, c->partial, c->nblk.seg_ptr,
                           c->cfg.temp_calc, c->scal);
    }
}
