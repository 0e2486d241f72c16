/*
 * CPU oracle for the CBS time-step hot path.  TEST INFRASTRUCTURE ONLY:
 * used by the tests, the smoke check and bench.py's CPU-baseline leg as the
 * checker; the product path never calls it.
 *
 * Restates the reference's per-step algorithm with its floating-point
 * operation order (see cbs_oracle.h).  Citations are reference file:line.
 * Every binary expression below keeps the reference's left-to-right grouping;
 * compile with -ffp-contract=off so no multiply-add is fused.
 *
 * Built with -fopenmp (oracle/Makefile: libcbs_oracle_omp.so, the CPU
 * baseline's timing variant), the loops the reference runs in parallel run
 * in parallel with the reference's pragmas (ORC_OMP below): the Step 1 and
 * apply colour loops, the apply's diagonal, the CG dot products, initial
 * residual, update and direction, and the Step 4 update.  Everything else
 * stays serial, as in the reference.  At one thread that build gives the
 * same bits as the serial one; with more, the dot products' reduction order
 * (and so the last bits of the solve) depends on the thread count, as the
 * reference's does (ConjugateGradient.cpp:245).
 */
#include "cbs_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double Real;

#ifdef _OPENMP
#define ORC_OMP(x) _Pragma(x)
#else
#define ORC_OMP(x)
#endif

#define IN(s, ie, a) ((s)->intma[((size_t)(ie) - 1) * 4 + (size_t)(a) - 1])
#define DN(s, ie, d, a) ((s)->dNkdx[((size_t)(ie) - 1) * 12 + ((size_t)(d) - 1) * 4 + (size_t)(a) - 1])
#define ANN(s, c, is, ie) ((s)->annxf[(((size_t)(ie) - 1) * 4 + (size_t)(is) - 1) * 4 + (size_t)(c) - 1])
#define FE(s, is, ie) ((s)->fedge[((size_t)(ie) - 1) * 4 + (size_t)(is) - 1])
#define V3(arr, i, ip) ((arr)[((size_t)(ip) - 1) * 3 + (size_t)(i) - 1])
#define A1(arr, i) ((arr)[(size_t)(i) - 1])
#define ISD(s, k, ib) ((s)->iside[((size_t)(ib) - 1) * 6 + (size_t)(k) - 1])
#define FNRM(s, k, ib) ((s)->face_norm[((size_t)(ib) - 1) * 4 + (size_t)(k) - 1])

/* Local face f is opposite local node f (CBSStateSI.hpp:342-365). */
static const int IPPN1[5][4] = {{0, 0, 0, 0}, {0, 3, 2, 4}, {0, 3, 4, 1}, {0, 1, 4, 2}, {0, 1, 2, 3}};

/* P1 integration constants (RunConfig.hpp:303-322). */
static const Real FCON1 = 1.0 / 24.0, FCON2 = 1.0 / 12.0, FCON3 = 1.0 / 6.0, FCON4 = 1.0 / 3.0;
static const Real FDIF1 = 1.0 / 6.0, FDIF2 = 1.0 / 3.0;
static const Real MASS_FACTOR = 1.0 / 24.0;

enum {
    BC_500 = 500, BC_501 = 501, BC_502 = 502, BC_503 = 503, BC_504 = 504, BC_506 = 506,
    BC_507 = 507, BC_508 = 508, BC_510 = 510, BC_511 = 511, BC_520 = 520, BC_530 = 530,
    BC_532 = 532, BC_901 = 901, BC_902 = 902
};

static char g_err[512];

const char *orc_last_error(void) { return g_err; }

static int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return 1;
}

/* std::min / std::max / std::clamp semantics (<algorithm>). */
static inline Real rmin(Real a, Real b) { return (b < a) ? b : a; }
static inline Real rmax(Real a, Real b) { return (a < b) ? b : a; }
static inline Real rclamp(Real v, Real lo, Real hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
static inline int finite(Real v) { return isfinite(v); }

static int is_fluid(const orc_state *s, int ie) { return A1(s->mat_elem, ie) == 0; }

/* touches_fluid / touches_solid node masks (e.g. TimeStep.cpp:213-242). */
static void touch_masks(const orc_state *s, char *tf, char *ts)
{
    memset(tf, 0, (size_t)s->npoin + 1);
    memset(ts, 0, (size_t)s->npoin + 1);
    for (int ie = 1; ie <= s->nelem; ++ie) {
        const int fluid = is_fluid(s, ie);
        for (int in = 1; in <= 4; ++in) {
            const int ip = IN(s, ie, in);
            if (fluid) tf[ip] = 1; else ts[ip] = 1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Coloring::build (include/cbs/parallel/Coloring.hpp:27-97)                 */
/* ------------------------------------------------------------------------ */
int32_t orc_coloring(int32_t nelem, int32_t npoin, const int32_t *intma,
                     int32_t *color_ptr, int32_t *color_elem)
{
    int32_t *cnt = calloc((size_t)npoin + 2, sizeof(int32_t));
    int32_t *n2e_ptr = calloc((size_t)npoin + 2, sizeof(int32_t));
    int32_t *fill_ = calloc((size_t)npoin + 2, sizeof(int32_t));
    int32_t *n2e = malloc(sizeof(int32_t) * (size_t)nelem * 4 + 4);
    int32_t *color = malloc(sizeof(int32_t) * ((size_t)nelem + 1));
    int32_t *stamp = malloc(sizeof(int32_t) * ((size_t)nelem + 2));
    int32_t ncolor = 0;
    for (int ie = 1; ie <= nelem; ++ie)
        for (int in = 0; in < 4; ++in) ++cnt[intma[(size_t)(ie - 1) * 4 + in]];
    for (int ip = 1; ip <= npoin; ++ip) n2e_ptr[ip + 1] = n2e_ptr[ip] + cnt[ip];
    for (int ie = 1; ie <= nelem; ++ie)
        for (int in = 0; in < 4; ++in) {
            const int ip = intma[(size_t)(ie - 1) * 4 + in];
            n2e[n2e_ptr[ip] + fill_[ip]] = ie;
            ++fill_[ip];
        }
    for (int ie = 0; ie <= nelem; ++ie) color[ie] = -1;
    for (int c = 0; c < nelem + 2; ++c) stamp[c] = -1;
    for (int ie = 1; ie <= nelem; ++ie) {
        for (int in = 0; in < 4; ++in) {
            const int ip = intma[(size_t)(ie - 1) * 4 + in];
            for (int k = n2e_ptr[ip]; k < n2e_ptr[ip + 1]; ++k) {
                const int cj = color[n2e[k]];
                if (cj >= 0) stamp[cj] = ie;
            }
        }
        int c = 0;
        while (stamp[c] == ie) ++c;
        color[ie] = c;
        if (c + 1 > ncolor) ncolor = c + 1;
    }
    memset(color_ptr, 0, sizeof(int32_t) * ((size_t)ncolor + 1));
    for (int ie = 1; ie <= nelem; ++ie) ++color_ptr[color[ie] + 1];
    for (int c = 0; c < ncolor; ++c) color_ptr[c + 1] += color_ptr[c];
    int32_t *off = malloc(sizeof(int32_t) * ((size_t)ncolor + 1));
    memcpy(off, color_ptr, sizeof(int32_t) * ((size_t)ncolor + 1));
    for (int ie = 1; ie <= nelem; ++ie) color_elem[off[color[ie]]++] = ie;
    free(off); free(cnt); free(n2e_ptr); free(fill_); free(n2e); free(color); free(stamp);
    return ncolor;
}

/* ------------------------------------------------------------------------ */
/* PressureAssembly::buildElementPressureTerms (PressureAssembly.cpp:97-178) */
/* ------------------------------------------------------------------------ */
static Real grad_dot(const orc_state *s, int ie, int a, int b)
{
    return DN(s, ie, 1, a) * DN(s, ie, 1, b) + DN(s, ie, 2, a) * DN(s, ie, 2, b)
         + DN(s, ie, 3, a) * DN(s, ie, 3, b);
}

int orc_build_element_pressure_terms(orc_state *s)
{
    memset(s->pdiagE, 0, sizeof(Real) * 4 * (size_t)s->nelem);
    memset(s->gstifE, 0, sizeof(Real) * 6 * (size_t)s->nelem);
    for (int ie = 1; ie <= s->nelem; ++ie) {
        if (!is_fluid(s, ie)) continue;
        const Real dj = A1(s->detJ, ie);
        if (dj <= 0.0 || !finite(dj))
            return fail("PressureAssembly::buildElementPressureTerms - invalid detJ at element %d", ie);
        const Real volume = dj / 6.0;
        for (int a = 1; a <= 4; ++a) s->pdiagE[(size_t)(ie - 1) * 4 + a - 1] = volume * grad_dot(s, ie, a, a);
        Real *g = s->gstifE + (size_t)(ie - 1) * 6;
        g[0] = volume * grad_dot(s, ie, 1, 2);
        g[1] = volume * grad_dot(s, ie, 1, 3);
        g[2] = volume * grad_dot(s, ie, 1, 4);
        g[3] = volume * grad_dot(s, ie, 2, 3);
        g[4] = volume * grad_dot(s, ie, 2, 4);
        g[5] = volume * grad_dot(s, ie, 3, 4);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* TimeStep (src/timestep/TimeStep.cpp:56-656)                               */
/* ------------------------------------------------------------------------ */
static Real node_speed(const orc_state *s, int ip)
{
    return sqrt(V3(s->unkno, 1, ip) * V3(s->unkno, 1, ip) + V3(s->unkno, 2, ip) * V3(s->unkno, 2, ip)
              + V3(s->unkno, 3, ip) * V3(s->unkno, 3, ip));
}

static Real safe_speed(Real v) { return rmax(v, 1.0e-14); }

static int standard_element_length(const orc_state *s, int ie, Real *h)
{
    *h = A1(s->alen_e, ie);
    if (*h <= 0.0 || !finite(*h))
        return fail("TimeStep - invalid alen_e at element %d. Run Preprocess::elementSize before timestep computation.", ie);
    return 0;
}

static int timestep_length(const orc_state *s, int ie, Real *h)
{
    if (s->htype == 1 || s->htype == 3) return standard_element_length(s, ie, h);
    if (s->htype == 2) {
        Real hs;
        if (standard_element_length(s, ie, &hs)) return 1;
        /* nodal_length_average (TimeStep.cpp:103-125) */
        Real ha = 0.0;
        int fallback = 0;
        for (int in = 1; in <= 4; ++in) {
            const int ip = IN(s, ie, in);
            if (A1(s->alen, ip) <= 0.0 || !finite(A1(s->alen, ip))) { fallback = 1; break; }
            ha += A1(s->alen, ip);
        }
        ha = fallback ? hs : ha / 4.0;
        *h = rmin(hs, ha);
        return 0;
    }
    return fail("TimeStep - htype must be 1, 2, or 3");
}

static Real timestep_speed(const orc_state *s, int ie)
{
    if (s->htype == 3) {
        Real u = 0.0, v = 0.0, w = 0.0;
        for (int in = 1; in <= 4; ++in) {
            const int ip = IN(s, ie, in);
            u += V3(s->unkno, 1, ip);
            v += V3(s->unkno, 2, ip);
            w += V3(s->unkno, 3, ip);
        }
        u /= 4.0; v /= 4.0; w /= 4.0;
        return sqrt(u * u + v * v + w * w);
    }
    Real vmax = 0.0;
    for (int in = 1; in <= 4; ++in) vmax = rmax(vmax, node_speed(s, IN(s, ie, in)));
    return vmax;
}

static int element_diffusivity(const orc_state *s, int ie, Real *out)
{
    Real diff = s->ani;
    if (s->dimensional_mode > 0 && s->material_properties_enabled > 0) {
        const Real al = A1(s->alpha_e, ie);
        if (al <= 0.0 || !finite(al))
            return fail("TimeStep - invalid thermal diffusivity alpha_e at element %d", ie);
        diff = al;
        if (is_fluid(s, ie)) {
            if (A1(s->rho_e, ie) <= 0.0 || A1(s->mu_e, ie) < 0.0)
                return fail("TimeStep - invalid fluid rho/mu at element %d", ie);
            const Real nu = A1(s->mu_e, ie) / A1(s->rho_e, ie);
            diff = rmax(diff, nu);
        }
    }
    if (diff <= 0.0 || !finite(diff))
        return fail("TimeStep - invalid timestep diffusivity at element %d", ie);
    *out = diff;
    return 0;
}

static int local_stability_dt(const orc_state *s, int ie, Real h, Real speed, Real *out)
{
    Real diffusivity = 0.0;
    if (element_diffusivity(s, ie, &diffusivity)) return 1;
    const Real advective_dt = is_fluid(s, ie) ? h / safe_speed(speed) : DBL_MAX;
    const Real diffusive_dt = h * h * (0.5 / diffusivity);
    Real dt = s->csafm * rmin(advective_dt, diffusive_dt);
    if (s->ilots == 2 && s->deltr > 0.0 && finite(s->deltr)) dt = rmin(dt, s->deltr);
    if (dt <= 0.0 || !finite(dt))
        return fail("TimeStep - invalid local stability timestep at element %d", ie);
    *out = dt;
    return 0;
}

static int validate_common_inputs(const orc_state *s)
{
    if (s->nelem < 1 || s->npoin < 1) return fail("TimeStep - problem sizes are not initialised");
    if (s->csafm <= 0.0 || !finite(s->csafm)) return fail("TimeStep - csafm must be positive");
    if (s->dtfix <= 0.0 || !finite(s->dtfix)) return fail("TimeStep - dtfix must be positive");
    if (s->htype < 1 || s->htype > 3) return fail("TimeStep - htype must be 1, 2, or 3");
    return 0;
}

/* computeNodalLocalTimeStep (TimeStep.cpp:600-656) */
static int nodal_local_timestep(orc_state *s)
{
    if (validate_common_inputs(s)) return 1;
    for (int ip = 1; ip <= s->npoin; ++ip) {
        const Real al = A1(s->alen, ip);
        const Real h = (al > 0.0 && finite(al)) ? al : 0.0;
        const Real speed = safe_speed(node_speed(s, ip));
        if (h <= 0.0) {
            A1(s->deltp, ip) = s->dtfix;
            A1(s->deltp2, ip) = s->dtfix;
            A1(s->beta1, ip) = 1.0;
            A1(s->vvis, ip) = 0.0;
            continue;
        }
        const Real advective_dt = h / speed;
        const Real diffusive_dt = h * h * (0.5 / s->ani);
        A1(s->vvis, ip) = diffusive_dt;
        Real dt = s->csafm * rmin(advective_dt, diffusive_dt);
        if (s->ilots == 2 && s->deltr > 0.0 && finite(s->deltr)) dt = rmin(dt, s->deltr);
        if (dt <= 0.0 || !finite(dt))
            return fail("TimeStep::computeNodalLocalTimeStep - invalid nodal timestep at node %d", ip);
        A1(s->deltp, ip) = dt;
        A1(s->deltp2, ip) = dt;
        if (s->beta_opt > 0) {
            const Real denom = advective_dt + diffusive_dt + 1.0e-30;
            A1(s->beta1, ip) = rclamp(diffusive_dt / denom, 0.0, 1.0);
        } else {
            A1(s->beta1, ip) = 1.0;
        }
    }
    return 0;
}

/* TimeStep::computeTimeStep (TimeStep.cpp:473-583) */
int orc_compute_timestep(orc_state *s, int32_t iitime)
{
    if (validate_common_inputs(s)) return 1;
    if (s->cbs_scheme != 1 && s->cbs_scheme != 0)
        return fail("TimeStep::computeTimeStep - cbs_scheme must be 1 semi-implicit or 0 explicit");
    if (s->cbs_scheme == 0)
        return fail("Fully explicit CBS3D mode was requested, but explicit CBS3D step kernels are not yet ported.");
    if (s->ilots <= -1 || s->dtfixed > 0 || (s->dtfix_end > 0 && iitime <= s->dtfix_end)) {
        for (int ie = 1; ie <= s->nelem; ++ie) A1(s->delte, ie) = s->dtfix;
        for (int ip = 1; ip <= s->npoin; ++ip) {
            A1(s->deltp, ip) = s->dtfix;
            A1(s->deltp2, ip) = s->dtfix;
            A1(s->beta1, ip) = 1.0;
        }
        s->dtreal = s->dtfix;
        return 0;
    }
    /* compute_velocity_helpers (:364-384) */
    for (int ip = 1; ip <= s->npoin; ++ip) {
        const Real speed = node_speed(s, ip);
        A1(s->velcp, ip) = speed;
        if (speed > 1.0e-14) {
            V3(s->unkno_unit_vec, 1, ip) = V3(s->unkno, 1, ip) / speed;
            V3(s->unkno_unit_vec, 2, ip) = V3(s->unkno, 2, ip) / speed;
            V3(s->unkno_unit_vec, 3, ip) = V3(s->unkno, 3, ip) / speed;
        } else {
            V3(s->unkno_unit_vec, 1, ip) = 0.0;
            V3(s->unkno_unit_vec, 2, ip) = 0.0;
            V3(s->unkno_unit_vec, 3, ip) = 0.0;
        }
    }
    /* compute_nodal_lengths_from_elements (:319-355) */
    for (int ip = 1; ip <= s->npoin; ++ip) A1(s->alen, ip) = DBL_MAX;
    for (int ie = 1; ie <= s->nelem; ++ie) {
        const Real dj = A1(s->detJ, ie);
        if (dj <= 0.0 || !finite(dj)) return fail("TimeStep - invalid detJ in nodal length; run Preprocess first");
        const Real volum = 0.5 * dj;
        for (int in = 1; in <= 4; ++in) {
            const int ip = IN(s, ie, in);
            const Real face_area = ANN(s, 4, in, ie);
            if (face_area <= 0.0 || !finite(face_area))
                return fail("TimeStep - invalid face area in nodal length; run Preprocess::getNormals first");
            const Real hite = volum / face_area;
            A1(s->alen, ip) = rmin(A1(s->alen, ip), hite);
        }
    }
    if (nodal_local_timestep(s)) return 1;
    Real dt_min = DBL_MAX;
    for (int ie = 1; ie <= s->nelem; ++ie) {
        Real h = 0.0, dt = 0.0;
        if (timestep_length(s, ie, &h)) return 1;
        const Real speed = timestep_speed(s, ie);
        if (local_stability_dt(s, ie, h, speed, &dt)) return 1;
        A1(s->delte, ie) = dt;
        dt_min = rmin(dt_min, dt);
    }
    if (dt_min <= 0.0 || !finite(dt_min))
        return fail("TimeStep::computeTimeStep - invalid global minimum timestep");
    if (s->transient_on > 0 || s->solver_opt > 1 || s->dtfixed == -1 || (s->ilots != 1 && s->ilots != 2)) {
        for (int ie = 1; ie <= s->nelem; ++ie) A1(s->delte, ie) = dt_min;
        for (int ip = 1; ip <= s->npoin; ++ip) {
            A1(s->deltp, ip) = dt_min;
            A1(s->deltp2, ip) = dt_min;
        }
    }
    if (s->rem_deltp > 0) {
        for (int ip = 1; ip <= s->npoin; ++ip) {
            if (A1(s->deltp1, ip) > 0.0) {
                A1(s->deltp, ip) = rmin(A1(s->deltp, ip), A1(s->deltp1, ip));
                A1(s->deltp2, ip) = rmin(A1(s->deltp2, ip), A1(s->deltp1, ip));
            }
            A1(s->deltp1, ip) = A1(s->deltp, ip);
        }
    } else {
        for (int ip = 1; ip <= s->npoin; ++ip) A1(s->deltp1, ip) = A1(s->deltp, ip);
    }
    s->dtreal = dt_min;
    return 0;
}

/* TimeStep::applyStep2PressureTimeStepCorrection (TimeStep.cpp:704-749) */
static int step2_timestep_correction(orc_state *s)
{
    if (s->step2_check < 1) return 0;
    if (s->csafm2 <= 0.0 || !finite(s->csafm2))
        return fail("TimeStep::applyStep2PressureTimeStepCorrection - csafm2 must be positive");
    for (int ie = 1; ie <= s->nelem; ++ie) {
        Real dt_face_min = A1(s->delte, ie);
        for (int in = 1; in <= 4; ++in) {
            const int ip = IN(s, ie, in);
            const Real pscale = fabs(A1(s->pres1, ip)) + s->epsilon1;
            const Real speed = node_speed(s, ip);
            if (speed > 1.0e-14 && pscale > 0.0) {
                Real h;
                if (A1(s->alen, ip) > 0.0) h = A1(s->alen, ip);
                else if (standard_element_length(s, ie, &h)) return 1;
                const Real candidate = s->csafm2 * h * pscale / speed;
                if (candidate > 0.0 && finite(candidate)) dt_face_min = rmin(dt_face_min, candidate);
            }
        }
        A1(s->delte, ie) = dt_face_min;
    }
    return 0;
}

/* TimeStep::updateLhsDiagonal (TimeStep.cpp:778-954) */
int orc_update_lhs_diagonal(orc_state *s)
{
    const int N = s->npoin;
    char *tf = malloc((size_t)N + 1), *ts = malloc((size_t)N + 1);
    int rc = 0;
    memset(s->elcoe2, 0, sizeof(Real) * (size_t)N);
    memset(s->elcoe2p, 0, sizeof(Real) * (size_t)N);
    memset(s->pdiag, 0, sizeof(Real) * (size_t)N);
    touch_masks(s, tf, ts);
    for (int ie = 1; ie <= s->nelem && !rc; ++ie) {
        const Real dt = A1(s->delte, ie);
        if (dt <= 0.0 || !finite(dt)) { rc = fail("TimeStep::updateLhsDiagonal - non-positive local timestep at element %d", ie); break; }
        const int fluid = is_fluid(s, ie);
        for (int in = 1; in <= 4; ++in) {
            const int ip = IN(s, ie, in);
            const Real mass = s->elcoe_e[(size_t)(ie - 1) * 4 + in - 1];
            if (mass <= 0.0 || !finite(mass)) { rc = fail("TimeStep::updateLhsDiagonal - invalid element-node mass"); break; }
            if (fluid) A1(s->elcoe2, ip) += mass / dt;
            A1(s->elcoe2p, ip) += A1(s->rho_cp_e, ie) * mass / dt;
        }
    }
    for (int ip = 1; ip <= N && !rc; ++ip) {
        const int fluid = tf[ip], solid = ts[ip];
        if (!fluid && !solid) { rc = fail("TimeStep::updateLhsDiagonal - orphan node detected at node %d", ip); break; }
        if (fluid) {
            if (A1(s->elcoe2, ip) <= 0.0 || !finite(A1(s->elcoe2, ip))) {
                rc = fail("TimeStep::updateLhsDiagonal - non-positive fluid momentum mass/time diagonal at node %d", ip);
                break;
            }
            A1(s->elcoe2, ip) = 1.0 / A1(s->elcoe2, ip);
        } else {
            A1(s->elcoe2, ip) = 0.0;
        }
        if (A1(s->elcoe2p, ip) <= 0.0 || !finite(A1(s->elcoe2p, ip))) {
            rc = fail("TimeStep::updateLhsDiagonal - non-positive capacitance/time diagonal at node %d", ip);
            break;
        }
        A1(s->elcoe2p, ip) = 1.0 / A1(s->elcoe2p, ip);
    }
    for (int ie = 1; ie <= s->nelem && !rc; ++ie) {
        const Real dt = A1(s->delte, ie);
        Real *g = s->gstif + (size_t)(ie - 1) * 6;
        if (!is_fluid(s, ie)) {
            for (int ig = 0; ig < 6; ++ig) g[ig] = 0.0;
            continue;
        }
        for (int ig = 0; ig < 6; ++ig) g[ig] = s->gstifE[(size_t)(ie - 1) * 6 + ig] * dt;
        for (int in = 1; in <= 4; ++in)
            A1(s->pdiag, IN(s, ie, in)) += s->pdiagE[(size_t)(ie - 1) * 4 + in - 1] * dt;
    }
    for (int ip = 1; ip <= N && !rc; ++ip) {
        if (tf[ip]) {
            if (fabs(A1(s->pdiag, ip)) <= 1.0e-14 || !finite(A1(s->pdiag, ip))) {
                rc = fail("TimeStep::updateLhsDiagonal - zero pressure diagonal on fluid-connected node %d", ip);
                break;
            }
        } else if (ts[ip]) {
            A1(s->pdiag, ip) = 0.0;
            A1(s->pres, ip) = 0.0;
            A1(s->rhs1, ip) = 0.0;
        }
    }
    for (int i = 1; i <= s->bc_fixed && !rc; ++i) {
        const int ip = A1(s->bc_list, i);
        if (ip < 1 || ip > N) { rc = fail("TimeStep::updateLhsDiagonal - fixed pressure node out of range"); break; }
        if (!tf[ip]) {
            rc = fail("TimeStep::updateLhsDiagonal - fixed pressure node is not connected to the fluid pressure domain, node %d", ip);
            break;
        }
        A1(s->pdiag, ip) = 1.0;
        A1(s->pres, ip) = A1(s->bc_values, i);
        A1(s->rhs1, ip) = A1(s->bc_values, i);
    }
    free(tf); free(ts);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Boundary (src/boundary/Boundary.cpp:94-723) and the Steps package         */
/* ------------------------------------------------------------------------ */
static int supported_bc(int bc)
{
    switch (bc) {
    case BC_500: case BC_501: case BC_502: case BC_503: case BC_504: case BC_506: case BC_507:
    case BC_508: case BC_510: case BC_511: case BC_520: case BC_530: case BC_532: case BC_901:
    case BC_902: return 1;
    default: return 0;
    }
}

static int validate_boundary_flags(const orc_state *s)
{
    for (int ib = 1; ib <= s->nboun; ++ib) {
        const int bc = ISD(s, 6, ib);
        if (!supported_bc(bc))
            return fail("Boundary - unsupported BC_ID %d at boundary face %d. This CBS3D boundary module supports the CBS2D++_SI BC family adapted to 3D.", bc, ib);
    }
    return 0;
}

static int unit_outward_normal(const orc_state *s, int ib, Real *nx, Real *ny, Real *nz)
{
    const Real area = FNRM(s, 4, ib);
    if (area <= 0.0 || !finite(area))
        return fail("Boundary - invalid boundary-face area in face_norm(4,ib). Preprocess::getNormals must be called before boundary application.");
    *nx = FNRM(s, 1, ib) / area;
    *ny = FNRM(s, 2, ib) / area;
    *nz = FNRM(s, 3, ib) / area;
    return 0;
}

static void set_velocity(orc_state *s, int ip, Real u, Real v, Real w)
{
    V3(s->unkno, 1, ip) = u;
    V3(s->unkno, 2, ip) = v;
    V3(s->unkno, 3, ip) = w;
}

/* Boundary.cpp:164-178; std::pow(y, n) with an int exponent calls pow(double, double). */
static Real bfs_profile_u(Real y)
{
    return 0.6624 * pow(y, 6) - 7.5547 * pow(y, 5) + 33.9 * pow(y, 4) - 75.283 * pow(y, 3)
         + 83.368 * pow(y, 2) - 37.793 * y + 2.6959;
}

static Real rectangular_profile_u(Real y) { return 6.0 * y * (1.0 - y); }

static int node_range(const orc_state *s, int ip, const char *ctx)
{
    if (ip < 1 || ip > s->npoin) return fail("%s - node index out of range", ctx);
    return 0;
}

/* Boundary::applySymmetry (Boundary.cpp:275-310) */
static int apply_symmetry(orc_state *s)
{
    if (validate_boundary_flags(s)) return 1;
    for (int ib = 1; ib <= s->nboun; ++ib) {
        if (ISD(s, 6, ib) != BC_506) continue;
        Real nx, ny, nz;
        if (unit_outward_normal(s, ib, &nx, &ny, &nz)) return 1;
        for (int in = 1; in <= 3; ++in) {
            const int ip = ISD(s, in, ib);
            if (node_range(s, ip, "Boundary::applySymmetry")) return 1;
            const Real un = V3(s->unkno, 1, ip) * nx + V3(s->unkno, 2, ip) * ny + V3(s->unkno, 3, ip) * nz;
            V3(s->unkno, 1, ip) -= un * nx;
            V3(s->unkno, 2, ip) -= un * ny;
            V3(s->unkno, 3, ip) -= un * nz;
        }
    }
    return 0;
}

/* Boundary::applyVelocity (Boundary.cpp:341-508) */
int orc_apply_velocity(orc_state *s)
{
    if (validate_boundary_flags(s)) return 1;
    for (int ib = 1; ib <= s->nboun; ++ib) {
        const int bc = ISD(s, 6, ib);
        Real nx = 0.0, ny = 0.0, nz = 0.0;
        if (bc == BC_511 && unit_outward_normal(s, ib, &nx, &ny, &nz)) return 1;
        for (int in = 1; in <= 3; ++in) {
            const int ip = ISD(s, in, ib);
            if (node_range(s, ip, "Boundary::applyVelocity")) return 1;
            if (bc == BC_500) set_velocity(s, ip, s->inlet_u, s->inlet_v, s->inlet_w);
            else if (bc == BC_501 || bc == BC_502) set_velocity(s, ip, 0.0, 0.0, 0.0);
            else if (bc == BC_503) set_velocity(s, ip, 1.0, 0.0, 0.0);
            else if (bc == BC_504 || bc == BC_520) continue;
            else if (bc == BC_506) continue;
            else if (bc == BC_507) set_velocity(s, ip, bfs_profile_u(V3(s->coord, 2, ip)), 0.0, 0.0);
            else if (bc == BC_508) set_velocity(s, ip, rectangular_profile_u(V3(s->coord, 2, ip)), 0.0, 0.0);
            else if (bc == BC_510) set_velocity(s, ip, s->inlet_u, s->inlet_v, s->inlet_w);
            else if (bc == BC_511)
                set_velocity(s, ip, -s->inlet_u_from_massflow * nx, -s->inlet_u_from_massflow * ny,
                             -s->inlet_u_from_massflow * nz);
            else if (bc == BC_530 || bc == BC_532 || bc == BC_901) set_velocity(s, ip, 0.0, 0.0, 0.0);
        }
    }
    for (int iw = 1; iw <= s->npoin_wall; ++iw) {
        const int ip = A1(s->wall_node_list, iw);
        if (ip >= 1 && ip <= s->npoin) set_velocity(s, ip, 0.0, 0.0, 0.0);
    }
    for (int ib = 1; ib <= s->nboun; ++ib) {
        if (ISD(s, 6, ib) != BC_500) continue;
        for (int in = 1; in <= 3; ++in) set_velocity(s, ISD(s, in, ib), s->inlet_u, s->inlet_v, s->inlet_w);
    }
    /* enforce_zero_velocity_on_material_solid_touch_nodes (:250-272) */
    char *tf = malloc((size_t)s->npoin + 1), *ts = malloc((size_t)s->npoin + 1);
    touch_masks(s, tf, ts);
    for (int ip = 1; ip <= s->npoin; ++ip)
        if (ts[ip]) set_velocity(s, ip, 0.0, 0.0, 0.0);
    free(tf); free(ts);
    return 0;
}

/* Boundary::applyOutletBackflowControl (Boundary.cpp:672-723) */
static int apply_outlet_backflow(orc_state *s)
{
    if (validate_boundary_flags(s)) return 1;
    char *tf = malloc((size_t)s->npoin + 1), *ts = malloc((size_t)s->npoin + 1);
    int rc = 0;
    touch_masks(s, tf, ts);
    for (int ib = 1; ib <= s->nboun && !rc; ++ib) {
        if (ISD(s, 6, ib) != BC_520) continue;
        Real nx, ny, nz;
        if ((rc = unit_outward_normal(s, ib, &nx, &ny, &nz))) break;
        for (int in = 1; in <= 3; ++in) {
            const int ip = ISD(s, in, ib);
            if ((rc = node_range(s, ip, "Boundary::applyOutletBackflowControl"))) break;
            if (ts[ip]) { set_velocity(s, ip, 0.0, 0.0, 0.0); continue; }
            const Real un = V3(s->unkno, 1, ip) * nx + V3(s->unkno, 2, ip) * ny + V3(s->unkno, 3, ip) * nz;
            if (un < 0.0) {
                V3(s->unkno, 1, ip) -= un * nx;
                V3(s->unkno, 2, ip) -= un * ny;
                V3(s->unkno, 3, ip) -= un * nz;
            }
        }
    }
    free(tf); free(ts);
    return rc;
}

/* Steps.cpp:102-121 */
static void enforce_zero_velocity_in_solid(orc_state *s)
{
    for (int ie = 1; ie <= s->nelem; ++ie) {
        if (is_fluid(s, ie)) continue;
        for (int in = 1; in <= 4; ++in) set_velocity(s, IN(s, ie, in), 0.0, 0.0, 0.0);
    }
}

/* apply_velocity_boundary_package (Steps.cpp:170-176) */
int orc_apply_velocity_package(orc_state *s)
{
    if (apply_symmetry(s)) return 1;
    if (orc_apply_velocity(s)) return 1;
    if (apply_outlet_backflow(s)) return 1;
    enforce_zero_velocity_in_solid(s);
    return 0;
}

/* Boundary::applyTemperature (Boundary.cpp:529-626) */
int orc_apply_temperature(orc_state *s)
{
    if (validate_boundary_flags(s)) return 1;
    for (int ib = 1; ib <= s->nboun; ++ib) {
        const int bc = ISD(s, 6, ib);
        for (int in = 1; in <= 3; ++in) {
            const int ip = ISD(s, in, ib);
            if (node_range(s, ip, "Boundary::applyTemperature")) return 1;
            if (bc == BC_501) A1(s->temperature, ip) = 1.0;
            else if (bc == BC_502 || bc == BC_503 || bc == BC_508) A1(s->temperature, ip) = 0.0;
            else if (bc == BC_510 || bc == BC_511) A1(s->temperature, ip) = s->inlet_temperature;
        }
    }
    return 0;
}

/* Boundary::applyPressure (Boundary.cpp:642-652) */
int orc_apply_pressure(orc_state *s)
{
    for (int i = 1; i <= s->bc_fixed; ++i) {
        const int ip = A1(s->bc_list, i);
        if (node_range(s, ip, "Boundary::applyPressure")) return 1;
        A1(s->pres, ip) = A1(s->bc_values, i);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Step 1: MomentumAssembly::assembleStep1Rhs (MomentumAssembly.cpp:144-554) */
/* ------------------------------------------------------------------------ */
static int momentum_diffusivity(const orc_state *s, int ie, Real *out)
{
    if (s->dimensional_mode > 0 && s->material_properties_enabled > 0) {
        if (A1(s->rho_e, ie) <= 0.0 || A1(s->mu_e, ie) < 0.0)
            return fail("MomentumAssembly::assembleStep1Rhs - invalid fluid rho/mu at element %d", ie);
        *out = A1(s->mu_e, ie) / A1(s->rho_e, ie);
        return 0;
    }
    *out = s->ani;
    return 0;
}

/* Element Step 1 RHS into lrhs[i][a] (1-based like the reference). */
static int step1_element(const orc_state *s, int ie, Real lrhs[4][5])
{
    Real lunkno[4][5];
    Real lunk[4][4][5];
    memset(lunkno, 0, sizeof lunkno);
    memset(lunk, 0, sizeof lunk);
    memset(lrhs, 0, sizeof(Real) * 4 * 5);
    for (int a = 1; a <= 4; ++a) {
        const int ip = IN(s, ie, a);
        for (int i = 1; i <= 3; ++i) lunkno[i][a] = V3(s->unkn1, i, ip);
    }
    for (int a = 1; a <= 4; ++a)
        for (int j = 1; j <= 3; ++j)
            for (int i = 1; i <= 3; ++i) lunk[i][j][a] = lunkno[j][a] * lunkno[i][a];
    const Real detj = A1(s->detJ, ie);

    /* step1_convective_galerkin (:144-220) */
    {
        Real lunksum[4][4], lunksum_face[4][4];
        memset(lunksum, 0, sizeof lunksum);
        memset(lunksum_face, 0, sizeof lunksum_face);
        const Real volume_factor = detj * FCON1;
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j)
                for (int in = 1; in <= 4; ++in) lunksum[i][j] += lunk[i][j][in];
        for (int a = 1; a <= 4; ++a)
            for (int i = 1; i <= 3; ++i)
                for (int j = 1; j <= 3; ++j) lrhs[i][a] += DN(s, ie, j, a) * volume_factor * lunksum[i][j];
        for (int is = 1; is <= 4; ++is) {
            if (FE(s, is, ie) == 0) continue;
            for (int i = 1; i <= 3; ++i)
                for (int j = 1; j <= 3; ++j) lunksum_face[i][j] = lunksum[i][j] - lunk[i][j][is];
            for (int j = 1; j <= 3; ++j) {
                const Real face_factor = ANN(s, j, is, ie) * FCON2;
                for (int fn = 1; fn <= 3; ++fn) {
                    const int a = IPPN1[is][fn];
                    for (int i = 1; i <= 3; ++i) lrhs[i][a] -= (lunksum_face[i][j] + lunk[i][j][a]) * face_factor;
                }
            }
        }
    }
    /* step1_characteristic_correction (:242-345) */
    {
        Real lunksum[4][4], lunksumk[4][4], umean[4];
        memset(lunksum, 0, sizeof lunksum);
        memset(lunksumk, 0, sizeof lunksumk);
        memset(umean, 0, sizeof umean);
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j)
                for (int a = 1; a <= 4; ++a) lunksum[i][j] += DN(s, ie, j, a) * lunk[i][j][a];
        for (int i = 1; i <= 3; ++i) {
            for (int a = 1; a <= 4; ++a) umean[i] += lunkno[i][a];
            umean[i] /= 4.0;
        }
        for (int i = 1; i <= 3; ++i) {
            Real sum_j = 0.0;
            for (int j = 1; j <= 3; ++j) sum_j += lunksum[i][j];
            for (int k = 1; k <= 3; ++k) lunksumk[i][k] = umean[k] * sum_j;
        }
        const Real dt = A1(s->delte, ie);
        if (dt <= 0.0 || !finite(dt))
            return fail("MomentumAssembly::assembleStep1Rhs - invalid element timestep at element %d", ie);
        const Real ldelte = 0.5 * dt * detj * FCON3;
        const Real ldelte2 = 0.5 * dt * FCON4;
        for (int a = 1; a <= 4; ++a)
            for (int k = 1; k <= 3; ++k)
                for (int i = 1; i <= 3; ++i) lrhs[i][a] -= DN(s, ie, k, a) * lunksumk[i][k] * ldelte;
        for (int is = 1; is <= 4; ++is) {
            if (FE(s, is, ie) == 0) continue;
            for (int k = 1; k <= 3; ++k) {
                const Real face_factor = ANN(s, k, is, ie) * ldelte2;
                for (int fn = 1; fn <= 3; ++fn) {
                    const int a = IPPN1[is][fn];
                    for (int i = 1; i <= 3; ++i) lrhs[i][a] += lunksumk[i][k] * face_factor;
                }
            }
        }
    }
    /* step1_diffusion (:369-432) */
    {
        Real ani = 0.0;
        if (momentum_diffusivity(s, ie, &ani)) return 1;
        Real dNuidxj[4][4];
        memset(dNuidxj, 0, sizeof dNuidxj);
        for (int i = 1; i <= 3; ++i)
            for (int a = 1; a <= 4; ++a)
                for (int j = 1; j <= 3; ++j) dNuidxj[i][j] += DN(s, ie, j, a) * lunkno[i][a];
        const Real volume_factor = ani * detj * FDIF1;
        for (int i = 1; i <= 3; ++i)
            for (int a = 1; a <= 4; ++a)
                for (int j = 1; j <= 3; ++j) lrhs[i][a] -= DN(s, ie, j, a) * dNuidxj[i][j] * volume_factor;
        const Real face_diffusion_factor = ani * FDIF2;
        for (int is = 1; is <= 4; ++is) {
            if (FE(s, is, ie) == 0) continue;
            for (int j = 1; j <= 3; ++j) {
                const Real face_factor = ANN(s, j, is, ie) * face_diffusion_factor;
                for (int i = 1; i <= 3; ++i)
                    for (int fn = 1; fn <= 3; ++fn) {
                        const int a = IPPN1[is][fn];
                        lrhs[i][a] += dNuidxj[i][j] * face_factor;
                    }
            }
        }
    }
    return 0;
}

int orc_assemble_step1_rhs(orc_state *s)
{
    memset(s->rhs, 0, sizeof(Real) * 3 * (size_t)s->npoin);
    if (s->dtreal <= 0.0 || !finite(s->dtreal))
        return fail("MomentumAssembly::assembleStep1Rhs - dtreal must be positive before Step 1 assembly");
    int bad_detj = 0;
    for (int c = 0; c < s->ncolor; ++c) {
        int err = 0;
        /* MomentumAssembly.cpp:480; within a colour no two elements share a node */
        ORC_OMP("omp parallel for schedule(static) reduction(|:err, bad_detj)")
        for (int k = s->color_ptr[c]; k < s->color_ptr[c + 1]; ++k) {
            const int ie = s->color_elem[k];
            if (!is_fluid(s, ie)) continue;
            const Real dj = A1(s->detJ, ie);
            if (dj <= 0.0 || !finite(dj)) { bad_detj = 1; continue; }
            Real lrhs[4][5];
            if (step1_element(s, ie, lrhs)) { err = 1; continue; }
            for (int a = 1; a <= 4; ++a) {
                const int ip = IN(s, ie, a);
                for (int i = 1; i <= 3; ++i) V3(s->rhs, i, ip) += lrhs[i][a];
            }
        }
        if (err) return 1;
    }
    if (bad_detj) return fail("MomentumAssembly::assembleStep1Rhs - invalid detJ at one or more elements");
    return 0;
}

/* Steps::step1SemiImplicit (Steps.cpp:273-284) */
int orc_step1(orc_state *s)
{
    if (orc_assemble_step1_rhs(s)) return 1;
    for (int ip = 1; ip <= s->npoin; ++ip)
        for (int i = 1; i <= 3; ++i)
            V3(s->unkno, i, ip) = V3(s->unkn1, i, ip) + V3(s->rhs, i, ip) * A1(s->elcoe2, ip);
    return orc_apply_velocity_package(s);
}

/* ------------------------------------------------------------------------ */
/* Step 2 RHS: PressureAssembly::assembleStep2Rhs (PressureAssembly.cpp:236-335) */
/* ------------------------------------------------------------------------ */
int orc_assemble_step2_rhs(orc_state *s)
{
    if (s->dtreal <= 0.0 || !finite(s->dtreal))
        return fail("PressureAssembly::assembleStep2Rhs - dtreal must be positive before Step 2 RHS assembly");
    memset(s->rhs1, 0, sizeof(Real) * (size_t)s->npoin);
    for (int ie = 1; ie <= s->nelem; ++ie) {
        if (!is_fluid(s, ie)) continue;
        const Real dj = A1(s->detJ, ie);
        if (dj <= 0.0 || !finite(dj))
            return fail("PressureAssembly::assembleStep2Rhs - invalid detJ at element %d", ie);
        const Real vol4 = dj * FCON1;
        Real u_sum[4] = {0.0, 0.0, 0.0, 0.0};
        for (int in = 1; in <= 4; ++in) {
            const int ip = IN(s, ie, in);
            u_sum[1] += V3(s->unkno, 1, ip);
            u_sum[2] += V3(s->unkno, 2, ip);
            u_sum[3] += V3(s->unkno, 3, ip);
        }
        u_sum[1] *= vol4;
        u_sum[2] *= vol4;
        u_sum[3] *= vol4;
        Real lrhs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int in = 1; in <= 4; ++in)
            lrhs[in] = DN(s, ie, 1, in) * u_sum[1] + DN(s, ie, 2, in) * u_sum[2] + DN(s, ie, 3, in) * u_sum[3];
        for (int is = 1; is <= 4; ++is) {
            if (FE(s, is, ie) != 1) continue;
            Real face_u_sum[4] = {0.0, 0.0, 0.0, 0.0};
            for (int i = 1; i <= 3; ++i) {
                const int ip = IN(s, ie, IPPN1[is][i]);
                face_u_sum[1] += V3(s->unkn1, 1, ip);
                face_u_sum[2] += V3(s->unkn1, 2, ip);
                face_u_sum[3] += V3(s->unkn1, 3, ip);
            }
            for (int idim = 1; idim <= 3; ++idim) {
                const Real normal_factor = ANN(s, idim, is, ie) * FCON2;
                for (int i = 1; i <= 3; ++i) {
                    const int ln = IPPN1[is][i];
                    const int ip = IN(s, ie, ln);
                    lrhs[ln] -= (face_u_sum[idim] + V3(s->unkn1, idim, ip)) * normal_factor;
                }
            }
        }
        for (int in = 1; in <= 4; ++in) A1(s->rhs1, IN(s, ie, in)) += lrhs[in];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* MatrixVectorCalc::pressureMultiply (MatrixVectorCalc.cpp:64-205)          */
/* ------------------------------------------------------------------------ */
int orc_pressure_multiply(const orc_state *s, const double *x, double *y)
{
    memset(y, 0, sizeof(Real) * (size_t)s->npoin);
    for (int c = 0; c < s->ncolor; ++c) {
        ORC_OMP("omp parallel for schedule(static)")  /* MatrixVectorCalc.cpp:158 */
        for (int k = s->color_ptr[c]; k < s->color_ptr[c + 1]; ++k) {
            const int ie = s->color_elem[k];
            if (A1(s->mat_elem, ie) != 0) continue;
            const int i1 = IN(s, ie, 1), i2 = IN(s, ie, 2), i3 = IN(s, ie, 3), i4 = IN(s, ie, 4);
            const Real x1 = A1(x, i1), x2 = A1(x, i2), x3 = A1(x, i3), x4 = A1(x, i4);
            const Real *g = s->gstif + (size_t)(ie - 1) * 6;
            const Real a12 = g[0], a13 = g[1], a14 = g[2], a23 = g[3], a24 = g[4], a34 = g[5];
            A1(y, i1) += a12 * x2 + a13 * x3 + a14 * x4;
            A1(y, i2) += a12 * x1 + a23 * x3 + a24 * x4;
            A1(y, i3) += a13 * x1 + a23 * x2 + a34 * x4;
            A1(y, i4) += a14 * x1 + a24 * x2 + a34 * x3;
        }
    }
    ORC_OMP("omp parallel for schedule(static)")  /* MatrixVectorCalc.cpp:198 */
    for (int ip = 1; ip <= s->npoin; ++ip) A1(y, ip) += A1(s->pdiag, ip) * A1(x, ip);
    for (int i = 1; i <= s->bc_fixed; ++i) {
        const int ip = A1(s->bc_list, i);
        if (ip < 1 || ip > s->npoin)
            return fail("MatrixVectorCalc::pressureMultiply - fixed pressure node out of range");
        A1(y, ip) = A1(x, ip);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ConjugateGradient::solvePressure (ConjugateGradient.cpp:98-869)           */
/* ------------------------------------------------------------------------ */
typedef struct {
    char *active, *fixed; /* indexed by 1-based node id */
    Real *fixed_values;   /* 1-based */
} cg_masks;

static int is_free(const cg_masks *m, int ip) { return m->active[ip] && !m->fixed[ip]; }

static Real dot_free(const orc_state *s, const cg_masks *m, const Real *a, const Real *b)
{
    Real sum = 0.0;
    ORC_OMP("omp parallel for reduction(+:sum) schedule(static)")  /* ConjugateGradient.cpp:245 */
    for (int ip = 1; ip <= s->npoin; ++ip)
        if (is_free(m, ip)) sum += A1(a, ip) * A1(b, ip);
    return sum;
}

static Real max_abs_free(const orc_state *s, const cg_masks *m, const Real *r)
{
    Real maxv = 0.0;
    for (int ip = 1; ip <= s->npoin; ++ip)
        if (is_free(m, ip)) maxv = rmax(maxv, fabs(A1(r, ip)));
    return maxv;
}

static void zero_non_free(const orc_state *s, const cg_masks *m, Real *x)
{
    for (int ip = 1; ip <= s->npoin; ++ip)
        if (!is_free(m, ip)) A1(x, ip) = 0.0;
}

static void impose_constraints(orc_state *s, const cg_masks *m)
{
    for (int ip = 1; ip <= s->npoin; ++ip) {
        if (!m->active[ip]) {
            A1(s->pres, ip) = 0.0;
            A1(s->rhs1, ip) = 0.0;
            A1(s->pdiag, ip) = 0.0;
            continue;
        }
        if (m->fixed[ip]) {
            A1(s->pres, ip) = m->fixed_values[ip];
            A1(s->rhs1, ip) = m->fixed_values[ip];
            A1(s->pdiag, ip) = 1.0;
        }
    }
}

static int apply_preconditioner(const orc_state *s, const cg_masks *m, const Real *r, Real *z)
{
    memset(z, 0, sizeof(Real) * (size_t)s->npoin);
    if (s->cg_preconditioner == 0) {
        for (int ip = 1; ip <= s->npoin; ++ip) {
            if (!is_free(m, ip)) continue;
            if (!finite(A1(r, ip)))
                return fail("ConjugateGradient::solvePressure - invalid pressure residual at node %d", ip);
            A1(z, ip) = A1(r, ip);
        }
        return 0;
    }
    if (s->cg_preconditioner != 1)
        return fail("ConjugateGradient::solvePressure - cg_preconditioner must be 0 or 1");
    for (int ip = 1; ip <= s->npoin; ++ip) {
        if (!is_free(m, ip)) { A1(z, ip) = 0.0; continue; }
        const Real diag = A1(s->pdiag, ip);
        if (!finite(diag) || diag <= 1.0e-30)
            return fail("ConjugateGradient::solvePressure - invalid Jacobi pressure diagonal at active node %d", ip);
        if (!finite(A1(r, ip)))
            return fail("ConjugateGradient::solvePressure - invalid pressure residual at node %d", ip);
        A1(z, ip) = A1(r, ip) / diag;
    }
    return 0;
}

static int convergence_reached(const orc_state *s, Real rel_l2, Real max_abs, int *out)
{
    if (s->cg_conv_test == 1) { *out = max_abs <= s->absToler; return 0; }
    if (s->cg_conv_test == 2) { *out = rel_l2 <= s->relToler; return 0; }
    if (s->cg_conv_test == 3) { *out = max_abs <= s->absToler || rel_l2 <= s->relToler; return 0; }
    return fail("ConjugateGradient::solvePressure - cg_conv_test must be 1, 2, or 3");
}

int orc_solve_pressure(orc_state *s, orc_cg_result *res)
{
    memset(res, 0, sizeof *res);
    if (s->cg_max_iter < 1) return fail("ConjugateGradient::solvePressure - cg_max_iter must be positive");
    if (s->relToler <= 0.0 || s->absToler <= 0.0)
        return fail("ConjugateGradient::solvePressure - relToler and absToler must be positive");
    if (s->cg_preconditioner < 0 || s->cg_preconditioner > 1)
        return fail("ConjugateGradient::solvePressure - cg_preconditioner must be 0 or 1");
    if (s->cg_conv_test < 1 || s->cg_conv_test > 3)
        return fail("ConjugateGradient::solvePressure - cg_conv_test must be 1, 2, or 3");
    const int N = s->npoin;
    cg_masks m;
    m.active = calloc((size_t)N + 1, 1);
    m.fixed = calloc((size_t)N + 1, 1);
    m.fixed_values = calloc((size_t)N + 1, sizeof(Real));
    Real *r = calloc((size_t)N, sizeof(Real)), *z = calloc((size_t)N, sizeof(Real));
    Real *p = calloc((size_t)N, sizeof(Real)), *Ap = calloc((size_t)N, sizeof(Real));
    int rc = 0;
    /* build_pressure_active_mask (:164-205) */
    for (int ie = 1; ie <= s->nelem; ++ie) {
        if (A1(s->mat_elem, ie) != 0) continue;
        for (int in = 1; in <= 4; ++in) m.active[IN(s, ie, in)] = 1;
    }
    int active_count = 0;
    for (int ip = 1; ip <= N; ++ip) active_count += m.active[ip] != 0;
    if (active_count == 0) { rc = fail("ConjugateGradient::solvePressure - no fluid-connected pressure nodes found"); goto done; }
    /* build_fixed_pressure_mask (:207-241) */
    for (int i = 1; i <= s->bc_fixed; ++i) {
        const int ip = A1(s->bc_list, i);
        if (ip < 1 || ip > N) { rc = fail("ConjugateGradient::solvePressure - fixed pressure node out of range"); goto done; }
        if (!m.active[ip]) { rc = fail("ConjugateGradient::solvePressure - fixed pressure node is not fluid-connected: node %d", ip); goto done; }
        m.fixed[ip] = 1;
        m.fixed_values[ip] = A1(s->bc_values, i);
    }
    impose_constraints(s, &m);
    if ((rc = orc_pressure_multiply(s, s->pres, Ap))) goto done;
    ORC_OMP("omp parallel for schedule(static)")  /* ConjugateGradient.cpp:634 */
    for (int ip = 1; ip <= N; ++ip) A1(r, ip) = m.active[ip] ? A1(s->rhs1, ip) - A1(Ap, ip) : 0.0;
    zero_non_free(s, &m, r);
    zero_non_free(s, &m, Ap);
    res->initial_l2 = sqrt(dot_free(s, &m, r, r));
    if (res->initial_l2 <= 1.0e-300) {
        res->final_l2 = res->initial_l2;
        res->final_relative_l2 = 0.0;
        res->final_max_abs = max_abs_free(s, &m, r);
        res->converged = 1;
        goto done;
    }
    res->final_l2 = res->initial_l2;
    res->final_relative_l2 = 1.0;
    res->final_max_abs = max_abs_free(s, &m, r);
    {
        int conv = 0;
        if ((rc = convergence_reached(s, res->final_relative_l2, res->final_max_abs, &conv))) goto done;
        if (conv) { res->converged = 1; goto done; }
    }
    if ((rc = apply_preconditioner(s, &m, r, z))) goto done;
    for (int ip = 1; ip <= N; ++ip) A1(p, ip) = A1(z, ip);
    zero_non_free(s, &m, p);
    Real rz_old = dot_free(s, &m, r, z);
    if (rz_old <= 1.0e-300 || !finite(rz_old)) {
        rc = fail("ConjugateGradient::solvePressure - invalid initial preconditioned residual dot product. rz_old=%g", rz_old);
        goto done;
    }
    for (int iter = 1; iter <= s->cg_max_iter; ++iter) {
        if ((rc = orc_pressure_multiply(s, p, Ap))) goto done;
        zero_non_free(s, &m, Ap);
        const Real pAp = dot_free(s, &m, p, Ap);
        if (fabs(pAp) <= 1.0e-300 || !finite(pAp)) { rc = fail("ConjugateGradient::solvePressure - breakdown: invalid pAp"); goto done; }
        const Real alpha = rz_old / pAp;
        ORC_OMP("omp parallel for schedule(static)")  /* ConjugateGradient.cpp:762 */
        for (int ip = 1; ip <= N; ++ip) {
            if (is_free(&m, ip)) {
                A1(s->pres, ip) += alpha * A1(p, ip);
                A1(r, ip) -= alpha * A1(Ap, ip);
            }
        }
        impose_constraints(s, &m);
        zero_non_free(s, &m, r);
        res->iterations = iter;
        res->final_l2 = sqrt(dot_free(s, &m, r, r));
        res->final_max_abs = max_abs_free(s, &m, r);
        res->final_relative_l2 = res->final_l2 / res->initial_l2;
        int conv = 0;
        if ((rc = convergence_reached(s, res->final_relative_l2, res->final_max_abs, &conv))) goto done;
        if (conv) { res->converged = 1; break; }
        if ((rc = apply_preconditioner(s, &m, r, z))) goto done;
        const Real rz_new = dot_free(s, &m, r, z);
        if (rz_new <= 1.0e-300 || !finite(rz_new)) {
            rc = fail("ConjugateGradient::solvePressure - invalid updated preconditioned residual dot product. rz_new=%g", rz_new);
            goto done;
        }
        const Real beta = rz_new / rz_old;
        ORC_OMP("omp parallel for schedule(static)")  /* ConjugateGradient.cpp:834 */
        for (int ip = 1; ip <= N; ++ip) A1(p, ip) = is_free(&m, ip) ? A1(z, ip) + beta * A1(p, ip) : 0.0;
        rz_old = rz_new;
    }
done:
    free(m.active); free(m.fixed); free(m.fixed_values); free(r); free(z); free(p); free(Ap);
    return rc;
}

/* Steps::step2SemiImplicit (Steps.cpp:305-350) */
int orc_step2(orc_state *s)
{
    if (orc_assemble_step2_rhs(s)) return 1;
    orc_cg_result res;
    if (s->solver_opt == 1) {
        if (orc_solve_pressure(s, &res)) return 1;
    } else if (s->solver_opt == 2) {
        return fail("Steps::step2SemiImplicit - banded pressure solver path is not ported for CBS3D yet");
    } else if (s->solver_opt == 3) {
        return fail("Steps::step2SemiImplicit - solver_opt=3 requests PETSc, but this executable was built without CBS3D_USE_PETSC");
    } else {
        return fail("Steps::step2SemiImplicit - unknown pressure solver option");
    }
    s->last_cg_iterations = res.iterations;
    s->last_cg_initial_l2 = res.initial_l2;
    s->last_cg_final_l2 = res.final_l2;
    s->last_cg_relative_l2 = res.final_relative_l2;
    s->last_cg_max_abs = res.final_max_abs;
    if (!res.converged) return fail("Steps::step2SemiImplicit - pressure solver failed to converge");
    return orc_apply_pressure(s);
}

/* Steps::step3SemiImplicit (Steps.cpp:377-428) */
int orc_step3(orc_state *s)
{
    memset(s->rhs, 0, sizeof(Real) * 3 * (size_t)s->npoin);
    for (int ie = 1; ie <= s->nelem; ++ie) {
        if (!is_fluid(s, ie)) continue;
        const Real dj = A1(s->detJ, ie);
        if (dj <= 0.0 || !finite(dj)) return fail("Steps::step3SemiImplicit - invalid detJ at element %d", ie);
        Real grad_pres[4] = {0.0, 0.0, 0.0, 0.0};
        for (int idim = 1; idim <= 3; ++idim)
            for (int a = 1; a <= 4; ++a) grad_pres[idim] += DN(s, ie, idim, a) * A1(s->pres, IN(s, ie, a));
        const Real vol4 = dj * FCON1;
        for (int a = 1; a <= 4; ++a) {
            const int ip = IN(s, ie, a);
            for (int idim = 1; idim <= 3; ++idim) V3(s->rhs, idim, ip) -= vol4 * grad_pres[idim];
        }
    }
    for (int ip = 1; ip <= s->npoin; ++ip)
        for (int i = 1; i <= 3; ++i) V3(s->unkno, i, ip) += V3(s->rhs, i, ip) * A1(s->elcoe2, ip);
    return orc_apply_velocity_package(s);
}

/* ------------------------------------------------------------------------ */
/* Step 4: EnergyAssembly::assembleStep4Rhs (EnergyAssembly.cpp:118-611)      */
/* ------------------------------------------------------------------------ */
int orc_assemble_step4_rhs(orc_state *s)
{
    memset(s->rhs1, 0, sizeof(Real) * (size_t)s->npoin);
    for (int ie = 1; ie <= s->nelem; ++ie) {
        const Real dj = A1(s->detJ, ie);
        if (dj <= 0.0 || !finite(dj)) return fail("EnergyAssembly::assembleStep4Rhs - invalid detJ at element %d", ie);
        const Real rcp = A1(s->rho_cp_e, ie), kk = A1(s->k_e, ie), q = A1(s->Qvol_e, ie);
        if (rcp <= 0.0 || !finite(rcp)) return fail("EnergyAssembly::assembleStep4Rhs - invalid rho*Cp at element %d", ie);
        if (kk <= 0.0 || !finite(kk)) return fail("EnergyAssembly::assembleStep4Rhs - invalid thermal conductivity at element %d", ie);
        if (!finite(q)) return fail("EnergyAssembly::assembleStep4Rhs - invalid volumetric source at element %d", ie);
        Real dTdx = 0.0, dTdy = 0.0, dTdz = 0.0;
        for (int a = 1; a <= 4; ++a) {
            const Real T = A1(s->temperature1, IN(s, ie, a));
            dTdx += DN(s, ie, 1, a) * T;
            dTdy += DN(s, ie, 2, a) * T;
            dTdz += DN(s, ie, 3, a) * T;
        }
        Real lrhs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (is_fluid(s, ie)) {
            /* add_fluid_convection (:214-269) */
            const Real adv_factor = rcp * dj / 120.0;
            Real u_sum = 0.0, v_sum = 0.0, w_sum = 0.0;
            for (int b = 1; b <= 4; ++b) {
                const int ip = IN(s, ie, b);
                u_sum += V3(s->unkno, 1, ip);
                v_sum += V3(s->unkno, 2, ip);
                w_sum += V3(s->unkno, 3, ip);
            }
            for (int a = 1; a <= 4; ++a) {
                const int ip = IN(s, ie, a);
                const Real uw = u_sum + V3(s->unkno, 1, ip);
                const Real vw = v_sum + V3(s->unkno, 2, ip);
                const Real ww = w_sum + V3(s->unkno, 3, ip);
                const Real advT = uw * dTdx + vw * dTdy + ww * dTdz;
                lrhs[a] -= adv_factor * advT;
            }
            /* add_fluid_convection_stabilisation (:297-357) */
            const Real dt = A1(s->delte, ie);
            if (dt <= 0.0 || !finite(dt))
                return fail("EnergyAssembly::assembleStep4Rhs - element timestep must be positive for thermal stabilisation at element %d", ie);
            Real ubar = 0.0, vbar = 0.0, wbar = 0.0;
            for (int a = 1; a <= 4; ++a) {
                const int ip = IN(s, ie, a);
                ubar += V3(s->unkno, 1, ip);
                vbar += V3(s->unkno, 2, ip);
                wbar += V3(s->unkno, 3, ip);
            }
            ubar /= 4.0; vbar /= 4.0; wbar /= 4.0;
            const Real advT = ubar * dTdx + vbar * dTdy + wbar * dTdz;
            const Real volume = dj / 6.0;
            const Real tau_factor = 0.5 * dt * rcp * volume;
            for (int a = 1; a <= 4; ++a) {
                const Real u_gradNa = ubar * DN(s, ie, 1, a) + vbar * DN(s, ie, 2, a) + wbar * DN(s, ie, 3, a);
                lrhs[a] += tau_factor * u_gradNa * advT;
            }
        }
        /* add_thermal_diffusion (:386-412) */
        {
            const Real volume = dj / 6.0;
            for (int a = 1; a <= 4; ++a) {
                const Real g = DN(s, ie, 1, a) * dTdx + DN(s, ie, 2, a) * dTdy + DN(s, ie, 3, a) * dTdz;
                lrhs[a] -= kk * volume * g;
            }
        }
        /* add_volumetric_source (:437-458) */
        if (q != 0.0) {
            const Real source = q * dj * MASS_FACTOR;
            for (int a = 1; a <= 4; ++a) lrhs[a] += source;
        }
        for (int a = 1; a <= 4; ++a) A1(s->rhs1, IN(s, ie, a)) += lrhs[a];
    }
    /* add_prescribed_heat_flux (:484-532) */
    if (s->heat_flux_bc != 0.0) {
        for (int ib = 1; ib <= s->nboun; ++ib) {
            if (ISD(s, 6, ib) != BC_532) continue;
            const Real area = FNRM(s, 4, ib);
            if (area <= 0.0 || !finite(area))
                return fail("EnergyAssembly::assembleStep4Rhs - invalid heat-flux boundary face area");
            const Real contribution = s->heat_flux_bc * area / 3.0;
            for (int in = 1; in <= 3; ++in) {
                const int ip = ISD(s, in, ib);
                if (ip < 1 || ip > s->npoin)
                    return fail("EnergyAssembly::assembleStep4Rhs - heat-flux boundary node out of range");
                A1(s->rhs1, ip) += contribution;
            }
        }
    }
    return 0;
}

/* Steps::step4Energy (Steps.cpp:448-467) */
int orc_step4(orc_state *s)
{
    if (s->temp_calc < 1) return 0;
    if (orc_assemble_step4_rhs(s)) return 1;
    ORC_OMP("omp parallel for schedule(static)")  /* Steps.cpp:459 */
    for (int ip = 1; ip <= s->npoin; ++ip)
        A1(s->temperature, ip) = A1(s->temperature1, ip) + A1(s->rhs1, ip) * A1(s->elcoe2p, ip);
    return orc_apply_temperature(s);
}

/* Solver::updateVelocityMagnitude (Solver.cpp:689-699) */
void orc_update_velocity_magnitude(orc_state *s)
{
    for (int ip = 1; ip <= s->npoin; ++ip) {
        const Real u = V3(s->unkno, 1, ip), v = V3(s->unkno, 2, ip), w = V3(s->unkno, 3, ip);
        A1(s->velocity, ip) = sqrt(u * u + v * v + w * w + 1.0e-16);
    }
}

/* ------------------------------------------------------------------------ */
/* Convergence::evaluate (Convergence.cpp:78-349)                            */
/* ------------------------------------------------------------------------ */
static Real safe_inverse_weight(Real inv_mass)
{
    if (fabs(inv_mass) <= 1.0e-300 || !finite(inv_mass)) return 1.0;
    return 1.0 / inv_mass;
}

int orc_convergence(orc_state *s)
{
    const Real eps_norm = 1.0e-30;
    const int N = s->npoin;
    char *tf = malloc((size_t)N + 1), *ts = malloc((size_t)N + 1);
    touch_masks(s, tf, ts);
    for (int ip = 1; ip <= N; ++ip)
        if (!tf[ip] && !ts[ip]) { free(tf); free(ts); return fail("Convergence - orphan node not touched by any material element"); }
    Real *hb = s->hb;
    for (int i = 0; i < 15; ++i) hb[i] = 0.0;
    for (int ip = 1; ip <= N; ++ip) {
        const int vel_active = tf[ip] && !ts[ip];
        for (int idim = 1; idim <= 3; ++idim) {
            if (!vel_active) continue;
            const int base = 3 * (idim - 1);
            const Real old_value = V3(s->unkn1, idim, ip), new_value = V3(s->unkno, idim, ip);
            const Real delta = new_value - old_value;
            hb[base + 0] += delta * delta;
            hb[base + 1] += new_value * new_value;
            const Real rhs_scale = delta * safe_inverse_weight(A1(s->elcoe2, ip));
            hb[base + 2] += rhs_scale * rhs_scale;
        }
        if (tf[ip]) {
            const Real delta = A1(s->pres, ip) - A1(s->pres1, ip);
            hb[9] += delta * delta;
            hb[10] += A1(s->pres, ip) * A1(s->pres, ip);
            hb[11] += delta * delta;
        }
        if (s->temp_calc > 0 && (tf[ip] || ts[ip])) {
            const Real delta = A1(s->temperature, ip) - A1(s->temperature1, ip);
            hb[12] += delta * delta;
            hb[13] += A1(s->temperature, ip) * A1(s->temperature, ip);
            const Real rhs_scale = delta * safe_inverse_weight(A1(s->elcoe2p, ip));
            hb[14] += rhs_scale * rhs_scale;
        }
    }
    for (int base = 0; base <= 6; base += 3) {
        hb[base + 0] = sqrt(hb[base + 0] / (hb[base + 1] + eps_norm));
        hb[base + 1] = sqrt(hb[base + 1]);
        hb[base + 2] = sqrt(hb[base + 2]);
    }
    hb[9] = sqrt(hb[9] / (hb[10] + eps_norm));
    hb[10] = sqrt(hb[10]);
    hb[11] = sqrt(hb[11]);
    if (s->temp_calc > 0) {
        hb[12] = sqrt(hb[12] / (hb[13] + eps_norm));
        hb[13] = sqrt(hb[13]);
        hb[14] = sqrt(hb[14]);
    } else {
        hb[12] = hb[13] = hb[14] = 0.0;
    }
    free(tf); free(ts);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Solver::advanceOneStep (src/solver/Solver.cpp:609-680)                    */
/* ------------------------------------------------------------------------ */
int orc_advance_one_step(orc_state *s, int32_t iitime)
{
    const size_t N = (size_t)s->npoin;
    memcpy(s->unkn1, s->unkno, sizeof(Real) * 3 * N);
    memcpy(s->pres1, s->pres, sizeof(Real) * N);
    memcpy(s->temperature1, s->temperature, sizeof(Real) * N);
    if (orc_compute_timestep(s, iitime)) return 1;
    if (s->transient_on > 0) s->rtime += s->dtreal;
    if (orc_update_lhs_diagonal(s)) return 1;
    if (s->cbs_scheme != 1) return fail("Steps::step1 - cbs_scheme must be 1 or 0");
    if (orc_step1(s)) return 1;
    if (s->step2_check > 0) {
        if (step2_timestep_correction(s)) return 1;
        if (orc_update_lhs_diagonal(s)) return 1;
    }
    if (orc_step2(s)) return 1;
    if (orc_step3(s)) return 1;
    if (orc_step4(s)) return 1;
    orc_update_velocity_magnitude(s);
    return orc_convergence(s);
}
