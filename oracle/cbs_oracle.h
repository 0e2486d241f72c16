/*
 * CPU oracle for the CBS time-step hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference solver's per-step algorithm
 * (CBS3D++_SI, src/solver, src/assembly, src/linalg, src/timestep,
 * src/boundary, include/cbs/parallel/Coloring.hpp), written to reproduce the
 * reference's floating-point operation order exactly at OMP_NUM_THREADS=1:
 * element loops in the same order, dot products summed in ascending node
 * order, no FMA contraction (build with -ffp-contract=off).  Each function
 * cites the reference lines it follows.
 *
 * Only the tests, the smoke check and the CPU-baseline leg of bench.py may
 * use this code, and only as the checker.  The product path never calls it.
 *
 * Array conventions: every array is 0-based in memory, laid out exactly like
 * the reference's 1-based column-major Array1D/2D/3D with element 0 dropped
 * (include/cbs/core/Array.hpp:40-212).  Node and element ids stored inside
 * arrays (intma, iside, bc_list, wall_node_list, color_elem) keep the
 * reference's 1-based values.
 */
#ifndef CBS_ORACLE_H
#define CBS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state {
    /* sizes */
    int32_t npoin, nelem, nboun, ncolor, bc_fixed, npoin_wall;
    /* run controls (include/cbs/core/RunConfig.hpp) */
    int32_t temp_calc, cbs_scheme, ilots, htype, dtfixed, dtfix_end, transient_on;
    int32_t solver_opt, rem_deltp, beta_opt, step2_check;
    int32_t dimensional_mode, material_properties_enabled;
    int32_t cg_preconditioner, cg_conv_test, cg_max_iter;
    double csafm, dtfix, deltr, ani, relToler, absToler, dtreal, rtime;
    double csafm2, epsilon1;
    double inlet_u, inlet_v, inlet_w, inlet_temperature, inlet_u_from_massflow, heat_flux_bc;
    /* mesh and static geometry */
    const int32_t *intma;      /* (4, nelem) */
    const int32_t *iside;      /* (6, nboun) */
    const int32_t *fedge;      /* (4, nelem) */
    const int32_t *mat_elem;   /* (nelem) */
    const int32_t *wall_node_list; /* (npoin_wall) */
    const int32_t *bc_list;    /* (bc_fixed) */
    const int32_t *color_ptr;  /* (ncolor+1), 0-based offsets */
    const int32_t *color_elem; /* (nelem), 1-based element ids */
    const double *coord;       /* (3, npoin) */
    const double *dNkdx;       /* (12, nelem): (e-1)*12 + (dim-1)*4 + (a-1) */
    const double *detJ;        /* (nelem) */
    const double *annxf;       /* (4, 4, nelem) */
    const double *face_norm;   /* (4, nboun) */
    const double *elcoe_e;     /* (4, nelem) */
    const double *alen_e;      /* (nelem) */
    const double *bc_values;   /* (bc_fixed) */
    const double *rho_e, *mu_e, *k_e, *rho_cp_e, *alpha_e, *Qvol_e; /* (nelem) */
    double *pdiagE;            /* (4, nelem) */
    double *gstifE;            /* (6, nelem) */
    /* solution and work state */
    double *unkno, *unkn1;     /* (3, npoin) */
    double *pres, *pres1, *temperature, *temperature1;
    double *rhs;               /* (3, npoin) */
    double *rhs1;
    double *delte;             /* (nelem) */
    double *deltp, *deltp1, *deltp2, *beta1, *vvis, *velcp, *alen;
    double *unkno_unit_vec;    /* (3, npoin) */
    double *elcoe2, *elcoe2p, *pdiag;
    double *gstif;             /* (6, nelem) */
    double *velocity;
    double *hb;                /* (15) */
    /* last pressure solve (Steps.cpp:337-341) */
    int32_t last_cg_iterations;
    double last_cg_initial_l2, last_cg_final_l2, last_cg_relative_l2, last_cg_max_abs;
} orc_state;

typedef struct orc_cg_result {
    int32_t iterations, converged;
    double initial_l2, final_l2, final_relative_l2, final_max_abs;
} orc_cg_result;

/* Error text of the last failing call (mirrors the reference's exception text). */
const char *orc_last_error(void);

/* Coloring::build (include/cbs/parallel/Coloring.hpp:27-97).  color_ptr must
 * hold nelem+1 entries; returns ncolor, or -1 on error. */
int32_t orc_coloring(int32_t nelem, int32_t npoin, const int32_t *intma,
                     int32_t *color_ptr, int32_t *color_elem);

/* Every function below returns 0 on success, nonzero after a failure that the
 * reference reports by throwing. */
int orc_build_element_pressure_terms(orc_state *s);
int orc_compute_timestep(orc_state *s, int32_t iitime);
int orc_update_lhs_diagonal(orc_state *s);
int orc_assemble_step1_rhs(orc_state *s);
int orc_step1(orc_state *s);
int orc_assemble_step2_rhs(orc_state *s);
int orc_pressure_multiply(const orc_state *s, const double *x, double *y);
int orc_solve_pressure(orc_state *s, orc_cg_result *result);
int orc_step2(orc_state *s);
int orc_step3(orc_state *s);
int orc_assemble_step4_rhs(orc_state *s);
int orc_step4(orc_state *s);
int orc_apply_velocity_package(orc_state *s);
int orc_apply_velocity(orc_state *s);
int orc_apply_temperature(orc_state *s);
int orc_apply_pressure(orc_state *s);
void orc_update_velocity_magnitude(orc_state *s);
int orc_convergence(orc_state *s);
/* Solver::advanceOneStep (src/solver/Solver.cpp:609-680). */
int orc_advance_one_step(orc_state *s, int32_t iitime);

#ifdef __cplusplus
}
#endif
#endif
