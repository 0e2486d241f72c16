/*
 * cbs_hip.h - C ABI of the MI355X (gfx950) implementation of the CBS
 * time-step hot path of CBS3D++_SI.
 *
 * The library owns device copies of the mesh, geometry and solution state of
 * one case (a "context") and runs the reference's per-step algorithm on the
 * GPU: time step, LHS diagonals, Steps 1-4, the matrix-free Jacobi-PCG
 * pressure solve, boundary conditions and convergence norms.  Every entry
 * point below names the reference function it replaces (file:line in the
 * reference tree).
 *
 * Conventions
 *  - Plain pointers and sizes only.  Host arrays are borrowed for the duration
 *    of a call and never retained.
 *  - Arrays use the reference's memory layout (include/cbs/core/Array.hpp:
 *    1-based, column-major) with the unused element 0 dropped, i.e. a caller
 *    holding a CBSStateSI passes `arr.raw().data() + 1`.  Ids stored inside
 *    arrays (intma, iside, bc_list, wall_node_list, color_elem) keep the
 *    reference's 1-based values.
 *  - One host thread per context; calls on one context are not reentrant.
 *    Each call returns with its GPU work complete.
 *  - Every call returns a cbs_status; never aborts.  The message of the last
 *    failure is available from cbs_hip_last_error().
 */
#ifndef CBS_HIP_H
#define CBS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBS_HIP_ABI_VERSION 2

/* CBS_HIP_ABI_VERSION of the loaded library, to be compared with the
 * header's before any other call. */
int32_t cbs_hip_abi_version(void);

/* Status codes.  Each maps to one class of std::runtime_error the reference
 * throws (messages are reproduced in cbs_hip_last_error). */
typedef enum cbs_status {
    CBS_OK = 0,
    CBS_ERR_BAD_ARGUMENT = 1,        /* invalid pointer, size or option */
    CBS_ERR_HIP = 2,                 /* HIP runtime failure (allocation, launch) */
    CBS_ERR_CG_NOT_CONVERGED = 3,    /* Steps.cpp:343-347 */
    CBS_ERR_CG_BREAKDOWN = 4,        /* |pAp| <= 1e-300 or non-finite: ConjugateGradient.cpp:749-753 */
    CBS_ERR_JACOBI_DIAGONAL = 5,     /* invalid Jacobi diagonal: ConjugateGradient.cpp:349-357 */
    CBS_ERR_NONFINITE_RESIDUAL = 6,  /* non-finite r or invalid (r,z): ConjugateGradient.cpp:358-365, 718-723 */
    CBS_ERR_INVALID_TIMESTEP = 7,    /* TimeStep.cpp:292-297, 800-805, 845-866, 912-917 */
    CBS_ERR_INVALID_GEOMETRY = 8,    /* detJ / face area / orphan node */
    CBS_ERR_UNSUPPORTED = 9,         /* option the reference rejects (explicit CBS, solver_opt 2/3) */
    CBS_ERR_INVALID_MATERIAL = 10,   /* rho, mu, k, rho*cp: EnergyAssembly.cpp:80-108 */
    CBS_ERR_IO = 11                  /* output file cannot be opened or written: Post.cpp:539-542, 787-790, 896-899, 1147-1150 */
} cbs_status;

/* Run controls of the hot path: the fields of RunConfig
 * (include/cbs/core/RunConfig.hpp:38-323) after MeshIO::readAll and
 * Solver::initialise. */
typedef struct cbs_config {
    int32_t temp_calc, cbs_scheme, ilots, htype, dtfixed, dtfix_end, transient_on;
    int32_t solver_opt, rem_deltp, beta_opt, step2_check;
    int32_t dimensional_mode, material_properties_enabled, mass_flow_inlet_enabled;
    int32_t convection_type, ntime;
    int32_t cg_preconditioner, cg_conv_test, cg_max_iter;
    int32_t pnode;
    double csafm, dtfix, deltr, csafm2, epsilon1;
    double re, pr, ri, ani;
    double relToler, absToler;
    double inlet_u, inlet_v, inlet_w, inlet_temperature, inlet_density, inlet_mass_flow_rate;
    double inlet_u_from_massflow, outlet_pressure_gauge, heat_flux_bc;
    double cinf[5];
} cbs_config;

/* Mesh, geometry and boundary data of CBSStateSI after Solver::initialise
 * (src/solver/Solver.cpp:242-302).  Sizes: E = nelem, N = npoin, B = nboun. */
typedef struct cbs_mesh_desc {
    int32_t npoin, nelem, nboun;
    int32_t ncolor;               /* Coloring::build (Coloring.hpp:27-97) */
    int32_t npoin_wall;           /* Preprocess::wallDetermination */
    int32_t bc_fixed;             /* Preprocess::detectPressureBoundaryNodes */
    const int32_t *intma;         /* (4, E) node ids */
    const double *coord;          /* (3, N) */
    const int32_t *mat_elem;      /* (E) 0 fluid, >0 solid */
    const double *rho_e, *mu_e, *k_e, *rho_cp_e, *alpha_e, *Qvol_e; /* (E) */
    const double *dNkdx;          /* (12, E): (e-1)*12 + (dim-1)*4 + (a-1) */
    const double *detJ;           /* (E) = 6 * volume */
    const double *annxf;          /* (4, 4, E): A_f n_f and A_f per local face */
    const int32_t *fedge;         /* (4, E): 0 interior, 1 velocity BC, 2 other exterior */
    const double *elcoe_e;        /* (4, E) lumped element-node mass */
    const double *alen_e;         /* (E) element length */
    const int32_t *iside;         /* (6, B): n1 n2 n3 local_face parent bc */
    const double *face_norm;      /* (4, B) */
    const int32_t *wall_node_list;/* (npoin_wall) */
    const int32_t *bc_list;       /* (bc_fixed) */
    const double *bc_values;      /* (bc_fixed) */
    const int32_t *color_ptr;     /* (ncolor+1) 0-based offsets into color_elem */
    const int32_t *color_elem;    /* (E) element ids, colour-major */
} cbs_mesh_desc;

/* Pressure-solver precision forms. */
typedef enum cbs_precision {
    CBS_PCG_FP64 = 0,   /* the reference algorithm in FP64 */
    CBS_PCG_MIXED = 1,  /* FP32 apply, FP32 direction, FP64 iterate/residual/reductions, FP64 true-residual refinement */
    CBS_PCG_AMG = 2     /* FP64 CG preconditioned by one FP32 smoothed-aggregation multigrid V-cycle (build-only) */
} cbs_precision;

typedef struct cbs_cg_options {
    int32_t precision;          /* cbs_precision */
    int32_t max_refinements;    /* multi-RHS mixed: true-residual refreshes per column (0 = default 50) */
    double inner_rel_tol;       /* multi-RHS mixed: recursive-residual drop between refreshes (0 = default 0.1) */
} cbs_cg_options;

/* ConjugateGradient::Result (include/cbs/linalg/ConjugateGradient.hpp:36-68),
 * field for field, plus the mixed form's refinements and the algorithmic
 * bytes moved.
 * total_seconds: wall time of the solve.  setup_seconds: wall time until
 * the first iteration (solver-order gathers, operator values, constraints,
 * initial residual; the first solve of a context also runs the apply
 * autotuner and, for CBS_PCG_AMG, builds the hierarchy).
 * matvec, dot, preconditioner, vector_update and constraint_seconds: filled
 * when cbs_hip_set_profiling is on (0 otherwise), from HIP events on the
 * context's stream.  One-off work (constraints, the mixed form's true
 * residuals) is its measured device time; the iterations' wall time (total
 * less setup and the one-off work) is split among the phases in proportion
 * to the device time of their kernels in sampled iterations (the iteration
 * in the middle of each chunk of graph replays, or every 8th without
 * graphs), so the phases add up to the solve as the reference's timers do.
 * The kernels fuse phases the reference keeps apart: the apply also
 * forms the (p, Ap) partials (matvec), the update kernel applies the Jacobi
 * step and forms the dot partials (vector_update), the finishing kernels
 * reduce the partials (dot); so preconditioner_seconds is 0 for the Jacobi
 * forms and the V-cycles' time for CBS_PCG_AMG.  Multi-RHS results and a
 * batch's per-member results (cbs_hip_get_member_cg_result) leave setup and
 * the phase timers at 0. */
typedef struct cbs_cg_result {
    int32_t iterations;
    int32_t converged;
    double initial_l2, final_l2, final_relative_l2, final_max_abs;
    double total_seconds, setup_seconds, matvec_seconds, dot_seconds, preconditioner_seconds,
        vector_update_seconds, constraint_seconds;
    int32_t matvec_calls, dot_calls, preconditioner_calls;
    int32_t refinements;        /* mixed form: outer FP64 cycles */
    double algorithmic_bytes;   /* sum over iterations of the per-iteration byte model */
} cbs_cg_result;

/* State scalars of one context. */
typedef struct cbs_scalars {
    double dtreal, rtime;
    double hb[15];              /* Convergence::evaluate (Convergence.cpp:276-349) */
    int32_t last_cg_iterations;
    double last_cg_initial_l2, last_cg_final_l2, last_cg_relative_l2, last_cg_max_abs;
} cbs_scalars;

/* Per-node and per-element arrays of the context, by name. */
typedef enum cbs_field {
    CBS_F_UNKNO = 0,        /* (3, N) */
    CBS_F_UNKN1,            /* (3, N) */
    CBS_F_PRES,             /* (N) */
    CBS_F_PRES1,            /* (N) */
    CBS_F_TEMPERATURE,      /* (N) */
    CBS_F_TEMPERATURE1,     /* (N) */
    CBS_F_RHS,              /* (3, N) */
    CBS_F_RHS1,             /* (N) */
    CBS_F_DELTE,            /* (E) */
    CBS_F_DELTP,            /* (N) */
    CBS_F_DELTP1,           /* (N) */
    CBS_F_DELTP2,           /* (N) */
    CBS_F_BETA1,            /* (N) */
    CBS_F_VVIS,             /* (N) */
    CBS_F_VELCP,            /* (N) */
    CBS_F_ALEN,             /* (N) */
    CBS_F_UNKNO_UNIT_VEC,   /* (3, N) */
    CBS_F_ELCOE2,           /* (N) */
    CBS_F_ELCOE2P,          /* (N) */
    CBS_F_PDIAG,            /* (N) */
    CBS_F_GSTIF,            /* (6, E) */
    CBS_F_GSTIFE,           /* (6, E) */
    CBS_F_PDIAGE,           /* (4, E) */
    CBS_F_VELOCITY,         /* (N) */
    CBS_F_COUNT
} cbs_field;

typedef struct cbs_hip_ctx cbs_hip_ctx;

/* Message of the last failure on ctx (or of the last failed create when ctx
 * is NULL).  Valid until the next call on that context. */
const char *cbs_hip_last_error(const cbs_hip_ctx *ctx);

/* Number of values of a field for this context (components included). */
int64_t cbs_hip_field_size(const cbs_hip_ctx *ctx, int32_t field);

/* Tiles per case that contexts created after the call aim at (default
 * 512).  The apply's element tiles are as large as the LDS allows (about
 * 24,000 tets), but a case is cut into at least this many tiles where that
 * leaves each 512 tets or more, so a small mesh still spreads over the
 * chip's CUs.  1 gives LDS-sized tiles only.  Results differ between tilings
 * only in the rounding of the assembled operator's sums.  Returns the
 * previous target (-1 for tiles < 1).  No reference counterpart. */
int32_t cbs_hip_set_tile_target(int32_t tiles);

/* Meshes with at least this many elements, in contexts created after the
 * call, store the element passes' slot terms (LHS, Step 3) staged
 * per block in LDS and written as whole 128-byte lines; smaller ones store
 * each term directly (default 4,194,304: below it the passes are
 * latency-bound and the staging costs more than it saves).  The results
 * are the same bits either way.  Returns the previous value (-1 for
 * elems < 0).  No reference counterpart. */
int32_t cbs_hip_set_line_group_min_elems(int32_t elems);

/* Creates a context on the current HIP device and uploads the mesh.  Builds
 * the device layouts: the fluid colour-major element stream of the pressure
 * apply, node incidence lists in element order, boundary node programmes and
 * pressure masks.  Replaces the device-independent state of
 * Solver::initialise (src/solver/Solver.cpp:242-302).  Fields start at zero. */
int cbs_hip_create(const cbs_config *cfg, const cbs_mesh_desc *mesh, cbs_hip_ctx **out);
int cbs_hip_destroy(cbs_hip_ctx *ctx);

/* Batched ensemble: one context over nmember independent cases ("members",
 * e.g. lid-driven cavities of different sizes and Reynolds numbers), each
 * advanced exactly as its own context would advance it.  The device arrays
 * hold the disjoint union of the members (member k's nodes and elements
 * follow those of members 0..k-1; cbs_hip_member_range gives the offsets)
 * and every launch covers all members; the time-step minimum, dtreal, the
 * convergence norms and every PCG scalar are per member, and each member's
 * reductions are decomposed exactly as in a context of its own, so its
 * results are bit-identical to an unbatched run.  Members may differ in
 * mesh, ani and re (and pnode, ntime); every other run control must equal
 * member 0's.  On a batch context, fields are the union arrays; the
 * step/solve entry points act on all members (a member that converges
 * early idles until the others finish); cbs_hip_get_scalars reports member
 * 0 and a cbs_cg_result reports the member with the most iterations, with
 * converged set only when every member converged; pressure solves are FP64
 * only and cbs_hip_solve_pressure_multi is unavailable.  The reference runs
 * one case per process (Solver::run, src/solver/Solver.cpp:137-225): the
 * batch is that loop over cases, fused. */
int cbs_hip_create_batch(int32_t nmember, const cbs_config *cfgs, const cbs_mesh_desc *meshes, cbs_hip_ctx **out);
int cbs_hip_member_count(const cbs_hip_ctx *ctx, int32_t *nmember);
int cbs_hip_member_range(const cbs_hip_ctx *ctx, int32_t member, int32_t *node_offset, int32_t *npoin,
                         int32_t *elem_offset, int32_t *nelem);
/* Per-member state scalars and last pressure-solve result
 * (Steps.cpp:337-341 copies the latter into s.last_cg_*). */
int cbs_hip_get_member_scalars(cbs_hip_ctx *ctx, int32_t member, cbs_scalars *out);
int cbs_hip_get_member_cg_result(cbs_hip_ctx *ctx, int32_t member, cbs_cg_result *out);

/* Copy one field between host memory and the context (count values). */
int cbs_hip_set_field(cbs_hip_ctx *ctx, int32_t field, const double *host, int64_t count);
int cbs_hip_get_field(cbs_hip_ctx *ctx, int32_t field, double *host, int64_t count);

/* Wall-clock seconds of the phases of the context's construction, in order:
 * validation, connectivity and flags, colour-major slots, tiles,
 * incidences, boundary programmes, uploads, assembled operator, node slots,
 * node vectors (with the final synchronisation).  Writes up to cap values
 * and returns the number of phases (-1 on a null context). */
int cbs_hip_create_seconds(const cbs_hip_ctx *ctx, double *seconds, int32_t cap);

/* Coloring::build (include/cbs/parallel/Coloring.hpp:27-97) on the device,
 * with the reference's result bit for bit: greedy first fit in element
 * order, then colour-major order with elements ascending inside a colour.
 * Needs no context.  intma is (4, nelem) with 1-based ids; color_ptr holds
 * CBS_HIP_MAX_COLOURS + 1 entries (ncolor + 1 are written, 0-based offsets);
 * color_elem holds nelem element ids (1-based); levels (may be NULL)
 * receives the number of dependency levels the device walked.  Meshes that
 * need more than CBS_HIP_MAX_COLOURS colours return CBS_ERR_UNSUPPORTED (the
 * host colouring has no limit).  The message of a failure is
 * cbs_hip_last_error(NULL). */
#define CBS_HIP_MAX_COLOURS 256
int cbs_hip_coloring(int32_t nelem, int32_t npoin, const int32_t *intma, int32_t *ncolor, int32_t *color_ptr,
                     int32_t *color_elem, int32_t *levels);

/* cbs_hip_coloring without any allocation.  No counterpart in the reference:
 * this is for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  The same colouring, bit for bit, on
 * device arrays: intma_dev (4, nelem) 1-based ids, color_ptr_dev
 * (CBS_HIP_MAX_COLOURS + 1 entries, all written: 0-based offsets, constant
 * past ncolor) and color_elem_dev (nelem 1-based ids, colour-major); ncolor
 * and levels (may be NULL) are host outputs.  Its workspace is one fixed
 * device array per GPU of cbs_hip_coloring_noalloc_scratch_bytes() (42 MB),
 * held until the process ends: meshes of at most 1,048,576 elements and
 * 524,288 nodes (CBS_ERR_BAD_ARGUMENT beyond; cbs_hip_coloring has no such
 * limit).  It runs on a stream of its own and returns with the work done.
 * Calls on one GPU share its workspace: calls that overlap on one GPU
 * race and must be serialised by the caller.  The message of a failure is
 * cbs_hip_last_error(NULL). */
int cbs_hip_coloring_noalloc(int32_t nelem, int32_t npoin, const int32_t *intma_dev, int32_t *ncolor,
                             int32_t *color_ptr_dev, int32_t *color_elem_dev, int32_t *levels);
int64_t cbs_hip_coloring_noalloc_scratch_bytes(void);

/* Output snapshots: the device side of Post::writeSolution
 * (src/io/Post.cpp:842-860) for the writer of cbs_post.h.
 * cbs_hip_snapshot_begin copies pres, temperature, unkno and the velocity
 * magnitude within device memory on the context's stream (the next step may
 * then overwrite them), then to pinned host memory on a second stream, and
 * returns without waiting for the download: unlike the other calls it leaves
 * GPU work running.  A context has two slots; begin blocks until one is
 * released.  Single-case contexts only (CBS_ERR_UNSUPPORTED for a batch).
 * cbs_hip_snapshot_wait(snap) blocks until the download is complete;
 * cbs_hip_snapshot_field(snap, f) then gives the host copy of f (CBS_F_PRES,
 * CBS_F_TEMPERATURE, CBS_F_UNKNO, CBS_F_VELOCITY; NULL otherwise);
 * cbs_hip_snapshot_release(snap) returns the slot.  wait and release take the
 * snapshot as void* so that they can be passed to
 * cbs_post_write_solution_async as its ready and done callbacks, and may be
 * called from the writer's thread.  Release every snapshot (flush the writer)
 * before destroying the context. */
typedef struct cbs_snapshot cbs_snapshot;
int cbs_hip_snapshot_begin(cbs_hip_ctx *ctx, cbs_snapshot **out);
int cbs_hip_snapshot_wait(void *snapshot);
const double *cbs_hip_snapshot_field(const cbs_snapshot *snapshot, int32_t field);
int cbs_hip_snapshot_release(void *snapshot);
int cbs_hip_get_scalars(cbs_hip_ctx *ctx, cbs_scalars *out);
int cbs_hip_set_dtreal(cbs_hip_ctx *ctx, double dtreal);

/* PressureAssembly::buildElementPressureTerms (PressureAssembly.cpp:136-178). */
int cbs_hip_build_element_pressure_terms(cbs_hip_ctx *ctx);

/* TimeStep::computeTimeStep (TimeStep.cpp:473-583). */
int cbs_hip_compute_timestep(cbs_hip_ctx *ctx, int32_t iitime);

/* TimeStep::updateLhsDiagonal (TimeStep.cpp:778-954). */
int cbs_hip_update_lhs_diagonal(cbs_hip_ctx *ctx);

/* MomentumAssembly::assembleStep1Rhs (MomentumAssembly.cpp:462-554). */
int cbs_hip_assemble_step1_rhs(cbs_hip_ctx *ctx);

/* Steps::step1SemiImplicit (Steps.cpp:273-284). */
int cbs_hip_step1(cbs_hip_ctx *ctx);

/* PressureAssembly::assembleStep2Rhs (PressureAssembly.cpp:236-335). */
int cbs_hip_assemble_step2_rhs(cbs_hip_ctx *ctx);

/* MatrixVectorCalc::pressureMultiply (MatrixVectorCalc.cpp:125-205) on host
 * vectors of length npoin, with the context's current gstif and pdiag. */
int cbs_hip_pressure_multiply(cbs_hip_ctx *ctx, const double *x, double *y);

/* The same on device vectors (npoin doubles each), stream-ordered on the
 * context's stream; returns after the apply completes. */
int cbs_hip_pressure_multiply_device(cbs_hip_ctx *ctx, const double *x_dev, double *y_dev);

/* ConjugateGradient::solvePressure (ConjugateGradient.cpp:584-869) on the
 * context's rhs1 / pres / pdiag / gstif.  opts may be NULL (FP64). */
int cbs_hip_solve_pressure(cbs_hip_ctx *ctx, const cbs_cg_options *opts, cbs_cg_result *result);

/* Multi-right-hand-side form of ConjugateGradient::solvePressure: nrhs
 * (1..64) independent systems A x_j = b_j with the context's current
 * operator and constraints, each column running the reference's Jacobi-PCG
 * recurrence and convergence test.  B and X are (npoin, nrhs) column blocks
 * (column j at offset j*npoin); X holds the initial guesses on entry and the
 * solutions on return.  Up to 32 columns are solved together (more run in
 * passes of 32); the assembled operator's stream is shared by the columns
 * of a pass.  opts->precision selects FP64 or the mixed FP32/FP64
 * refinement form.  results receives nrhs entries. */
int cbs_hip_solve_pressure_multi(cbs_hip_ctx *ctx, int32_t nrhs, const double *B, double *X,
                                 const cbs_cg_options *opts, cbs_cg_result *results);

/* Steps::step2SemiImplicit (Steps.cpp:305-350): RHS, solve, Result copy,
 * non-convergence error, Boundary::applyPressure. */
int cbs_hip_step2(cbs_hip_ctx *ctx, const cbs_cg_options *opts, cbs_cg_result *result);

/* Steps::step3SemiImplicit (Steps.cpp:377-428). */
int cbs_hip_step3(cbs_hip_ctx *ctx);

/* EnergyAssembly::assembleStep4Rhs (EnergyAssembly.cpp:566-611). */
int cbs_hip_assemble_step4_rhs(cbs_hip_ctx *ctx);

/* Steps::step4Energy (Steps.cpp:448-467). */
int cbs_hip_step4(cbs_hip_ctx *ctx);

/* Boundary application.  which: 0 the Steps velocity package
 * (applySymmetry, applyVelocity, applyOutletBackflowControl, solid zeroing:
 * Steps.cpp:170-176), 1 Boundary::applyVelocity (Boundary.cpp:341-508),
 * 2 Boundary::applyTemperature (:529-626), 3 Boundary::applyPressure (:642-652). */
int cbs_hip_apply_boundary(cbs_hip_ctx *ctx, int32_t which);

/* Solver::updateVelocityMagnitude (Solver.cpp:689-699). */
int cbs_hip_update_velocity_magnitude(cbs_hip_ctx *ctx);

/* Convergence::evaluate (Convergence.cpp:276-349). */
int cbs_hip_convergence(cbs_hip_ctx *ctx);

/* Sizes of a context (fluid elements = elements streamed by the apply). */
typedef struct cbs_info {
    int32_t npoin, nelem, nfluid, ncolor;
    int64_t device_bytes;       /* device memory held by the context */
} cbs_info;
int cbs_hip_info(const cbs_hip_ctx *ctx, cbs_info *out);

/* Plain device copy of `bytes` (read + write counted) repeated `reps`
 * times on the context's stream; returns the achieved GB/s.  Calibration
 * for the HBM counters and the achievable-bandwidth figure of the bench; no
 * reference counterpart. */
int cbs_hip_bandwidth_probe(cbs_hip_ctx *ctx, int64_t bytes, int32_t reps, double *gbs);

/* Waits for all work on the context's stream. */
int cbs_hip_synchronize(cbs_hip_ctx *ctx);

/* Device timing of the pressure applies inside pressure solves: when on,
 * every 8th apply sweep of a solve's iterations is bracketed by HIP events on
 * the context's stream (sweeps launched after the solve converged, which
 * return at once, are not counted).  cbs_hip_apply_timer returns the
 * accumulated milliseconds and the number of timed sweeps for one precision
 * (CBS_PCG_FP64 or CBS_PCG_MIXED for the FP32 sweeps) and optionally resets
 * both accumulators.  No reference counterpart: this is the measurement hook
 * of the roofline figure. */
int cbs_hip_set_profiling(cbs_hip_ctx *ctx, int32_t on);

/* Launch shape of the PCG apply for one precision, as chosen by the apply
 * autotuner at the context's first pressure solve (threads per row tile and
 * column pairs loaded per batch, returned in *prefetch), and the
 * autotuner's timing of each of its six candidates in milliseconds
 * ((1024,8), (1024,4), (512,8), (512,4), (256,8), (256,4); zeros before
 * tuning).  Any pointer may be null.  No reference counterpart. */
int cbs_hip_apply_config(cbs_hip_ctx *ctx, int32_t precision, int32_t *threads, int32_t *prefetch,
                         float *candidate_ms);
/* Fixes the launch shape of one precision's PCG apply instead of tuning it
 * (threads 256, 512 or 1024; batch 4 or 8 column pairs).  Results do not
 * depend on the shape: every row's sum and every tile's (p, Ap) partial is
 * taken in an order fixed by the operator's layout.  No reference
 * counterpart. */
int cbs_hip_set_apply_config(cbs_hip_ctx *ctx, int32_t precision, int32_t threads, int32_t batch);
/* How one solver form (precision CBS_PCG_FP64 or CBS_PCG_MIXED) forms the
 * PCG iteration's direction p = z + beta p: *fused 1 inside the apply, while
 * it stages p (one launch fewer per iteration), 0 in a kernel of its own
 * after the update, -1 not yet decided.  By default the first solve of a
 * context times both on its scratch vectors and keeps the faster:
 * *ms_separate and *ms_fused are those timings (apply plus direction, per
 * iteration; 0 before timing or when pinned).  Both ways give the same
 * directions bit for bit.  Any pointer may be null.  No reference
 * counterpart. */
int cbs_hip_direction_fusion(cbs_hip_ctx *ctx, int32_t precision, int32_t *fused, float *ms_separate,
                             float *ms_fused);
/* Pins the choice above (mode 1 fused, 0 separate) or returns it to the
 * timed default (mode -1).  No reference counterpart. */
int cbs_hip_set_direction_fusion(cbs_hip_ctx *ctx, int32_t precision, int32_t mode);
int cbs_hip_apply_timer(cbs_hip_ctx *ctx, int32_t precision, double *ms, int64_t *count, int32_t reset);

/* PCG iterations as HIP graphs (on by default): each solver form (FP64,
 * mixed) captures one block of 8 iterations (apply, update, convergence
 * test, direction) at its first solve and replays it, instead of launching
 * every kernel.  The block holds the same kernels with the same arguments
 * in the same order, so results are bit-identical either way; iterations
 * past convergence return at once in both.  cbs_hip_set_graphs(ctx, 0)
 * switches to per-kernel launches (and releases the graphs);
 * cbs_hip_graphs_active reports which forms currently replay a captured
 * block (a form whose capture failed falls back to launches).  No reference
 * counterpart. */
int cbs_hip_set_graphs(cbs_hip_ctx *ctx, int32_t on);

/* The CBS_PCG_AMG form (build-only; no reference counterpart; the reference
 * selects its pressure solvers through solver_opt, Steps.cpp:313-335):
 * the reference's CG recurrence and convergence test with z = M r, M one
 * V-cycle of smoothed-aggregation algebraic multigrid in FP32 (one damped
 * Jacobi sweep before and after each coarse correction, dense inverse on
 * the coarsest level), instead of Jacobi.  The hierarchy is built on the
 * host at the context's first CBS_PCG_AMG solve from the current operator;
 * later solves scale its coarse levels by the ratio of the current to the
 * set-up sum of pdiag (exact under a uniform time step; otherwise still a
 * symmetric positive definite preconditioner).  Single-case contexts only.
 * cbs_hip_amg_info reports the set-up time, the number of levels and the
 * stored nonzeros of the coarse operators, prolongations and restrictions
 * (zeros before the first such solve). */
int cbs_hip_amg_info(const cbs_hip_ctx *ctx, double *setup_seconds, int32_t *levels, int64_t *stored);
int cbs_hip_graphs_active(const cbs_hip_ctx *ctx, int32_t *fp64, int32_t *mixed);

/* The PCG applies the pressure operator of MatrixVectorCalc::pressureMultiply
 * (MatrixVectorCalc.cpp:125-205) in assembled form: each node row holds
 * a_ij = sum of gstif over the fluid elements holding edge (i, j) (summed in
 * a fixed element order at every solve) and pdiag on the diagonal, stored
 * as sliced ELLPACK over row tiles whose x is staged in LDS (16-bit
 * columns).  cbs_hip_operator_info reports its shape: row tiles, the
 * largest tile's staged nodes, staged nodes over all tiles (rows plus
 * halos), off-diagonal nonzeros, stored entries (with slice padding) and
 * the element contributions summed into them; after the first multi-RHS
 * solve also the tiles of its multi-column product (parts of the row tiles
 * whose staging fits 32 bytes per node) and the largest part's staged
 * nodes (both 0 before).  cbs_hip_apply_assembled
 * runs that apply on host vectors of length npoin with the context's
 * current gstif and pdiag: y = free ? A x : 0 (free = touches fluid, not a
 * fixed pressure node), in FP64 (CBS_PCG_FP64) or with FP32 values and
 * vectors (CBS_PCG_MIXED; x rounded to FP32, y returned widened).  On free
 * nodes it equals cbs_hip_pressure_multiply up to summation order.  No
 * reference counterpart: the shape feeds the roofline, the apply is a
 * test hook. */
typedef struct cbs_operator_info {
    int32_t row_tiles, max_staged;
    int64_t staged_nodes, nnz, entries, contributions;
    int32_t product_parts, product_max_staged;
} cbs_operator_info;
int cbs_hip_operator_info(const cbs_hip_ctx *ctx, cbs_operator_info *out);
int cbs_hip_apply_assembled(cbs_hip_ctx *ctx, int32_t precision, const double *x, double *y);

/* ---- Adjoint of the CBS step (build-only; no reference counterpart) ----
 *
 * Forward-model extensions, off by default.  With both off every entry point
 * computes the reference's step bit for bit (the extension code is not on
 * the path).
 *  - cbs_hip_set_body_force: a nodal body force f (3, N), an acceleration:
 *    the Step 1 predictor (Steps.cpp:132-142) becomes
 *    u* = u1 + (rhs + m f) elcoe2 with m the node's lumped fluid mass, i.e.
 *    u* gains dt f.  f = NULL switches it off.
 *  - cbs_hip_set_inflow_coefficients: 16 coefficients c added to the BC-508
 *    inflow profile 6y(1-y) (Boundary.cpp:180-183):
 *    u_x = 6y(1-y) + sum_q c_q phi_q, phi_(4(j-1)+k-1) = sin(j pi yh) sin(k pi zh),
 *    j, k = 1..4, (yh, zh) the inlet coordinates normalised to its bounding
 *    box; nodes whose last prescription is not 508 (walls) are unchanged.
 *    n = 0 restores the profile.  cbs_hip_inflow_basis returns the affected
 *    nodes (1-based) and phi at each (16 per node).
 * Adjoint run (dtfixed = 1 required: the minimum of the stability time step
 * over elements is not differentiable):
 *  - cbs_hip_adjoint_forward applies Solver::initialise's boundary values
 *    (Solver.cpp:296-300) to the current fields, advances nsteps CBS steps
 *    (the energy step runs if enabled but does not enter the objective) with
 *    every pressure solve converged to rel_tol, keeps the state every
 *    checkpoint_every steps, and returns
 *      Phi = sum over the last objective_steps steps of
 *            sum_e nu_e V_e sum_ij (du_i/dx_j)^2  (fluid elements, P1 gradients).
 *  - cbs_hip_adjoint_backward returns seed * dPhi/df (3, N) and
 *    seed * dPhi/dc (16) by the discrete adjoint: transposed boundary
 *    packages, Step 3, Step 2 RHS and Step 1 element terms (hand-written
 *    vector-Jacobian products) and one adjoint pressure solve per step with
 *    the same operator; segments between checkpoints are recomputed.  The
 *    context is left at the end of the forward run. */
typedef struct cbs_adjoint_options {
    int32_t nsteps;             /* CBS steps of the run */
    int32_t objective_steps;    /* Phi sums over the last this-many steps */
    int32_t checkpoint_every;   /* 0: 10 */
    double rel_tol;             /* relative residual of every forward and adjoint pressure solve; 0: 1e-12 */
} cbs_adjoint_options;

int cbs_hip_set_body_force(cbs_hip_ctx *ctx, const double *force);
int cbs_hip_set_inflow_coefficients(cbs_hip_ctx *ctx, const double *coef, int32_t n);
int cbs_hip_inflow_basis(cbs_hip_ctx *ctx, int32_t *nnodes, int32_t *node, double *phi);
int cbs_hip_adjoint_forward(cbs_hip_ctx *ctx, const cbs_adjoint_options *opts, double *objective);
int cbs_hip_adjoint_backward(cbs_hip_ctx *ctx, double seed, double *grad_force, double *grad_inflow);

/* Solver::advanceOneStep (Solver.cpp:609-680): one full CBS step. */
int cbs_hip_advance_one_step(cbs_hip_ctx *ctx, int32_t iitime, const cbs_cg_options *opts,
                             cbs_cg_result *cg);

#ifdef __cplusplus
}
#endif
#endif
