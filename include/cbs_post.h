/*
 * cbs_post.h - C ABI of the solution output of a CBS run, taken off the time
 * step's critical path.
 *
 * Replaces the file output of Post (src/io/Post.cpp, include/cbs/io/Post.hpp)
 * with the same directory layout, file names, residual CSV text and PVD text,
 * and VTU files that carry the same arrays (names, types, order, masks,
 * values) as the reference's, written as VTK XML with the arrays in one
 * appended raw-binary block instead of ASCII:
 *
 *   <root>/<case>/<case>_residuals.csv          Post.cpp:534-558, 768-812
 *   <root>/<case>/<case>_step_XXXXXXXX.vtu      Post.cpp:887-1130
 *   <root>/<case>/<case>.pvd                    Post.cpp:1133-1182
 *
 * The reference formats every value as text with 16 significant digits; the
 * binary arrays hold the doubles themselves, so formatting one with "%.16g"
 * gives the reference's text for it.
 *
 * The writer runs on the host.  Device fields reach it through snapshots
 * (cbs_hip_snapshot_* in cbs_hip.h): a snapshot copies the fields within
 * device memory on the compute stream, then to pinned host memory on a copy
 * stream while the next steps run; cbs_post_write_solution_async hands the
 * snapshot to a background thread that waits for the copy, encodes and writes
 * the file, rewrites the PVD, and releases the snapshot.
 *
 * Conventions are those of cbs_hip.h (reference array layout with element 0
 * dropped, 1-based ids inside arrays, cbs_status codes).  One host thread
 * drives a writer; the background thread is internal.
 */
#ifndef CBS_POST_H
#define CBS_POST_H

#include <stdint.h>

#include "cbs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cbs_post cbs_post;

/* Output controls of RunConfig (include/cbs/core/RunConfig.hpp:160-173),
 * read from .par line 44 (src/io/MeshIO.cpp:321-356). */
typedef struct cbs_post_config {
    int32_t residual_log_enabled;          /* write <case>_residuals.csv */
    int32_t residual_log_every;            /* every k-th iteration when k > 1 */
    int32_t vtu_output_enabled;
    int32_t vtu_output_every_iterations;   /* 0: off */
    double vtu_output_every_sim_time;      /* transient runs; 0: off */
    int32_t transient_on;
    int32_t write_boundary_debug_arrays;   /* adds the node_bc_id_debug array */
} cbs_post_config;

/* The mesh arrays the VTU files hold (layouts as in cbs_mesh_desc). */
typedef struct cbs_post_mesh {
    int32_t npoin, nelem, nboun;
    const double *coord;       /* (3, N) */
    const int32_t *intma;      /* (4, E), 1-based node ids */
    const int32_t *iside;      /* (6, B): n1 n2 n3 local_face parent bc */
    const int32_t *mat_elem;   /* (E) 0 fluid, > 0 solid */
} cbs_post_mesh;

/* The solution arrays of one output (borrowed until the write completes). */
typedef struct cbs_post_fields {
    const double *pres;         /* (N) */
    const double *temperature;  /* (N) */
    const double *unkno;        /* (3, N) velocity */
    const double *velocity;     /* (N) velocity magnitude (Solver.cpp:689-699) */
} cbs_post_fields;

/* Callbacks of an asynchronous write, called on the writer thread:
 * ready(user) before the fields are read (e.g. cbs_hip_snapshot_wait, which
 * returns nonzero on failure), done(user) when they are no longer needed
 * (e.g. cbs_hip_snapshot_release).  done is called exactly once, also after
 * a failure. */
typedef int (*cbs_post_ready_fn)(void *user);
typedef int (*cbs_post_done_fn)(void *user);

/* Message of the last failing cbs_post_* call in this thread, or of the first
 * failed background write (reported by cbs_post_flush). */
const char *cbs_post_last_error(void);

/* Builds the static part of every VTU file (points, cells, node masks, cell
 * data; Post.cpp:436-514, 907-1011, 1070-1122) once per run.  root is the
 * output root (the reference's is "output", Post.cpp:61-64).  Fails with
 * CBS_ERR_BAD_ARGUMENT and "Post::writeVTU - element connectivity node out of
 * range" when intma holds an id outside [1, npoin]. */
int cbs_post_create(const cbs_post_mesh *mesh, const cbs_post_config *cfg, const char *root,
                    const char *case_name, cbs_post **out);

/* Post::initialiseRunOutputs (Post.cpp:731-754): creates <root>/<case>/,
 * resets the output history, writes the residual CSV header (when enabled)
 * and, when VTU output is enabled, the iteration-0 solution from `fields`
 * (written before the call returns). */
int cbs_post_initialise_run_outputs(cbs_post *post, double rtime, const cbs_post_fields *fields);

/* Post::writeResidualRow (Post.cpp:768-812): appends the row of iteration
 * iitime from the step's scalars (cbs_hip_get_scalars) when the residual log
 * is enabled and iitime is due. */
int cbs_post_write_residual_row(cbs_post *post, int32_t iitime, const cbs_scalars *sc);

/* Post::shouldWriteVTU (Post.cpp:821-840), with the same side effect on the
 * next simulated-time output.  *due receives 0 or 1. */
int cbs_post_should_write_vtu(cbs_post *post, int32_t iitime, double rtime, int32_t *due);

/* Post::writeSolution (Post.cpp:842-860): writes the VTU of iteration iitime,
 * appends (iitime, rtime) to the output history and rewrites the PVD.  The
 * call returns with the files written. */
int cbs_post_write_solution(cbs_post *post, int32_t iitime, double rtime, const cbs_post_fields *fields);

/* The same, on the writer thread: the history entry is recorded now (so
 * cbs_post_last_output sees it), the VTU is written to a temporary name and
 * renamed when complete, and the PVD is rewritten after it.  `fields` and the
 * arrays it points to must stay valid until done(user) is called. */
int cbs_post_write_solution_async(cbs_post *post, int32_t iitime, double rtime, const cbs_post_fields *fields,
                                  cbs_post_ready_fn ready, cbs_post_done_fn done, void *user);

/* Iteration of the last recorded output, or -1 (Solver.cpp:211-221 writes the
 * final solution unless it already is the last one). */
int cbs_post_last_output(cbs_post *post, int32_t *iitime);

/* Waits for every pending write; returns the status of the first failed
 * background write since the previous flush. */
int cbs_post_flush(cbs_post *post);

/* Bytes and seconds spent by the writer: VTU bytes written, seconds spent
 * encoding and writing VTU files (background and foreground). */
int cbs_post_stats(cbs_post *post, int64_t *vtu_files, int64_t *vtu_bytes, double *write_seconds);

/* Flushes and frees the writer. */
void cbs_post_destroy(cbs_post *post);

#ifdef __cplusplus
}
#endif

#endif /* CBS_POST_H */
