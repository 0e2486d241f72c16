/*
 * cbs_case.h - C ABI of the host-side case set-up that feeds cbs_hip.h.
 *
 * A case holds the CBSStateSI arrays a run needs before its first step:
 * mesh, materials, initial fields, and the one-time geometry and boundary
 * classification of Solver::initialise (src/solver/Solver.cpp:242-302):
 * Preprocess::* (src/preprocess/Preprocess.cpp) and Coloring::build
 * (include/cbs/parallel/Coloring.hpp:27-97).  The host code is C++20 with
 * OpenMP on the element loops whose results do not depend on order.
 *
 * Arrays follow the conventions of cbs_hip.h (reference layout, element 0
 * dropped, 1-based ids inside arrays).
 */
#ifndef CBS_CASE_H
#define CBS_CASE_H

#include <stdint.h>

#include "cbs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cbs_case cbs_case;

/* One row of the .matprop file (src/io/MeshIO.cpp:1523-1623). */
typedef struct cbs_material {
    int32_t id;
    double rho, cp, k, mu, qvol;
} cbs_material;

/* Message of the last failing cbs_case_* call in this thread. */
const char *cbs_case_last_error(void);

/* Builds a case from in-memory mesh data, as MeshIO::readAll would leave it
 * (src/io/MeshIO.cpp:572-609 and MeshIO::initialiseFields :1627-1701).
 *   par       .par controls (only the input fields are read; ani, temp_calc
 *             and the derived fields are set here)
 *   coord     (3, npoin)
 *   intma     (4, nelem) positively oriented tetrahedra
 *   bface     (5, nboun): n1 n2 n3 parent solver_bc (after the .bco mapping)
 *   mat_elem  (nelem) material ids, NULL when material files are disabled
 *   mats      nmat material rows, NULL when material files are disabled */
int cbs_case_create(const cbs_config *par, int32_t nelem, int32_t npoin, int32_t nboun,
                    const double *coord, const int32_t *intma, const int32_t *bface,
                    const int32_t *mat_elem, int32_t nmat, const cbs_material *mats,
                    cbs_case **out);

/* Reads <base>.plt/.bco/.par[/.material/.matprop] (MeshIO::readAll). */
int cbs_case_read(const char *case_base, cbs_case **out);

/* Solver::initialise geometry and classification: validateBoundaryFlags,
 * shapeFunctionDerivatives, assignBoundaryFaceNumbers, getNormals,
 * massMatrix, classifyFaceEdges, elementSize, wallDetermination,
 * computeMassFlowInletVelocity, detectPressureBoundaryNodes, Coloring::build.
 * The initial boundary values are applied on the device afterwards
 * (cbs_hip_apply_boundary). */
int cbs_case_setup(cbs_case *c);

/* Effective run controls and the mesh descriptor for cbs_hip_create.  The
 * descriptor points into the case and stays valid while the case lives. */
int cbs_case_config(const cbs_case *c, cbs_config *out);
int cbs_case_mesh_desc(const cbs_case *c, cbs_mesh_desc *out);

/* Binary mesh files.  <base>.cbsmesh holds the content of <base>.plt (and
 * of <base>.material when that file exists) as raw little-endian arrays with
 * a checksum: magic "CBSMESH1", uint32 version 1, uint32 flags (bit 0:
 * material ids), int32 nelem, npoin, nboun, 0; double coord[3 npoin];
 * int32 intma[4 nelem]; int32 bface[5 nboun] (n1 n2 n3 parent raw_side_id);
 * [int32 mat[nelem]]; uint64 checksum.  cbs_case_read reads it in place of
 * the ASCII files when it exists (.par, .bco and .matprop stay text), with
 * the same result, and refuses one older than a <base>.plt or
 * <base>.material beside it.  cbs_case_write_mesh_binary converts <base>.plt (and
 * .material) into <base>.cbsmesh and refuses to overwrite one;
 * cbs_case_write_mesh_arrays_binary writes one from arrays (mat_elem may be
 * NULL). */
int cbs_case_write_mesh_binary(const char *base);
int cbs_case_write_mesh_arrays_binary(const char *path, int32_t nelem, int32_t npoin, int32_t nboun,
                                      const double *coord, const int32_t *intma, const int32_t *bface,
                                      const int32_t *mat_elem);

/* cbs_case_setup without its last phase, Coloring::build: for a caller that
 * colours on the device (cbs_hip_coloring in cbs_hip.h) and hands the
 * result to cbs_case_set_coloring. */
int cbs_case_setup_without_coloring(cbs_case *c);

/* Installs a colouring (colour-major element ids, 1-based; 0-based offsets
 * color_ptr[0..ncolor]).  Checks that the classes are non-empty and
 * partition the elements; it does not check that they are the reference's
 * first-fit colouring (cbs_hip_coloring's are). */
int cbs_case_set_coloring(cbs_case *c, int32_t ncolor, const int32_t *color_ptr, const int32_t *color_elem);

/* Borrow an array by name ("unkno", "pres", "temperature", "dNkdx", ...):
 * dtype is 'd' (double) or 'i' (int32).  "setup_seconds" holds the wall-clock
 * seconds of the phases of the last cbs_case_setup, in Solver::initialise's
 * order (validate BCs, shape functions, boundary faces, normals, mass matrix,
 * face edges, element size, walls, mass-flow inlet, pressure nodes,
 * colouring). */
int cbs_case_array(const cbs_case *c, const char *name, const void **data, int64_t *count,
                   char *dtype);

void cbs_case_destroy(cbs_case *c);

#ifdef __cplusplus
}
#endif
#endif
