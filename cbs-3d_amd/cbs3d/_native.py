"""ctypes declarations of include/cbs_hip.h, include/cbs_case.h and include/cbs_post.h."""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libcbs3d_host.so")
HIP_LIB = os.path.join(HERE, "libcbs3d_hip.so")

_I = C.c_int32
_D = C.c_double
_PI = C.POINTER(C.c_int32)
_PD = C.POINTER(C.c_double)

CBS_OK = 0
# CBS_HIP_ABI_VERSION of include/cbs_hip.h that these structures mirror
HIP_ABI_VERSION = 2
STATUS_NAMES = {
    0: "CBS_OK", 1: "CBS_ERR_BAD_ARGUMENT", 2: "CBS_ERR_HIP", 3: "CBS_ERR_CG_NOT_CONVERGED",
    4: "CBS_ERR_CG_BREAKDOWN", 5: "CBS_ERR_JACOBI_DIAGONAL", 6: "CBS_ERR_NONFINITE_RESIDUAL",
    7: "CBS_ERR_INVALID_TIMESTEP", 8: "CBS_ERR_INVALID_GEOMETRY", 9: "CBS_ERR_UNSUPPORTED",
    10: "CBS_ERR_INVALID_MATERIAL", 11: "CBS_ERR_IO",
}


class Config(C.Structure):
    _fields_ = [(n, _I) for n in (
        "temp_calc", "cbs_scheme", "ilots", "htype", "dtfixed", "dtfix_end", "transient_on",
        "solver_opt", "rem_deltp", "beta_opt", "step2_check",
        "dimensional_mode", "material_properties_enabled", "mass_flow_inlet_enabled",
        "convection_type", "ntime", "cg_preconditioner", "cg_conv_test", "cg_max_iter", "pnode")] + \
        [(n, _D) for n in (
            "csafm", "dtfix", "deltr", "csafm2", "epsilon1", "re", "pr", "ri", "ani",
            "relToler", "absToler", "inlet_u", "inlet_v", "inlet_w", "inlet_temperature",
            "inlet_density", "inlet_mass_flow_rate", "inlet_u_from_massflow",
            "outlet_pressure_gauge", "heat_flux_bc")] + [("cinf", _D * 5)]

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n, _ in self._fields_ if n != "cinf"}
        d["cinf"] = tuple(self.cinf)
        return d


class MeshDesc(C.Structure):
    _fields_ = [(n, _I) for n in ("npoin", "nelem", "nboun", "ncolor", "npoin_wall", "bc_fixed")] + [
        ("intma", _PI), ("coord", _PD), ("mat_elem", _PI),
        ("rho_e", _PD), ("mu_e", _PD), ("k_e", _PD), ("rho_cp_e", _PD), ("alpha_e", _PD), ("Qvol_e", _PD),
        ("dNkdx", _PD), ("detJ", _PD), ("annxf", _PD), ("fedge", _PI), ("elcoe_e", _PD), ("alen_e", _PD),
        ("iside", _PI), ("face_norm", _PD), ("wall_node_list", _PI), ("bc_list", _PI), ("bc_values", _PD),
        ("color_ptr", _PI), ("color_elem", _PI)]


class Material(C.Structure):
    _fields_ = [("id", _I), ("rho", _D), ("cp", _D), ("k", _D), ("mu", _D), ("qvol", _D)]


class AdjointOptions(C.Structure):
    _fields_ = [("nsteps", _I), ("objective_steps", _I), ("checkpoint_every", _I), ("rel_tol", _D)]


class CgOptions(C.Structure):
    _fields_ = [("precision", _I), ("max_refinements", _I), ("inner_rel_tol", _D)]


class CgResult(C.Structure):
    _fields_ = [("iterations", _I), ("converged", _I), ("initial_l2", _D), ("final_l2", _D),
                ("final_relative_l2", _D), ("final_max_abs", _D), ("total_seconds", _D),
                ("setup_seconds", _D), ("matvec_seconds", _D), ("dot_seconds", _D),
                ("preconditioner_seconds", _D), ("vector_update_seconds", _D), ("constraint_seconds", _D),
                ("matvec_calls", _I), ("dot_calls", _I),
                ("preconditioner_calls", _I), ("refinements", _I), ("algorithmic_bytes", _D)]


class Info(C.Structure):
    _fields_ = [("npoin", _I), ("nelem", _I), ("nfluid", _I), ("ncolor", _I), ("device_bytes", C.c_int64)]


class OperatorInfo(C.Structure):
    """Shape of the assembled operator the PCG applies (cbs_operator_info)."""
    _fields_ = [("row_tiles", _I), ("max_staged", _I), ("staged_nodes", C.c_int64), ("nnz", C.c_int64),
                ("entries", C.c_int64), ("contributions", C.c_int64), ("product_parts", _I),
                ("product_max_staged", _I)]


class Scalars(C.Structure):
    _fields_ = [("dtreal", _D), ("rtime", _D), ("hb", _D * 15), ("last_cg_iterations", _I),
                ("last_cg_initial_l2", _D), ("last_cg_final_l2", _D), ("last_cg_relative_l2", _D),
                ("last_cg_max_abs", _D)]


class PostConfig(C.Structure):
    _fields_ = [("residual_log_enabled", _I), ("residual_log_every", _I), ("vtu_output_enabled", _I),
                ("vtu_output_every_iterations", _I), ("vtu_output_every_sim_time", _D), ("transient_on", _I),
                ("write_boundary_debug_arrays", _I)]


class PostMesh(C.Structure):
    _fields_ = [("npoin", _I), ("nelem", _I), ("nboun", _I), ("coord", _PD), ("intma", _PI), ("iside", _PI),
                ("mat_elem", _PI)]


class PostFields(C.Structure):
    _fields_ = [("pres", _PD), ("temperature", _PD), ("unkno", _PD), ("velocity", _PD)]


FIELDS = ["unkno", "unkn1", "pres", "pres1", "temperature", "temperature1", "rhs", "rhs1",
          "delte", "deltp", "deltp1", "deltp2", "beta1", "vvis", "velcp", "alen",
          "unkno_unit_vec", "elcoe2", "elcoe2p", "pdiag", "gstif", "gstifE", "pdiagE", "velocity"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

_host = None
_hip = None


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(HOST_LIB)
        lib.cbs_case_last_error.restype = C.c_char_p
        lib.cbs_case_create.argtypes = [C.POINTER(Config), _I, _I, _I, _PD, _PI, _PI, _PI, _I,
                                        C.POINTER(Material), C.POINTER(C.c_void_p)]
        lib.cbs_case_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        lib.cbs_case_setup.argtypes = [C.c_void_p]
        lib.cbs_case_config.argtypes = [C.c_void_p, C.POINTER(Config)]
        lib.cbs_case_mesh_desc.argtypes = [C.c_void_p, C.POINTER(MeshDesc)]
        lib.cbs_case_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_char)]
        lib.cbs_case_destroy.argtypes = [C.c_void_p]
        lib.cbs_case_destroy.restype = None
        lib.cbs_case_setup_without_coloring.argtypes = [C.c_void_p]
        lib.cbs_case_set_coloring.argtypes = [C.c_void_p, _I, _PI, _PI]
        P = C.c_void_p
        lib.cbs_post_last_error.restype = C.c_char_p
        lib.cbs_post_create.argtypes = [C.POINTER(PostMesh), C.POINTER(PostConfig), C.c_char_p, C.c_char_p,
                                        C.POINTER(P)]
        lib.cbs_post_initialise_run_outputs.argtypes = [P, _D, C.POINTER(PostFields)]
        lib.cbs_post_write_residual_row.argtypes = [P, _I, C.POINTER(Scalars)]
        lib.cbs_post_should_write_vtu.argtypes = [P, _I, _D, _PI]
        lib.cbs_post_write_solution.argtypes = [P, _I, _D, C.POINTER(PostFields)]
        lib.cbs_post_write_solution_async.argtypes = [P, _I, _D, C.POINTER(PostFields), P, P, P]
        lib.cbs_post_last_output.argtypes = [P, _PI]
        lib.cbs_post_flush.argtypes = [P]
        lib.cbs_post_stats.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        lib.cbs_post_destroy.argtypes = [P]
        lib.cbs_post_destroy.restype = None
        _host = lib
    return _host


def hip_lib():
    """Load the HIP library.  Raises when it is missing: there is no CPU fallback."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(f"{HIP_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(HIP_LIB)
        P = C.c_void_p
        lib.cbs_hip_abi_version.restype = _I
        lib.cbs_hip_abi_version.argtypes = []
        if lib.cbs_hip_abi_version() != HIP_ABI_VERSION:
            raise RuntimeError(f"{HIP_LIB} has C-ABI version {lib.cbs_hip_abi_version()}, this binding expects "
                               f"{HIP_ABI_VERSION}; rebuild with __graft_entry__.build()")
        lib.cbs_hip_last_error.restype = C.c_char_p
        lib.cbs_hip_last_error.argtypes = [P]
        lib.cbs_hip_set_tile_target.restype = _I
        lib.cbs_hip_set_tile_target.argtypes = [_I]
        lib.cbs_hip_set_line_group_min_elems.restype = _I
        lib.cbs_hip_set_line_group_min_elems.argtypes = [_I]
        lib.cbs_hip_field_size.restype = C.c_int64
        lib.cbs_hip_field_size.argtypes = [P, _I]
        lib.cbs_hip_create.argtypes = [C.POINTER(Config), C.POINTER(MeshDesc), C.POINTER(P)]
        lib.cbs_hip_destroy.argtypes = [P]
        lib.cbs_hip_set_field.argtypes = [P, _I, _PD, C.c_int64]
        lib.cbs_hip_get_field.argtypes = [P, _I, _PD, C.c_int64]
        lib.cbs_hip_get_scalars.argtypes = [P, C.POINTER(Scalars)]
        lib.cbs_hip_set_dtreal.argtypes = [P, _D]
        for n in ("build_element_pressure_terms", "update_lhs_diagonal", "assemble_step1_rhs",
                  "step1", "assemble_step2_rhs", "step3", "assemble_step4_rhs", "step4",
                  "update_velocity_magnitude", "convergence"):
            getattr(lib, "cbs_hip_" + n).argtypes = [P]
        lib.cbs_hip_compute_timestep.argtypes = [P, _I]
        lib.cbs_hip_apply_boundary.argtypes = [P, _I]
        lib.cbs_hip_pressure_multiply.argtypes = [P, _PD, _PD]
        lib.cbs_hip_pressure_multiply_device.argtypes = [P, C.c_void_p, C.c_void_p]
        lib.cbs_hip_solve_pressure.argtypes = [P, C.POINTER(CgOptions), C.POINTER(CgResult)]
        lib.cbs_hip_solve_pressure_multi.argtypes = [P, _I, _PD, _PD, C.POINTER(CgOptions), C.POINTER(CgResult)]
        lib.cbs_hip_step2.argtypes = [P, C.POINTER(CgOptions), C.POINTER(CgResult)]
        lib.cbs_hip_info.argtypes = [P, C.POINTER(Info)]
        lib.cbs_hip_bandwidth_probe.argtypes = [P, C.c_int64, _I, C.POINTER(C.c_double)]
        lib.cbs_hip_synchronize.argtypes = [P]
        lib.cbs_hip_set_profiling.argtypes = [P, _I]
        lib.cbs_hip_apply_config.argtypes = [P, _I, _PI, _PI, C.POINTER(C.c_float)]
        lib.cbs_hip_operator_info.argtypes = [P, C.POINTER(OperatorInfo)]
        lib.cbs_hip_apply_assembled.argtypes = [P, _I, _PD, _PD]
        lib.cbs_hip_set_graphs.argtypes = [P, _I]
        lib.cbs_hip_amg_info.argtypes = [P, _PD, _PI, C.POINTER(C.c_int64)]
        lib.cbs_hip_set_apply_config.argtypes = [P, _I, _I, _I]
        lib.cbs_hip_direction_fusion.argtypes = [P, _I, _PI, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.cbs_hip_set_direction_fusion.argtypes = [P, _I, _I]
        lib.cbs_hip_graphs_active.argtypes = [P, _PI, _PI]
        lib.cbs_hip_create_batch.argtypes = [_I, C.POINTER(Config), C.POINTER(MeshDesc), C.POINTER(P)]
        lib.cbs_hip_member_count.argtypes = [P, _PI]
        lib.cbs_hip_member_range.argtypes = [P, _I, _PI, _PI, _PI, _PI]
        lib.cbs_hip_get_member_scalars.argtypes = [P, _I, C.POINTER(Scalars)]
        lib.cbs_hip_get_member_cg_result.argtypes = [P, _I, C.POINTER(CgResult)]
        lib.cbs_hip_set_body_force.argtypes = [P, C.c_void_p]
        lib.cbs_hip_set_inflow_coefficients.argtypes = [P, C.c_void_p, _I]
        lib.cbs_hip_inflow_basis.argtypes = [P, _PI, C.c_void_p, C.c_void_p]
        lib.cbs_hip_adjoint_forward.argtypes = [P, C.POINTER(AdjointOptions), C.POINTER(C.c_double)]
        lib.cbs_hip_adjoint_backward.argtypes = [P, _D, C.c_void_p, C.c_void_p]
        lib.cbs_hip_apply_timer.argtypes = [P, _I, _PD, C.POINTER(C.c_int64), _I]
        lib.cbs_hip_advance_one_step.argtypes = [P, _I, C.POINTER(CgOptions), C.POINTER(CgResult)]
        lib.cbs_hip_coloring.argtypes = [_I, _I, _PI, _PI, _PI, _PI, _PI]
        # device pointers for intma, color_ptr and color_elem
        lib.cbs_hip_coloring_noalloc.argtypes = [_I, _I, C.c_void_p, _PI, C.c_void_p, C.c_void_p, _PI]
        lib.cbs_hip_coloring_noalloc_scratch_bytes.argtypes = []
        lib.cbs_hip_coloring_noalloc_scratch_bytes.restype = C.c_int64
        lib.cbs_hip_create_seconds.argtypes = [P, _PD, _I]
        lib.cbs_hip_snapshot_begin.argtypes = [P, C.POINTER(P)]
        lib.cbs_hip_snapshot_wait.argtypes = [P]
        lib.cbs_hip_snapshot_field.argtypes = [P, _I]
        lib.cbs_hip_snapshot_field.restype = _PD
        lib.cbs_hip_snapshot_release.argtypes = [P]
        _hip = lib
    return _hip
