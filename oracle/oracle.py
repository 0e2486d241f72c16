"""ctypes binding of the CPU oracle (oracle/libcbs_oracle.so).  TEST INFRASTRUCTURE.

Only the tests, the smoke check and bench.py's CPU-baseline leg use this
module, as the checker.  `OracleState` wraps the C `orc_state`: it owns numpy
arrays laid out like the reference's arrays (0-based memory, 1-based ids
inside intma/iside/bc_list/wall_node_list/color_elem) and passes pointers.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcbs_oracle.so")
# the same source built with the reference's OpenMP pragmas active (the CPU
# baseline's timing variant; not bit-identical above one thread)
OMP_LIB_PATH = os.path.join(HERE, "libcbs_oracle_omp.so")

_I = C.c_int32
_D = C.c_double
_PI = C.POINTER(C.c_int32)
_PD = C.POINTER(C.c_double)

INT_FIELDS = ["npoin", "nelem", "nboun", "ncolor", "bc_fixed", "npoin_wall",
              "temp_calc", "cbs_scheme", "ilots", "htype", "dtfixed", "dtfix_end", "transient_on",
              "solver_opt", "rem_deltp", "beta_opt", "step2_check",
              "dimensional_mode", "material_properties_enabled",
              "cg_preconditioner", "cg_conv_test", "cg_max_iter"]
DBL_FIELDS = ["csafm", "dtfix", "deltr", "ani", "relToler", "absToler", "dtreal", "rtime",
              "csafm2", "epsilon1",
              "inlet_u", "inlet_v", "inlet_w", "inlet_temperature", "inlet_u_from_massflow",
              "heat_flux_bc"]
INT_ARRAYS = ["intma", "iside", "fedge", "mat_elem", "wall_node_list", "bc_list",
              "color_ptr", "color_elem"]
DBL_CONST = ["coord", "dNkdx", "detJ", "annxf", "face_norm", "elcoe_e", "alen_e", "bc_values",
             "rho_e", "mu_e", "k_e", "rho_cp_e", "alpha_e", "Qvol_e", "pdiagE", "gstifE"]
# (name, components per node or element, 'n' node / 'e' element / 'h' fixed)
DBL_STATE = [("unkno", 3, "n"), ("unkn1", 3, "n"), ("pres", 1, "n"), ("pres1", 1, "n"),
             ("temperature", 1, "n"), ("temperature1", 1, "n"), ("rhs", 3, "n"), ("rhs1", 1, "n"),
             ("delte", 1, "e"), ("deltp", 1, "n"), ("deltp1", 1, "n"), ("deltp2", 1, "n"),
             ("beta1", 1, "n"), ("vvis", 1, "n"), ("velcp", 1, "n"), ("alen", 1, "n"),
             ("unkno_unit_vec", 3, "n"), ("elcoe2", 1, "n"), ("elcoe2p", 1, "n"),
             ("pdiag", 1, "n"), ("gstif", 6, "e"), ("velocity", 1, "n"), ("hb", 15, "h")]


class _OrcState(C.Structure):
    _fields_ = ([(n, _I) for n in INT_FIELDS] + [(n, _D) for n in DBL_FIELDS]
                + [(n, _PI) for n in INT_ARRAYS]
                + [(n, _PD) for n in DBL_CONST[:14]] + [("pdiagE", _PD), ("gstifE", _PD)]
                + [(n, _PD) for n, _, _ in DBL_STATE]
                + [("last_cg_iterations", _I), ("last_cg_initial_l2", _D), ("last_cg_final_l2", _D),
                   ("last_cg_relative_l2", _D), ("last_cg_max_abs", _D)])


class CgResult(C.Structure):
    _fields_ = [("iterations", _I), ("converged", _I), ("initial_l2", _D), ("final_l2", _D),
                ("final_relative_l2", _D), ("final_max_abs", _D)]


_libs = {}


def build() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib(openmp: bool = False):
    """The oracle library: serial (the checker) or its OpenMP build (timing)."""
    if openmp not in _libs:
        path = OMP_LIB_PATH if openmp else LIB_PATH
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_coloring.restype = _I
        L.orc_coloring.argtypes = [_I, _I, _PI, _PI, _PI]
        for name in ["orc_build_element_pressure_terms", "orc_update_lhs_diagonal",
                     "orc_assemble_step1_rhs", "orc_step1", "orc_assemble_step2_rhs",
                     "orc_step2", "orc_step3", "orc_assemble_step4_rhs", "orc_step4",
                     "orc_apply_velocity_package", "orc_apply_velocity", "orc_apply_temperature",
                     "orc_apply_pressure", "orc_convergence"]:
            getattr(L, name).argtypes = [C.POINTER(_OrcState)]
        L.orc_compute_timestep.argtypes = [C.POINTER(_OrcState), _I]
        L.orc_advance_one_step.argtypes = [C.POINTER(_OrcState), _I]
        L.orc_pressure_multiply.argtypes = [C.POINTER(_OrcState), _PD, _PD]
        L.orc_solve_pressure.argtypes = [C.POINTER(_OrcState), C.POINTER(CgResult)]
        L.orc_update_velocity_magnitude.argtypes = [C.POINTER(_OrcState)]
        _libs[openmp] = L
    return _libs[openmp]


class OracleError(RuntimeError):
    pass


def coloring(intma: np.ndarray, npoin: int):
    """Coloring::build: returns (ncolor, color_ptr, color_elem)."""
    intma = np.ascontiguousarray(intma, dtype=np.int32).reshape(-1)
    nelem = intma.size // 4
    ptr = np.zeros(nelem + 1, np.int32)
    elem = np.zeros(nelem, np.int32)
    nc = lib().orc_coloring(nelem, npoin, intma.ctypes.data_as(_PI), ptr.ctypes.data_as(_PI),
                            elem.ctypes.data_as(_PI))
    return int(nc), ptr[: nc + 1].copy(), elem


class OracleState:
    """The reference's CBSStateSI arrays for the hot path, driven by the C oracle."""

    def __init__(self, arrays: dict, cfg: dict, openmp: bool = False):
        self.a = {}
        self._lib = lib(openmp)
        npoin, nelem = int(cfg["npoin"]), int(cfg["nelem"])
        self._s = _OrcState()
        for n in INT_FIELDS:
            setattr(self._s, n, int(cfg.get(n, 0)))
        for n in DBL_FIELDS:
            setattr(self._s, n, float(cfg.get(n, 0.0)))
        for n in INT_ARRAYS:
            v = arrays.get(n)
            v = np.zeros(1, np.int32) if v is None or np.size(v) == 0 else v
            self.a[n] = np.ascontiguousarray(v, dtype=np.int32).reshape(-1).copy()
            setattr(self._s, n, self.a[n].ctypes.data_as(_PI))
        for n in DBL_CONST:
            v = arrays.get(n)
            if v is None or np.size(v) == 0:
                v = np.zeros({"pdiagE": 4 * nelem, "gstifE": 6 * nelem}.get(n, 1))
            self.a[n] = np.ascontiguousarray(v, dtype=np.float64).reshape(-1).copy()
            setattr(self._s, n, self.a[n].ctypes.data_as(_PD))
        for n, comp, kind in DBL_STATE:
            size = comp * (npoin if kind == "n" else nelem if kind == "e" else 1)
            v = arrays.get(n)
            arr = np.zeros(size) if v is None else np.ascontiguousarray(v, np.float64).reshape(-1).copy()
            assert arr.size == size, (n, arr.size, size)
            self.a[n] = arr
            setattr(self._s, n, arr.ctypes.data_as(_PD))

    def __getitem__(self, name):
        return self.a[name]

    def __getattr__(self, name):
        if name in INT_FIELDS or name in DBL_FIELDS or name.startswith("last_cg"):
            return getattr(self._s, name)
        raise AttributeError(name)

    def _call(self, fn, *args):
        rc = getattr(self._lib, fn)(C.byref(self._s), *args)
        if rc:
            raise OracleError(self._lib.orc_last_error().decode())

    def build_element_pressure_terms(self):
        self._call("orc_build_element_pressure_terms")

    def compute_timestep(self, it):
        self._call("orc_compute_timestep", it)

    def update_lhs_diagonal(self):
        self._call("orc_update_lhs_diagonal")

    def assemble_step1_rhs(self):
        self._call("orc_assemble_step1_rhs")

    def step1(self):
        self._call("orc_step1")

    def assemble_step2_rhs(self):
        self._call("orc_assemble_step2_rhs")

    def step2(self):
        self._call("orc_step2")

    def step3(self):
        self._call("orc_step3")

    def assemble_step4_rhs(self):
        self._call("orc_assemble_step4_rhs")

    def step4(self):
        self._call("orc_step4")

    def apply_velocity_package(self):
        self._call("orc_apply_velocity_package")

    def apply_velocity(self):
        self._call("orc_apply_velocity")

    def apply_temperature(self):
        self._call("orc_apply_temperature")

    def apply_pressure(self):
        self._call("orc_apply_pressure")

    def convergence(self):
        self._call("orc_convergence")

    def update_velocity_magnitude(self):
        self._lib.orc_update_velocity_magnitude(C.byref(self._s))

    def advance_one_step(self, it):
        self._call("orc_advance_one_step", it)

    def begin_step(self):
        self.a["unkn1"][:] = self.a["unkno"]
        self.a["pres1"][:] = self.a["pres"]
        self.a["temperature1"][:] = self.a["temperature"]

    def pressure_multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        rc = self._lib.orc_pressure_multiply(C.byref(self._s), x.ctypes.data_as(_PD), y.ctypes.data_as(_PD))
        if rc:
            raise OracleError(self._lib.orc_last_error().decode())
        return y

    def solve_pressure(self) -> CgResult:
        res = CgResult()
        rc = self._lib.orc_solve_pressure(C.byref(self._s), C.byref(res))
        if rc:
            raise OracleError(self._lib.orc_last_error().decode())
        return res
