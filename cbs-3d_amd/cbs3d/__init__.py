"""MI355X-native CBS time-step hot path of CBS3D++_SI.

Python mirror of the reference's solver interface over the C ABI in
include/cbs_hip.h (device) and include/cbs_case.h (host set-up):

    case  = CaseSetup.from_case(cases.workload("cavity20"))   # MeshIO + Preprocess + Coloring
    ctx   = HipContext(case)                                  # device state
    ctx.initialise()                                          # initial BCs (Solver::initialise)
    ctx.buildElementPressureTerms()                           # Solver::preparePressureSystem
    for it in range(1, ntime + 1):
        result = ctx.advanceOneStep(it)                       # Solver::advanceOneStep

The methods keep the reference's names (TimeStep::computeTimeStep ->
computeTimeStep, Steps::step1 -> step1, ConjugateGradient::solvePressure ->
solvePressure, ...).  Failures raise CbsError carrying the status code and
the reference's message class.  There is no CPU fallback: without the built
HIP library every device call raises.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import cases
from ._native import (CBS_OK, FIELD_ID, FIELDS, STATUS_NAMES, AdjointOptions, CgOptions, CgResult, Config,
                      Info, Material, MeshDesc, OperatorInfo, Scalars, hip_lib, host_lib)

__all__ = ["CaseSetup", "HipContext", "BatchContext", "CbsError", "cases", "stability_dt", "FP64", "MIXED", "AMG"]

FP64 = 0
MIXED = 1
AMG = 2  # FP64 CG preconditioned by an FP32 smoothed-aggregation multigrid V-cycle (build-only)

_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int32)


class CbsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


def config_from_params(p: cases.CaseParams) -> Config:
    """The .par controls of a generated case as a cbs_config (RunConfig defaults elsewhere)."""
    c = Config()
    c.solver_opt = 1
    c.cbs_scheme = 1
    c.ntime = p.ntime
    c.transient_on = p.transient_on
    c.temp_calc = p.temp_calc
    c.ilots = p.ilots
    c.htype = p.htype
    c.dtfixed = p.dtfixed
    c.dtfix = p.dtfix
    c.beta_opt = p.beta_opt
    c.rem_deltp = p.rem_deltp
    c.dtfix_end = p.dtfix_end
    c.csafm = p.csafm
    c.deltr = p.deltr
    c.csafm2 = 2.0
    c.epsilon1 = 1.0e-6
    c.re = p.re
    c.pr = p.pr
    c.pnode = p.pnode
    c.relToler = p.rel_toler
    c.absToler = p.abs_toler
    c.cg_preconditioner = p.cg_preconditioner
    c.cg_conv_test = p.cg_conv_test
    c.cg_max_iter = p.cg_max_iter
    c.dimensional_mode = p.dimensional_mode
    c.material_properties_enabled = p.material_properties_enabled
    c.inlet_temperature = p.inlet_temperature
    c.inlet_density = p.inlet_density
    c.mass_flow_inlet_enabled = p.mass_flow_inlet_enabled
    c.inlet_mass_flow_rate = p.inlet_mass_flow_rate
    c.outlet_pressure_gauge = p.outlet_pressure_gauge
    c.inlet_u, c.inlet_v, c.inlet_w = p.inlet_u, p.inlet_v, p.inlet_w
    c.heat_flux_bc = p.heat_flux_bc
    for i, v in enumerate(p.cinf):
        c.cinf[i] = v
    return c


class CaseSetup:
    """Host-side CBSStateSI before the first step (include/cbs_case.h)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def from_case(cls, case: cases.Case, setup: bool = True, device_coloring: bool = False) -> "CaseSetup":
        lib = host_lib()
        cfg = config_from_params(case.params)
        coord = np.ascontiguousarray(case.coord, np.float64)
        intma = np.ascontiguousarray(case.intma, np.int32)
        bface = np.ascontiguousarray(case.bface, np.int32).copy()
        bface[:, 4] = case.solver_bc()
        mat_ptr, mats_ptr, nmat = None, None, 0
        keep = []
        if case.params.material_properties_enabled:
            mat = np.ascontiguousarray(getattr(case, "mat", np.zeros(case.nelem, np.int32)), np.int32)
            rows = getattr(case, "materials", None)
            if rows is None:
                rho, cp, k, mu, q = case.material
                rows = [(0, "fluid", "fluid", rho, cp, k, mu, q)]
            arr = (Material * len(rows))()
            for i, r in enumerate(rows):
                arr[i].id, arr[i].rho, arr[i].cp, arr[i].k, arr[i].mu, arr[i].qvol = (
                    r[0], r[3], r[4], r[5], r[6], r[7])
            keep += [mat, arr]
            mat_ptr, mats_ptr, nmat = mat.ctypes.data_as(_PI), arr, len(rows)
        h = C.c_void_p()
        rc = lib.cbs_case_create(C.byref(cfg), case.nelem, case.npoin, case.nboun,
                                 coord.ctypes.data_as(_PD), intma.ctypes.data_as(_PI),
                                 bface.ctypes.data_as(_PI), mat_ptr, nmat, mats_ptr, C.byref(h))
        if rc != CBS_OK:
            raise CbsError(rc, lib.cbs_case_last_error().decode())
        obj = cls(h)
        if setup:
            obj.setup(device_coloring=device_coloring)
        return obj

    @classmethod
    def read(cls, base: str, setup: bool = True) -> "CaseSetup":
        lib = host_lib()
        h = C.c_void_p()
        rc = lib.cbs_case_read(base.encode(), C.byref(h))
        if rc != CBS_OK:
            raise CbsError(rc, lib.cbs_case_last_error().decode())
        obj = cls(h)
        if setup:
            obj.setup()
        return obj

    def setup(self, device_coloring: bool = False):
        """Solver::initialise's set-up.  device_coloring: every phase on the
        host except Coloring::build, which runs on the GPU (cbs_hip_coloring,
        the same colouring bit for bit)."""
        lib = host_lib()
        if not device_coloring:
            rc = lib.cbs_case_setup(self._h)
            if rc != CBS_OK:
                raise CbsError(rc, lib.cbs_case_last_error().decode())
            return
        rc = lib.cbs_case_setup_without_coloring(self._h)
        if rc != CBS_OK:
            raise CbsError(rc, lib.cbs_case_last_error().decode())
        ncolor, ptr, elem, self.coloring_levels = device_coloring_of(self.array("intma"), self.mesh_desc().npoin)
        rc = lib.cbs_case_set_coloring(self._h, ncolor, ptr.ctypes.data_as(_PI), elem.ctypes.data_as(_PI))
        if rc != CBS_OK:
            raise CbsError(rc, lib.cbs_case_last_error().decode())

    def config(self) -> Config:
        c = Config()
        host_lib().cbs_case_config(self._h, C.byref(c))
        return c

    def mesh_desc(self) -> MeshDesc:
        d = MeshDesc()
        host_lib().cbs_case_mesh_desc(self._h, C.byref(d))
        return d

    def array(self, name: str) -> np.ndarray:
        data = C.c_void_p()
        cnt = C.c_int64()
        dt = C.c_char()
        rc = host_lib().cbs_case_array(self._h, name.encode(), C.byref(data), C.byref(cnt), C.byref(dt))
        if rc != CBS_OK:
            raise CbsError(rc, host_lib().cbs_case_last_error().decode())
        ctype = C.c_double if dt.value == b"d" else C.c_int32
        if cnt.value == 0:
            return np.zeros(0, np.float64 if dt.value == b"d" else np.int32)
        buf = (ctype * cnt.value).from_address(data.value)
        return np.ctypeslib.as_array(buf).copy()

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                host_lib().cbs_case_destroy(self._h)
            except Exception:  # interpreter shutdown: the module globals may be gone
                pass
            self._h = None


def _opts(precision: int = FP64, inner_rel_tol: float = 0.0, max_refinements: int = 0):
    o = CgOptions()
    o.precision = precision
    o.inner_rel_tol = inner_rel_tol
    o.max_refinements = max_refinements
    return o


class HipContext:
    """Device state of one case: the reference's per-step functions on the GPU."""

    def __init__(self, case: CaseSetup):
        self.lib = hip_lib()
        self.case = case  # keeps the host arrays alive while the context is created
        self.cfg = case.config()
        self.mesh = case.mesh_desc()
        self._h = C.c_void_p()
        rc = self.lib.cbs_hip_create(C.byref(self.cfg), C.byref(self.mesh), C.byref(self._h))
        if rc != CBS_OK:
            raise CbsError(rc, self.lib.cbs_hip_last_error(None).decode())
        self.npoin = self.mesh.npoin
        self.nelem = self.mesh.nelem
        info = Info()
        self._check(self.lib.cbs_hip_info(self._h, C.byref(info)))
        self.nfluid = info.nfluid
        self.ncolor = info.ncolor
        self.device_bytes = info.device_bytes
        self.set("unkno", case.array("unkno"))
        self.set("pres", case.array("pres"))
        self.set("temperature", case.array("temperature"))

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int):
        if rc != CBS_OK:
            raise CbsError(rc, self.lib.cbs_hip_last_error(self._h).decode())

    def set(self, name: str, values):
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        self._check(self.lib.cbs_hip_set_field(self._h, FIELD_ID[name], v.ctypes.data_as(_PD), v.size))

    def get(self, name: str) -> np.ndarray:
        n = self.lib.cbs_hip_field_size(self._h, FIELD_ID[name])
        out = np.empty(n, np.float64)
        self._check(self.lib.cbs_hip_get_field(self._h, FIELD_ID[name], out.ctypes.data_as(_PD), n))
        return out

    def scalars(self) -> Scalars:
        s = Scalars()
        self._check(self.lib.cbs_hip_get_scalars(self._h, C.byref(s)))
        return s

    CREATE_PHASES = ("validation", "connectivity", "colour_slots", "tiles", "incidences", "boundary_programmes",
                     "uploads", "assembled_operator", "node_slots", "node_vectors")

    def create_seconds(self) -> dict:
        """Wall-clock seconds of the phases of this context's construction."""
        buf = (C.c_double * 16)()
        n = self.lib.cbs_hip_create_seconds(self._h, buf, 16)
        return {name: round(buf[k], 4) for k, name in enumerate(self.CREATE_PHASES[:max(n, 0)])}

    def bandwidth_probe(self, nbytes: int = 1 << 30, reps: int = 20) -> float:
        """GB/s of a plain device copy (read + write) on the context's stream."""
        g = C.c_double()
        self._check(self.lib.cbs_hip_bandwidth_probe(self._h, nbytes, reps, C.byref(g)))
        return g.value

    def synchronize(self):
        self._check(self.lib.cbs_hip_synchronize(self._h))

    def set_profile(self, on: bool):
        """Device timing with HIP events inside pressure solves: the apply
        timer (apply_timer) and the phase timers of the CgResult."""
        self._check(self.lib.cbs_hip_set_profiling(self._h, 1 if on else 0))

    def apply_timer(self, precision: int = FP64, reset: bool = False):
        """(milliseconds, sweeps) accumulated for one precision's applies."""
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.lib.cbs_hip_apply_timer(self._h, precision, C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def reset_apply_timer(self):
        self.apply_timer(reset=True)

    def operator_info(self) -> dict:
        """Shape of the assembled operator of the PCG apply (row tiles, staged
        nodes, off-diagonal nonzeros, stored entries, contributions; the
        multi-RHS product's tiles once a multi-RHS solve has run)."""
        o = OperatorInfo()
        self._check(self.lib.cbs_hip_operator_info(self._h, C.byref(o)))
        return {n: getattr(o, n) for n, _ in o._fields_}

    def apply_assembled(self, x, precision: int = FP64) -> np.ndarray:
        """The PCG's apply on the assembled operator: free ? A x : 0."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._check(self.lib.cbs_hip_apply_assembled(self._h, precision, x.ctypes.data_as(_PD),
                                                     y.ctypes.data_as(_PD)))
        return y

    def set_apply_config(self, precision: int, threads: int, batch: int):
        """Fix the PCG apply's launch shape (threads per row tile, column pairs per load batch)."""
        self._check(self.lib.cbs_hip_set_apply_config(self._h, precision, threads, batch))

    def direction_fusion(self, precision: int = FP64):
        """(fused: 1 in the apply, 0 its own kernel, -1 undecided; timed ms separate; timed ms fused) of
        how the PCG iteration of one solver form forms its direction."""
        f = C.c_int32()
        a, b = C.c_float(), C.c_float()
        self._check(self.lib.cbs_hip_direction_fusion(self._h, precision, C.byref(f), C.byref(a), C.byref(b)))
        return f.value, round(a.value, 5), round(b.value, 5)

    def set_direction_fusion(self, precision: int, mode: int):
        """Pin the direction update in the apply (1) or in its own kernel (0), or back to timed (-1)."""
        self._check(self.lib.cbs_hip_set_direction_fusion(self._h, precision, mode))

    def amg_info(self) -> dict:
        """Multigrid hierarchy of the AMG form (after its first solve): levels,
        set-up seconds, stored nonzeros of the coarse operators, P and R."""
        secs = C.c_double()
        lev = C.c_int32()
        nnz = C.c_int64()
        self._check(self.lib.cbs_hip_amg_info(self._h, C.byref(secs), C.byref(lev), C.byref(nnz)))
        return {"levels": lev.value, "setup_seconds": secs.value, "stored_nonzeros": nnz.value}

    def set_graphs(self, on: bool):
        """PCG iterations as replays of a captured HIP graph block (default) or per-kernel launches."""
        self._check(self.lib.cbs_hip_set_graphs(self._h, 1 if on else 0))

    def graphs_active(self):
        """(FP64, mixed): whether each solver form replays a captured block."""
        a, b = C.c_int32(), C.c_int32()
        self._check(self.lib.cbs_hip_graphs_active(self._h, C.byref(a), C.byref(b)))
        return bool(a.value), bool(b.value)

    def apply_config(self, precision: int = FP64):
        """(threads per row tile, column pairs per load batch, candidate ms) of the tuned PCG apply."""
        t = C.c_int32()
        p = C.c_int32()
        ms = (C.c_float * 6)()
        self._check(self.lib.cbs_hip_apply_config(self._h, precision, C.byref(t), C.byref(p), ms))
        return t.value, p.value, [round(float(v), 4) for v in ms]

    def set_dtreal(self, dt: float):
        self._check(self.lib.cbs_hip_set_dtreal(self._h, dt))

    # -- Solver::initialise / preparePressureSystem -----------------------
    def initialise(self):
        """Initial boundary values of Solver::initialise (Solver.cpp:296-300)."""
        self.applyTemperature()
        self.applyVelocity()
        self.applyPressure()
        self.updateVelocityMagnitude()

    def buildElementPressureTerms(self):
        self._check(self.lib.cbs_hip_build_element_pressure_terms(self._h))

    # -- the reference's per-step functions --------------------------------
    def computeTimeStep(self, iitime: int):
        self._check(self.lib.cbs_hip_compute_timestep(self._h, iitime))

    def updateLhsDiagonal(self):
        self._check(self.lib.cbs_hip_update_lhs_diagonal(self._h))

    def assembleStep1Rhs(self):
        self._check(self.lib.cbs_hip_assemble_step1_rhs(self._h))

    def step1(self):
        self._check(self.lib.cbs_hip_step1(self._h))

    def assembleStep2Rhs(self):
        self._check(self.lib.cbs_hip_assemble_step2_rhs(self._h))

    def pressureMultiply(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._check(self.lib.cbs_hip_pressure_multiply(self._h, x.ctypes.data_as(_PD), y.ctypes.data_as(_PD)))
        return y

    def solvePressure(self, precision: int = FP64, **kw) -> CgResult:
        r = CgResult()
        self._check(self.lib.cbs_hip_solve_pressure(self._h, C.byref(_opts(precision, **kw)), C.byref(r)))
        return r

    def solvePressureMulti(self, B, X=None, precision: int = FP64, **kw):
        """Solve A x_j = b_j for the columns of B (shape (nrhs, npoin)); returns (X, results)."""
        B = np.ascontiguousarray(B, np.float64)
        nrhs = B.shape[0]
        X = np.zeros_like(B) if X is None else np.ascontiguousarray(X, np.float64).copy()
        res = (CgResult * nrhs)()
        self._check(self.lib.cbs_hip_solve_pressure_multi(self._h, nrhs, B.ctypes.data_as(_PD), X.ctypes.data_as(_PD),
                                                          C.byref(_opts(precision, **kw)), res))
        return X, list(res)

    def step2(self, precision: int = FP64, **kw) -> CgResult:
        r = CgResult()
        self._check(self.lib.cbs_hip_step2(self._h, C.byref(_opts(precision, **kw)), C.byref(r)))
        return r

    def step3(self):
        self._check(self.lib.cbs_hip_step3(self._h))

    def assembleStep4Rhs(self):
        self._check(self.lib.cbs_hip_assemble_step4_rhs(self._h))

    def step4(self):
        self._check(self.lib.cbs_hip_step4(self._h))

    def applyVelocityPackage(self):
        self._check(self.lib.cbs_hip_apply_boundary(self._h, 0))

    def applyVelocity(self):
        self._check(self.lib.cbs_hip_apply_boundary(self._h, 1))

    def applyTemperature(self):
        self._check(self.lib.cbs_hip_apply_boundary(self._h, 2))

    def applyPressure(self):
        self._check(self.lib.cbs_hip_apply_boundary(self._h, 3))

    def updateVelocityMagnitude(self):
        self._check(self.lib.cbs_hip_update_velocity_magnitude(self._h))

    def evaluateConvergence(self):
        self._check(self.lib.cbs_hip_convergence(self._h))

    def advanceOneStep(self, iitime: int, precision: int = FP64, **kw) -> CgResult:
        r = CgResult()
        self._check(self.lib.cbs_hip_advance_one_step(self._h, iitime, C.byref(_opts(precision, **kw)),
                                                      C.byref(r)))
        return r

    # -- adjoint workload (build-only forward-model extensions) -------------
    def set_body_force(self, f=None):
        """Nodal body force (N, 3), an acceleration added to the Step 1 predictor; None switches it off."""
        if f is None:
            self._check(self.lib.cbs_hip_set_body_force(self._h, None))
            return
        self._force = np.ascontiguousarray(f, np.float64).reshape(-1)
        assert self._force.size == 3 * self.npoin
        self._check(self.lib.cbs_hip_set_body_force(self._h, self._force.ctypes.data))

    def set_inflow_coefficients(self, c=None):
        """16 Fourier coefficients added to the BC-508 inflow profile; None restores the profile."""
        if c is None:
            self._check(self.lib.cbs_hip_set_inflow_coefficients(self._h, None, 0))
            return
        cc = np.ascontiguousarray(c, np.float64).reshape(-1)
        self._check(self.lib.cbs_hip_set_inflow_coefficients(self._h, cc.ctypes.data, cc.size))

    def inflow_basis(self):
        """(1-based nodes, phi (nodes, 16)) of the inflow controls."""
        n = C.c_int32()
        self._check(self.lib.cbs_hip_inflow_basis(self._h, C.byref(n), None, None))
        nodes = np.zeros(n.value, np.int32)
        phi = np.zeros((n.value, 16))
        self._check(self.lib.cbs_hip_inflow_basis(self._h, C.byref(n), nodes.ctypes.data, phi.ctypes.data))
        return nodes, phi

    def adjoint_forward(self, nsteps: int, objective_steps: int, checkpoint_every: int = 10,
                        rel_tol: float = 1e-12) -> float:
        """Initial BCs + nsteps CBS steps from the current fields; returns the objective Phi."""
        o = AdjointOptions(nsteps, objective_steps, checkpoint_every, rel_tol)
        phi = C.c_double()
        self._check(self.lib.cbs_hip_adjoint_forward(self._h, C.byref(o), C.byref(phi)))
        return phi.value

    def adjoint_backward(self, seed: float = 1.0):
        """seed * (dPhi/df (N, 3), dPhi/dc (16,)) of the last adjoint_forward."""
        gf = np.zeros(3 * self.npoin)
        gc = np.zeros(16)
        self._check(self.lib.cbs_hip_adjoint_backward(self._h, seed, gf.ctypes.data, gc.ctypes.data))
        return gf.reshape(self.npoin, 3), gc

    def close(self):
        if getattr(self, "_h", None):
            self.lib.cbs_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def device_coloring_of(intma: np.ndarray, npoin: int):
    """Coloring::build on the GPU (cbs_hip_coloring): (ncolor, color_ptr,
    color_elem, levels) for a (4, E) 1-based connectivity."""
    lib = hip_lib()
    intma = np.ascontiguousarray(intma, np.int32).reshape(-1)
    nelem = intma.size // 4
    ptr = np.zeros(257, np.int32)
    elem = np.zeros(nelem, np.int32)
    nc, lv = C.c_int32(), C.c_int32()
    rc = lib.cbs_hip_coloring(nelem, npoin, intma.ctypes.data_as(_PI), C.byref(nc), ptr.ctypes.data_as(_PI),
                              elem.ctypes.data_as(_PI), C.byref(lv))
    if rc != CBS_OK:
        raise CbsError(rc, lib.cbs_hip_last_error(None).decode())
    return nc.value, ptr[:nc.value + 1].copy(), elem, lv.value


def stability_dt(case, iitime: int = 1) -> float:
    """The reference's stability time step (TimeStep::computeTimeStep) of a
    case's first step, computed on the device from its initial fields."""
    setup = CaseSetup.from_case(case)
    ctx = HipContext(setup)
    try:
        ctx.initialise()
        ctx.computeTimeStep(iitime)
        return float(ctx.scalars().dtreal)
    finally:
        ctx.close()


class BatchContext(HipContext):
    """Device state of several independent cases advanced together.

    One context over the disjoint union of the members (cbs_hip_create_batch):
    every per-step launch covers all members, while the time step, norms and
    pressure-solve scalars stay per member, so member k evolves exactly as
    HipContext(cases[k]) would.  Fields are the union arrays; member(k, name)
    slices one member's part.  The reference runs one case per process
    (Solver::run, src/solver/Solver.cpp:137-225).
    """

    def __init__(self, cases_: list):
        self.lib = hip_lib()
        self.cases = list(cases_)  # keeps the host arrays alive while the context is created
        n = len(self.cases)
        self.cfgs = (Config * n)(*[c.config() for c in self.cases])
        self.meshes = (MeshDesc * n)(*[c.mesh_desc() for c in self.cases])
        self._h = C.c_void_p()
        rc = self.lib.cbs_hip_create_batch(n, self.cfgs, self.meshes, C.byref(self._h))
        if rc != CBS_OK:
            raise CbsError(rc, self.lib.cbs_hip_last_error(None).decode())
        self.members = n
        self.node_range, self.elem_range = [], []
        for k in range(n):
            no, npn, eo, ne = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            self._check(self.lib.cbs_hip_member_range(self._h, k, C.byref(no), C.byref(npn), C.byref(eo),
                                                      C.byref(ne)))
            self.node_range.append((no.value, no.value + npn.value))
            self.elem_range.append((eo.value, eo.value + ne.value))
        info = Info()
        self._check(self.lib.cbs_hip_info(self._h, C.byref(info)))
        self.npoin = info.npoin
        self.nelem = info.nelem
        self.nfluid = info.nfluid
        self.ncolor = info.ncolor
        self.device_bytes = info.device_bytes
        for name in ("unkno", "pres", "temperature"):
            self.set(name, np.concatenate([np.asarray(c.array(name), np.float64).reshape(-1) for c in self.cases]))

    def member(self, k: int, name: str) -> np.ndarray:
        """Member k's part of a union field (node fields per node, element fields per element)."""
        full = self.get(name)
        if name in ("delte", "gstif", "gstifE", "pdiagE"):
            a, b = self.elem_range[k]
            w = full.size // self.nelem
        else:
            a, b = self.node_range[k]
            w = full.size // self.npoin
        return full[a * w:b * w].copy()

    def member_scalars(self, k: int) -> Scalars:
        s = Scalars()
        self._check(self.lib.cbs_hip_get_member_scalars(self._h, k, C.byref(s)))
        return s

    def member_cg_result(self, k: int) -> CgResult:
        r = CgResult()
        self._check(self.lib.cbs_hip_get_member_cg_result(self._h, k, C.byref(r)))
        return r
