"""Solution output of a CBS run, off the time step's critical path.

Mirrors the file output of the reference's Post (src/io/Post.cpp,
include/cbs/io/Post.hpp) and the output part of Solver::run
(src/solver/Solver.cpp:137-224) over include/cbs_post.h:

    post = Post(case, PostConfig(...), root="output", case_name="cav")
    post.initialiseRunOutputs(ctx=ctx)            # Post.cpp:731-754
    post.writeResidualRow(it, ctx.scalars())      # Post.cpp:768-812
    if post.shouldWriteVTU(it, rtime):            # Post.cpp:821-840
        post.writeSolution(it, rtime, ctx=ctx)    # Post.cpp:842-860 (asynchronous)
    post.flush()

Residual CSV and PVD files carry the reference's text.  VTU files carry the
reference's arrays (names, types, order, masks, values) in one appended
raw-binary block; read_vtu() decodes them.  writeSolution(ctx=...) takes a
device snapshot (cbs_hip_snapshot_*) and lets the writer's thread download,
encode and write it while the next steps run.
"""

from __future__ import annotations

import ctypes as C
import re
from typing import Dict, Optional, Tuple

import numpy as np

from ._native import (CBS_OK, FIELD_ID, STATUS_NAMES, PostConfig, PostFields, PostMesh, Scalars, hip_lib,
                      host_lib)

__all__ = ["Post", "PostConfig", "PostError", "read_vtu", "run"]

_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int32)


class PostError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def _check(rc: int):
    if rc != CBS_OK:
        raise PostError(rc, host_lib().cbs_post_last_error().decode())


def config(residual_log_enabled=1, residual_log_every=1, vtu_output_enabled=1, vtu_output_every_iterations=100,
           vtu_output_every_sim_time=0.0, transient_on=0, write_boundary_debug_arrays=1) -> PostConfig:
    """Output controls; the defaults are RunConfig's (include/cbs/core/RunConfig.hpp:163-172)."""
    return PostConfig(residual_log_enabled, residual_log_every, vtu_output_enabled, vtu_output_every_iterations,
                      vtu_output_every_sim_time, transient_on, write_boundary_debug_arrays)


class Post:
    """One run's output writer.

    mesh is a CaseSetup (its arrays are borrowed; keep it alive) or a dict with
    coord (N, 3), intma (E, 4) 1-based, iside (B, 6) and mat_elem (E)."""

    def __init__(self, mesh, cfg: Optional[PostConfig] = None, root: str = "output", case_name: str = "case"):
        self.cfg = cfg if cfg is not None else config()
        self._keep = [mesh]
        if isinstance(mesh, dict):
            coord = np.ascontiguousarray(mesh["coord"], np.float64).reshape(-1)
            intma = np.ascontiguousarray(mesh["intma"], np.int32).reshape(-1)
            iside = np.ascontiguousarray(mesh["iside"], np.int32).reshape(-1)
            mat = np.ascontiguousarray(mesh["mat_elem"], np.int32).reshape(-1)
            self._keep += [coord, intma, iside, mat]
            m = PostMesh(coord.size // 3, mat.size, iside.size // 6, coord.ctypes.data_as(_PD),
                         intma.ctypes.data_as(_PI), iside.ctypes.data_as(_PI), mat.ctypes.data_as(_PI))
        else:
            d = mesh.mesh_desc()
            m = PostMesh(d.npoin, d.nelem, d.nboun, d.coord, d.intma, d.iside, d.mat_elem)
        self.npoin = m.npoin
        self._h = C.c_void_p()
        _check(host_lib().cbs_post_create(C.byref(m), C.byref(self.cfg), root.encode(), case_name.encode(),
                                          C.byref(self._h)))

    # -- the reference's entry points ------------------------------------
    def initialiseRunOutputs(self, rtime: float = 0.0, fields: Optional[Dict[str, np.ndarray]] = None, ctx=None):
        """Post::initialiseRunOutputs: directory, CSV header, iteration-0 VTU (written before returning)."""
        if ctx is not None:
            fields = _ctx_fields(ctx)
            rtime = ctx.scalars().rtime
        f, keep = _fields(fields, self.npoin) if fields is not None else (None, None)
        _check(host_lib().cbs_post_initialise_run_outputs(self._h, rtime, C.byref(f) if f is not None else None))

    def writeResidualRow(self, iitime: int, scalars: Scalars):
        _check(host_lib().cbs_post_write_residual_row(self._h, iitime, C.byref(scalars)))

    def shouldWriteVTU(self, iitime: int, rtime: float) -> bool:
        due = C.c_int32()
        _check(host_lib().cbs_post_should_write_vtu(self._h, iitime, rtime, C.byref(due)))
        return bool(due.value)

    def writeSolution(self, iitime: int, rtime: float, fields: Optional[Dict[str, np.ndarray]] = None, ctx=None):
        """Post::writeSolution.  With ctx: a device snapshot, written by the
        writer's thread (returns once the copies are queued).  With fields
        (host arrays): written before returning."""
        if ctx is not None:
            hip = hip_lib()
            snap = C.c_void_p()
            ctx._check(hip.cbs_hip_snapshot_begin(ctx._h, C.byref(snap)))
            f = PostFields(*(hip.cbs_hip_snapshot_field(snap, FIELD_ID[n])
                             for n in ("pres", "temperature", "unkno", "velocity")))
            _check(host_lib().cbs_post_write_solution_async(
                self._h, iitime, rtime, C.byref(f), C.cast(hip.cbs_hip_snapshot_wait, C.c_void_p),
                C.cast(hip.cbs_hip_snapshot_release, C.c_void_p), snap))
            return
        f, keep = _fields(fields, self.npoin)
        _check(host_lib().cbs_post_write_solution(self._h, iitime, rtime, C.byref(f)))

    def last_output(self) -> int:
        it = C.c_int32()
        _check(host_lib().cbs_post_last_output(self._h, C.byref(it)))
        return it.value

    def flush(self):
        """Waits for the pending writes; raises the first background failure."""
        _check(host_lib().cbs_post_flush(self._h))

    def stats(self) -> dict:
        n, b, s = C.c_int64(), C.c_int64(), C.c_double()
        _check(host_lib().cbs_post_stats(self._h, C.byref(n), C.byref(b), C.byref(s)))
        return {"vtu_files": n.value, "vtu_bytes": b.value, "write_seconds": s.value}

    def close(self):
        if getattr(self, "_h", None):
            host_lib().cbs_post_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _ctx_fields(ctx) -> Dict[str, np.ndarray]:
    return {n: ctx.get(n) for n in ("pres", "temperature", "unkno", "velocity")}


def _fields(fields: Dict[str, np.ndarray], npoin: int) -> Tuple[PostFields, list]:
    arrs = []
    for name, n in (("pres", npoin), ("temperature", npoin), ("unkno", 3 * npoin), ("velocity", npoin)):
        a = np.ascontiguousarray(fields[name], np.float64).reshape(-1)
        if a.size != n:
            raise ValueError(f"{name}: {a.size} values, expected {n}")
        arrs.append(a)
    return PostFields(*(a.ctypes.data_as(_PD) for a in arrs)), arrs


_DTYPES = {"Float64": np.float64, "Int32": np.int32, "UInt8": np.uint8}
_ARRAY = re.compile(r'<DataArray type="(\w+)"(?: Name="(\w+)")? NumberOfComponents="(\d+)" '
                    r'format="appended" offset="(\d+)"/>')


def read_vtu(path: str) -> Dict[str, np.ndarray]:
    """Arrays of a VTU file written by Post ("points" for the coordinates),
    each flat in file order."""
    with open(path, "rb") as f:
        data = f.read()
    start = data.index(b'<AppendedData encoding="raw">')
    base = data.index(b"_", start) + 1
    out = {}
    for typ, name, ncomp, off in _ARRAY.findall(data[:start].decode()):
        at = base + int(off)
        nbytes = int(np.frombuffer(data, np.uint64, 1, at)[0])
        dt = np.dtype(_DTYPES[typ])
        out[name or "points"] = np.frombuffer(data, dt, nbytes // dt.itemsize, at + 8).copy()
    return out


def run(ctx, post: Post, ntime: int, precision: int = 0, steady_min_iterations: int = 5, vel_check: int = 1,
        l2norm_vel_tolerance: float = 1.0e-6, temp_check: int = 0, l2norm_temp_tolerance: float = 1.0e-6,
        l2norm_pres_tolerance: float = 0.0, end_rtime: float = 0.0, **solve_kw) -> Tuple[int, str]:
    """The CBS loop of Solver::run (Solver.cpp:137-224) without the console
    output: steps, residual rows, VTU outputs (asynchronous), the stopping
    rules (Solver.cpp:191-203, 708-726; Convergence.cpp:399-418) and the
    final solution written exactly once.  Returns (last iteration, reason)."""
    cfg = ctx.cfg
    post.initialiseRunOutputs(ctx=ctx)
    last, reason = 0, "maximum iteration count reached"
    for it in range(1, ntime + 1):
        last = it
        ctx.advanceOneStep(it, precision=precision, **solve_kw)
        sc = ctx.scalars()
        post.writeResidualRow(it, sc)
        if post.shouldWriteVTU(it, sc.rtime):
            post.writeSolution(it, sc.rtime, ctx=ctx)
        if cfg.transient_on > 0 and end_rtime > 0.0 and sc.rtime >= end_rtime:
            reason = "transient end time reached"
            break
        if cfg.transient_on < 1 and it >= steady_min_iterations:
            hb = list(sc.hb)
            vres = max(hb[0], hb[3], hb[6])
            vel_ok = vel_check < 1 or vres < l2norm_vel_tolerance
            temp_ok = temp_check < 1 or cfg.temp_calc < 1 or hb[12] < l2norm_temp_tolerance
            pres_ok = l2norm_pres_tolerance <= 0.0 or hb[9] < l2norm_pres_tolerance
            if vel_ok and temp_ok and pres_ok:
                reason = "steady-state convergence reached"
                break
    if post.cfg.vtu_output_enabled > 0 and post.last_output() != last:
        post.writeSolution(last, ctx.scalars().rtime, ctx=ctx)
    post.flush()
    return last, reason
