"""Device snapshots (cbs_hip_snapshot_*) feeding the asynchronous writer
(cbs_post_write_solution_async) on the GPU.

The text and array formats are pinned against the reference on the CPU
(tests/test_post_output.py).  Here the device path runs a few steps with a
VTU output at every step written by the writer's thread from snapshots while
the next steps run; every written array must equal, bit for bit, the device
fields read synchronously after the same step (with the reference's masks),
and every CSV value must be the "%.16g" text of the step's scalars.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(name="gpost", transient=0):
    from cbs3d import CaseSetup, cases
    c = cases.cavity(5, re=100.0, jitter=0.1, seed=3, name=name)
    c.params.transient_on = transient
    return CaseSetup.from_case(c)


def test_async_snapshots_match_device_fields(tmp_path):
    from cbs3d import FP64, HipContext
    from cbs3d import post as cpost
    case = _case()
    ctx = HipContext(case)
    ctx.initialise()
    ctx.buildElementPressureTerms()
    w = cpost.Post(case, cpost.config(vtu_output_every_iterations=1), root=str(tmp_path), case_name="gpost")
    w.initialiseRunOutputs(ctx=ctx)
    want = {}
    rows = []
    for it in range(1, 6):
        ctx.advanceOneStep(it, precision=FP64)
        sc = ctx.scalars()
        w.writeResidualRow(it, sc)
        assert w.shouldWriteVTU(it, sc.rtime)
        w.writeSolution(it, sc.rtime, ctx=ctx)  # the next step starts before this file is written
        want[it] = {n: ctx.get(n) for n in ("pres", "temperature", "unkno", "velocity")}
        rows.append([str(it)] + ["%.16g" % v for v in (sc.rtime, sc.dtreal, *sc.hb)])
    w.flush()
    st = w.stats()
    w.close()
    out = os.path.join(str(tmp_path), "gpost")
    assert st["vtu_files"] == 6
    mask = None
    for it, f in want.items():
        v = cpost.read_vtu(os.path.join(out, f"gpost_step_{it:08d}.vtu"))
        pa = v["pressure_active"].astype(bool)
        va = v["velocity_active"].astype(bool)
        mask = pa
        assert np.array_equal(v["pressure"][pa], f["pres"][pa]) and np.all(v["pressure"][~pa] == 0)
        assert np.array_equal(v["temperature"], f["temperature"])
        u = f["unkno"].reshape(-1, 3)
        assert np.array_equal(v["velocity"].reshape(-1, 3)[va], u[va])
        assert np.array_equal(v["velocity_magnitude"][va], f["velocity"][va])
    assert mask is not None and mask.all()  # one fluid: every node carries pressure
    with open(os.path.join(out, "gpost_residuals.csv")) as fh:
        lines = fh.read().splitlines()[1:]
    assert len(lines) == len(rows)
    for line, want_row in zip(lines, rows):
        assert line.split(",")[:18] == want_row


def test_run_loop_writes_final_solution_once(tmp_path):
    from cbs3d import MIXED, HipContext
    from cbs3d import post as cpost
    case = _case("grun")
    ctx = HipContext(case)
    ctx.initialise()
    ctx.buildElementPressureTerms()
    w = cpost.Post(case, cpost.config(vtu_output_every_iterations=3), root=str(tmp_path), case_name="grun")
    last, reason = cpost.run(ctx, w, 7, precision=MIXED, vel_check=1, l2norm_vel_tolerance=1e-300)
    assert (last, reason) == (7, "maximum iteration count reached")
    names = sorted(os.listdir(os.path.join(str(tmp_path), "grun")))
    assert names == ["grun.pvd", "grun_residuals.csv"] + [f"grun_step_{k:08d}.vtu" for k in (0, 3, 6, 7)]
    v = cpost.read_vtu(os.path.join(str(tmp_path), "grun", "grun_step_00000007.vtu"))
    assert np.array_equal(v["pressure"], ctx.get("pres"))
    w.close()


def test_snapshot_slots_are_reused():
    """Two slots: a third snapshot waits for a release; released slots come
    back with the current fields."""
    import ctypes as C

    from cbs3d import HipContext
    from cbs3d._native import FIELD_ID, hip_lib
    case = _case("gslot")
    ctx = HipContext(case)
    ctx.initialise()
    lib = hip_lib()
    snaps = []
    for _ in range(2):
        s = C.c_void_p()
        ctx._check(lib.cbs_hip_snapshot_begin(ctx._h, C.byref(s)))
        snaps.append(s)
    assert snaps[0].value != snaps[1].value
    n = ctx.npoin
    for s in snaps:
        assert lib.cbs_hip_snapshot_wait(s) == 0
        p = lib.cbs_hip_snapshot_field(s, FIELD_ID["pres"])
        assert np.array_equal(np.ctypeslib.as_array(p, (n,)), ctx.get("pres"))
    assert lib.cbs_hip_snapshot_release(snaps[0]) == 0
    ctx.set("pres", np.arange(n, dtype=np.float64))
    s = C.c_void_p()
    ctx._check(lib.cbs_hip_snapshot_begin(ctx._h, C.byref(s)))
    assert s.value == snaps[0].value
    assert lib.cbs_hip_snapshot_wait(s) == 0
    p = lib.cbs_hip_snapshot_field(s, FIELD_ID["pres"])
    assert np.array_equal(np.ctypeslib.as_array(p, (n,)), np.arange(n, dtype=np.float64))
    assert lib.cbs_hip_snapshot_release(s) == 0
    assert lib.cbs_hip_snapshot_release(snaps[1]) == 0
    assert lib.cbs_hip_snapshot_release(snaps[1]) != 0  # not held
