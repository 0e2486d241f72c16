"""The ConjugateGradient::Result fields of cbs_cg_result on the device.

With profiling on, every solver form fills the phase timers with device
time from HIP events: matvec, dot, vector update and constraint time are
positive; the Jacobi forms fuse the preconditioner into the update
(preconditioner_seconds 0), the multigrid form spends it in its V-cycles;
setup ends before the solve does, and set-up and phases add up to the
solve's wall time.  With profiling off
only total and setup time are reported.
"""

import pytest

from cbs3d import AMG, FP64, MIXED, CaseSetup, HipContext, cases

pytestmark = pytest.mark.gpu

PHASES = ("matvec_seconds", "dot_seconds", "preconditioner_seconds", "vector_update_seconds",
          "constraint_seconds")


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    setup = CaseSetup.from_case(cases.cavity(24, re=400.0, jitter=0.1, seed=2, name="cav24"))
    c = HipContext(setup)
    c.initialise()
    c.buildElementPressureTerms()
    for a, b in (("unkn1", "unkno"), ("pres1", "pres"), ("temperature1", "temperature")):
        c.set(a, c.get(b))
    c.computeTimeStep(1)
    c.updateLhsDiagonal()
    c.step1()
    c.assembleStep2Rhs()
    c.rhs0, c.pres0 = c.get("rhs1"), c.get("pres")
    yield c
    c.close()


@pytest.mark.parametrize("form", [FP64, MIXED, AMG])
def test_phase_timers_with_profiling(ctx, form):
    for graphs in (True, False):
        ctx.set_graphs(graphs)
        ctx.set_profile(True)
        ctx.set("rhs1", ctx.rhs0)
        ctx.set("pres", ctx.pres0)
        r = ctx.solvePressure(form)
        ctx.set_profile(False)
        assert r.converged
        t = {k: getattr(r, k) for k in PHASES}
        log(f"[cg result] form {form} graphs {graphs}: {r.iterations} iterations, total {r.total_seconds:.4f} s, "
            f"setup {r.setup_seconds:.4f} s, " + ", ".join(f"{k} {v:.4f}" for k, v in t.items()))
        for k in ("matvec_seconds", "dot_seconds", "vector_update_seconds", "constraint_seconds"):
            assert t[k] > 0.0, k
        if form == AMG:
            assert t["preconditioner_seconds"] > 0.0
        else:
            assert t["preconditioner_seconds"] == 0.0
        assert 0.0 < r.setup_seconds < r.total_seconds
        assert r.setup_seconds + sum(t.values()) <= r.total_seconds * (1 + 1e-9)
    ctx.set_graphs(True)


def test_phase_timers_off_without_profiling(ctx):
    ctx.set("rhs1", ctx.rhs0)
    ctx.set("pres", ctx.pres0)
    r = ctx.solvePressure(FP64)
    assert r.converged and r.total_seconds > 0.0 and 0.0 < r.setup_seconds < r.total_seconds
    for k in PHASES:
        assert getattr(r, k) == 0.0, k
