"""PCG iterations replayed from captured HIP graphs against per-kernel launches.

Each solve replays the captured block of iterations many times, with the
solver's own eager work (constraints, true residuals, chunk-size decisions)
and host synchronisations between replays.  The replayed kernels are the
eager loop's kernels with the same arguments in the same order, so every
step of both solver forms must give the same bits and iteration counts."""

import numpy as np
import pytest

from cbs3d import FP64, MIXED, CaseSetup, HipContext, cases

pytestmark = pytest.mark.gpu


def run(case, precision, graphs, steps=3):
    ctx = HipContext(CaseSetup.from_case(case))
    ctx.set_graphs(graphs)
    ctx.set_profile(True)  # the timed block (events around an apply) is captured too
    ctx.initialise()
    ctx.buildElementPressureTerms()
    out = []
    for it in range(1, steps + 1):
        r = ctx.advanceOneStep(it, precision=precision)
        out.append((r.iterations, ctx.get("pres"), ctx.get("unkno")))
    active = ctx.graphs_active()[0 if precision == FP64 else 1]
    timed = ctx.apply_timer(precision)[1]
    return out, active, timed


@pytest.mark.parametrize("precision", [FP64, MIXED], ids=["fp64", "mixed"])
def test_graph_replay_matches_eager(precision):
    case = cases.cavity(10, re=250.0, jitter=0.15, seed=3, name="cav10j")
    eager, active_e, _ = run(case, precision, graphs=False)
    graph, active_g, timed = run(case, precision, graphs=True)
    assert not active_e
    assert active_g, "the PCG block was not captured"
    assert timed > 0, "no apply was timed from the captured block"
    for k, ((i1, p1, u1), (i2, p2, u2)) in enumerate(zip(eager, graph)):
        print(f"[gpu graphs] step {k + 1}: {i1} iterations eager, {i2} replayed", flush=True)
        assert i1 == i2
        assert np.array_equal(p1, p2)
        assert np.array_equal(u1, u2)


@pytest.mark.parametrize("precision", [FP64, MIXED], ids=["fp64", "mixed"])
def test_apply_launch_shape_does_not_change_results(precision):
    """Two contexts with different apply launch shapes (threads per row tile,
    column pairs per batch) give the same bits: row sums and the tiles'
    (p, Ap) partials are taken in an order fixed by the operator's layout."""
    case = cases.cavity(10, re=250.0, jitter=0.15, seed=3, name="cav10j")
    outs = []
    for threads, batch in ((1024, 8), (256, 4)):
        ctx = HipContext(CaseSetup.from_case(case))
        ctx.set_apply_config(FP64, threads, batch)
        ctx.set_apply_config(MIXED, threads, batch)
        ctx.initialise()
        ctx.buildElementPressureTerms()
        steps = []
        for it in range(1, 3):
            r = ctx.advanceOneStep(it, precision=precision)
            steps.append((r.iterations, ctx.get("pres")))
        assert ctx.apply_config(precision)[:2] == (threads, batch)
        outs.append(steps)
    for (i1, p1), (i2, p2) in zip(*outs):
        assert i1 == i2
        assert np.array_equal(p1, p2)


@pytest.mark.parametrize("precision", [FP64, MIXED], ids=["fp64", "mixed"])
def test_direction_fusion_does_not_change_results(precision):
    """The direction update formed inside the apply's staging and the one
    in its own kernel give the same bits (same expressions on the same
    values), with graphs and with the mixed form's refreshes; the default
    context times both at its first solve and records its choice."""
    case = cases.cavity(10, re=250.0, jitter=0.15, seed=3, name="cav10j")
    outs = []
    for mode in (0, 1, -1):
        ctx = HipContext(CaseSetup.from_case(case))
        if mode >= 0:
            ctx.set_direction_fusion(precision, mode)
        ctx.initialise()
        ctx.buildElementPressureTerms()
        steps = []
        for it in range(1, 4):
            r = ctx.advanceOneStep(it, precision=precision)
            steps.append((r.iterations, r.refinements, ctx.get("pres"), ctx.get("unkno")))
        fused, ms_sep, ms_fused = ctx.direction_fusion(precision)
        print(f"[gpu fusion] mode {mode}: fused {fused}, timed {ms_sep} ms separate, {ms_fused} ms fused, "
              f"iterations {[s[0] for s in steps]}", flush=True)
        if mode >= 0:
            assert fused == mode and ms_sep == 0.0 and ms_fused == 0.0
        else:
            assert fused in (0, 1) and ms_sep > 0.0 and ms_fused > 0.0
        outs.append(steps)
    for other in outs[1:]:
        for (i1, f1, p1, u1), (i2, f2, p2, u2) in zip(outs[0], other):
            assert (i1, f1) == (i2, f2)
            assert np.array_equal(p1, p2)
            assert np.array_equal(u1, u2)
