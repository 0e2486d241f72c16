"""Batched ensembles on the GPU: every member of a batch context evolves bit
for bit as a context of its own.

A batch holds the disjoint union of its members; each member's reductions
(time-step minimum, convergence norms, every PCG dot product) are cut into
the same blocks as in a context holding that member alone, and its tiles are
the same tiles, so the union run must reproduce each standalone run exactly
(same iterations, same bits) while sharing every launch.  The standalone
runs are themselves pinned to the oracle and the reference goldens by
test_gpu_parity.py.
"""

import numpy as np
import pytest

from helpers import assert_bits
from cbs3d import FP64, MIXED, BatchContext, CaseSetup, CbsError, HipContext, cases

pytestmark = pytest.mark.gpu

FIELDS = ["unkno", "pres", "temperature", "velocity", "delte", "deltp", "deltp1", "elcoe2", "pdiag", "rhs1"]


def log(msg):
    print(msg, flush=True)


def members():
    # different sizes, Reynolds numbers and jitter: different time steps,
    # colour counts and PCG iteration counts per member
    return [cases.cavity(4, re=80.0, name="m0"),
            cases.cavity(7, re=420.0, jitter=0.1, seed=5, name="m1"),
            cases.cavity(5, re=150.0, jitter=0.05, seed=6, name="m2")]


def run(ctx, nsteps):
    ctx.initialise()
    ctx.buildElementPressureTerms()
    return [ctx.advanceOneStep(it) for it in range(1, nsteps + 1)]


def test_batch_members_match_standalone_bitwise():
    cs = members()
    setups = [CaseSetup.from_case(c) for c in cs]
    nsteps = 4
    log("[batch] union run")
    batch = BatchContext(setups)
    assert batch.members == len(cs)
    res = run(batch, nsteps)
    for k, setup in enumerate(setups):
        log(f"[batch] standalone member {k}")
        ctx = HipContext(setup)
        ref = run(ctx, nsteps)
        mr = batch.member_cg_result(k)
        assert mr.iterations == ref[-1].iterations, (k, mr.iterations, ref[-1].iterations)
        assert mr.converged
        assert mr.final_relative_l2 == ref[-1].final_relative_l2
        for name in FIELDS:
            assert_bits(f"member {k} {name}", batch.member(k, name), ctx.get(name))
        ms, ss = batch.member_scalars(k), ctx.scalars()
        assert ms.dtreal == ss.dtreal
        assert list(ms.hb) == list(ss.hb)
        assert ms.last_cg_iterations == ss.last_cg_iterations
    # the batch's own result reports the slowest member and all converged
    assert res[-1].converged
    assert res[-1].iterations == max(batch.member_cg_result(k).iterations for k in range(len(cs)))


def test_batch_rejects_mixed_precision_solve():
    setups = [CaseSetup.from_case(c) for c in members()[:2]]
    batch = BatchContext(setups)
    batch.initialise()
    batch.buildElementPressureTerms()
    with pytest.raises(CbsError) as e:
        batch.advanceOneStep(1, precision=MIXED)
    assert "FP64 only" in str(e.value)


def test_single_member_batch_equals_context():
    setup = CaseSetup.from_case(members()[1])
    batch = BatchContext([setup])
    run(batch, 2)
    ctx = HipContext(setup)
    run(ctx, 2)
    for name in FIELDS:
        assert_bits(name, batch.get(name), ctx.get(name))
