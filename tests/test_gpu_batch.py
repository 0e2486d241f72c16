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
from cbs3d import AMG, FP64, MIXED, BatchContext, CaseSetup, CbsError, HipContext, cases

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


def run(ctx, nsteps, precision=FP64):
    ctx.initialise()
    ctx.buildElementPressureTerms()
    return [ctx.advanceOneStep(it, precision=precision) for it in range(1, nsteps + 1)]


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


def ensemble_smallest():
    """The members of the seeded ensemble workload (cases.ensemble: 64
    cavities, n ~ U{12..64}, Re log-uniform in [50, 1000], seed 3) with
    n <= 16, as the workload generates them."""
    params = cases.ensemble_params(64, 3)
    picked = [(k, n, re) for k, (n, re) in enumerate(params) if n <= 16]
    assert len(picked) >= 3
    return [cases.cavity(n, re=re, name=f"ensemble3_{k:02d}_n{n}") for k, n, re in picked]


def test_batch_mixed_members_match_standalone_bitwise():
    """The mixed form (FP32 apply, FP64 reductions, reliable updates) in a
    batch: each member pauses, refreshes its residual and finishes on its
    own schedule, and ends with the bits, iterations and refinements of its
    standalone mixed run."""
    cs = ensemble_smallest()
    setups = [CaseSetup.from_case(c) for c in cs]
    nsteps = 3
    log(f"[batch] mixed union run of {len(cs)} members: n = {[c.nx for c in cs]}")
    batch = BatchContext(setups)
    res = run(batch, nsteps, MIXED)
    assert res[-1].converged
    for k, setup in enumerate(setups):
        ctx = HipContext(setup)
        ref = run(ctx, nsteps, MIXED)
        mr = batch.member_cg_result(k)
        log(f"[batch] member {k} (n {cs[k].nx}): iterations {mr.iterations} vs {ref[-1].iterations}, "
            f"refinements {mr.refinements} vs {ref[-1].refinements}")
        assert mr.converged and ref[-1].converged
        assert mr.iterations == ref[-1].iterations
        assert mr.refinements == ref[-1].refinements
        assert mr.final_relative_l2 == ref[-1].final_relative_l2
        for name in FIELDS:
            assert_bits(f"member {k} {name}", batch.member(k, name), ctx.get(name))
        ctx.close()
    assert res[-1].iterations == max(batch.member_cg_result(k).iterations for k in range(len(cs)))
    batch.close()


def test_batch_rejects_multigrid_solve():
    setups = [CaseSetup.from_case(c) for c in members()[:2]]
    batch = BatchContext(setups)
    batch.initialise()
    batch.buildElementPressureTerms()
    with pytest.raises(CbsError) as e:
        batch.advanceOneStep(1, precision=AMG)
    assert "multigrid" in str(e.value)


def test_single_member_batch_equals_context():
    setup = CaseSetup.from_case(members()[1])
    batch = BatchContext([setup])
    run(batch, 2)
    ctx = HipContext(setup)
    run(ctx, 2)
    for name in FIELDS:
        assert_bits(name, batch.get(name), ctx.get(name))
