"""Host-side checks of batched ensembles (no GPU needed): the ensemble
generator and the validation of member run controls, which happens before
any device call."""

import ctypes as C

import numpy as np
import pytest

from cbs3d import BatchContext, CaseSetup, CbsError, cases
from cbs3d._native import hip_lib


def test_ensemble_generator_is_deterministic_and_in_range():
    a = cases.ensemble_params()
    assert a == cases.ensemble_params()
    assert len(a) == 64
    ns = np.array([n for n, _ in a])
    res = np.array([r for _, r in a])
    assert ns.min() >= 12 and ns.max() <= 64
    assert res.min() >= 50.0 and res.max() <= 1000.0
    # roughly uniform sizes, log-uniform Re: about half below the midpoints
    assert 16 <= int((ns <= 38).sum()) <= 48
    assert 16 <= int((res < np.sqrt(50.0 * 1000.0)).sum()) <= 48
    # other seeds give other ensembles
    assert cases.ensemble_params(seed=4) != a
    # members are the cavity recipe with their own size and Re
    small = cases.ensemble(count=3, nmin=3, nmax=5)
    for c, (n, re) in zip(small, cases.ensemble_params(count=3, nmin=3, nmax=5)):
        assert c.nx == c.ny == c.nz == n and c.params.re == re
        assert c.params.inlet_u == 1.0 and c.params.csafm == 0.5 and c.params.cg_conv_test == 2


def test_batch_rejects_members_with_other_run_controls():
    c0 = cases.cavity(3, re=100.0, name="a")
    c1 = cases.cavity(3, re=200.0, name="b")
    c1.params.csafm = 0.25  # a run control other than ani/re/pnode/ntime
    setups = [CaseSetup.from_case(c0), CaseSetup.from_case(c1)]
    with pytest.raises(CbsError) as e:
        BatchContext(setups)
    assert "member 1 differs from member 0" in str(e.value)


def test_batch_abi_rejects_bad_member_count():
    lib = hip_lib()
    h = C.c_void_p()
    assert lib.cbs_hip_create_batch(0, None, None, C.byref(h)) == 1
    assert b"member count" in lib.cbs_hip_last_error(None)
