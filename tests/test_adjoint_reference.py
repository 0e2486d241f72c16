"""The CPU reverse-mode reference of the adjoint workload (oracle/adjoint_ref.py).

Pinned two ways before it is used to check the device adjoint:
  * its forward run with no body force and no inflow coefficients is the
    reference's step: it tracks the C oracle (itself bitwise equal to the
    reference) step by step to 1e-10 relative (the oracle's PCG runs to
    1e-13; the reference restatement solves exactly);
  * its autograd gradients agree with central finite differences of its own
    forward run (h = 1e-6) on random directions of (f, c).
"""

import numpy as np
import pytest
import torch

import oracle as orc
from adjoint_ref import AdjointReference
from cbs3d import CaseSetup, cases

NAMES = ["intma", "coord", "mat_elem", "rho_e", "mu_e", "k_e", "rho_cp_e", "alpha_e", "Qvol_e",
         "dNkdx", "detJ", "annxf", "fedge", "iside", "face_norm", "elcoe_e", "alen_e",
         "wall_node_list", "bc_list", "bc_values", "color_ptr", "color_elem",
         "unkno", "pres", "temperature"]


def small_case(nsteps=4):
    c = cases.adjoint_channel(8, 4, 4, re=100.0, ntime=nsteps, dtfix=0.02, rel_toler=1e-13, name="adj8")
    setup = CaseSetup.from_case(c)
    d = setup.mesh_desc()
    cfg = setup.config().as_dict()
    cfg.update(npoin=d.npoin, nelem=d.nelem, nboun=d.nboun, ncolor=d.ncolor, bc_fixed=d.bc_fixed,
               npoin_wall=d.npoin_wall)
    return setup, {k: setup.array(k) for k in NAMES}, cfg


def test_reference_forward_tracks_oracle():
    setup, arrays, cfg = small_case()
    st = orc.OracleState(arrays, cfg)
    st.apply_temperature()
    st.apply_velocity()
    st.apply_pressure()
    st.build_element_pressure_terms()
    ref = AdjointReference(arrays, cfg)
    _, states = ref.run(arrays["unkno"], None, None, nsteps=4, objective_steps=1, return_states=True)
    for it in range(1, 5):
        st.advance_one_step(it)
        u_orc = st["unkno"].reshape(-1, 3)
        u_ref = states[it - 1][0].numpy()
        err = np.linalg.norm(u_ref - u_orc) / np.linalg.norm(u_orc)
        assert err < 1e-10, (it, err)
        p_orc = st["pres"]
        perr = np.linalg.norm(states[it - 1][1].numpy() - p_orc) / np.linalg.norm(p_orc)
        assert perr < 1e-9, (it, perr)


def test_reference_gradient_matches_central_differences():
    _, arrays, cfg = small_case()
    ref = AdjointReference(arrays, cfg)
    N = ref.N
    rng = np.random.default_rng(4)
    f0 = torch.tensor(rng.normal(0.0, 0.01, (N, 3)), dtype=torch.float64, requires_grad=True)
    c0 = torch.tensor(rng.normal(0.0, 0.02, 16), dtype=torch.float64, requires_grad=True)
    phi = ref.run(arrays["unkno"], f0, c0, nsteps=4, objective_steps=2)
    phi.backward()
    gf, gc = f0.grad.numpy(), c0.grad.numpy()
    assert np.abs(gc).max() > 0 and np.abs(gf).max() > 0
    h = 1e-6
    for k in range(4):
        df = rng.standard_normal((N, 3))
        dc = rng.standard_normal(16)
        nrm = np.sqrt((df ** 2).sum() + (dc ** 2).sum())
        df, dc = df / nrm, dc / nrm
        with torch.no_grad():
            fp = ref.run(arrays["unkno"], f0 + h * torch.tensor(df), c0 + h * torch.tensor(dc), 4, 2)
            fm = ref.run(arrays["unkno"], f0 - h * torch.tensor(df), c0 - h * torch.tensor(dc), 4, 2)
        fd = float(fp - fm) / (2 * h)
        ad = float((gf * df).sum() + (gc * dc).sum())
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-12) + 1e-12, (k, fd, ad)
