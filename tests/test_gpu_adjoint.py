"""The device adjoint of the CBS step (fixed time step), against
  * the reference step: with no body force and no inflow coefficients the
    adjoint run's forward is bit-identical to cbs_hip_advance_one_step;
  * the FP64 CPU reverse-mode reference (oracle/adjoint_ref.py, torch
    autograd + exact sparse solves; itself pinned to the oracle and to finite
    differences by tests/test_adjoint_reference.py): objective and both
    gradients;
  * central finite differences (h = 1e-6) of the device forward on 32 random
    directions of (f, c);
  * the torch.autograd binding.
"""

import numpy as np
import pytest
import torch

from adjoint_ref import AdjointReference
from cbs3d import CaseSetup, HipContext, cases, stability_dt
from cbs3d.adjoint import CbsObjective

pytestmark = pytest.mark.gpu

NAMES = ["intma", "coord", "mat_elem", "rho_e", "mu_e", "k_e", "rho_cp_e", "alpha_e", "Qvol_e",
         "dNkdx", "detJ", "annxf", "fedge", "iside", "face_norm", "elcoe_e", "alen_e",
         "wall_node_list", "bc_list", "bc_values", "color_ptr", "color_elem",
         "unkno", "pres", "temperature"]
NSTEPS, OBJ, K, TOL = 6, 3, 4, 1e-13


def log(msg):
    print(msg, flush=True)


def make(nx=8, ny=4, nz=4):
    c = cases.adjoint_channel(nx, ny, nz, re=100.0, ntime=NSTEPS, dtfix=0.02, rel_toler=TOL, name=f"adj{nx}")
    setup = CaseSetup.from_case(c)
    ctx = HipContext(setup)
    ctx.buildElementPressureTerms()
    return setup, ctx


def controls(n, seed=4):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 0.01, (n, 3)), rng.normal(0.0, 0.02, 16)


def forward(ctx, setup, f, c):
    ctx.set("unkno", setup.array("unkno"))
    ctx.set("pres", setup.array("pres"))
    ctx.set_body_force(f)
    ctx.set_inflow_coefficients(c)
    return ctx.adjoint_forward(NSTEPS, OBJ, K, TOL)


def test_forward_without_controls_is_the_reference_step():
    setup, ctx = make()
    log("[adjoint] forward without controls")
    ctx.adjoint_forward(NSTEPS, OBJ, K, TOL)
    u_adj, p_adj = ctx.get("unkno"), ctx.get("pres")
    _, ref = make()
    ref.initialise()
    for it in range(1, NSTEPS + 1):
        ref.advanceOneStep(it)
    assert np.array_equal(u_adj.view(np.uint64), ref.get("unkno").view(np.uint64))
    assert np.array_equal(p_adj.view(np.uint64), ref.get("pres").view(np.uint64))


def test_adjoint_matches_cpu_reverse_mode():
    setup, ctx = make()
    f, c = controls(ctx.npoin)
    log("[adjoint] device forward/backward")
    phi = forward(ctx, setup, f, c)
    gf, gc = ctx.adjoint_backward()
    d = setup.mesh_desc()
    cfg = setup.config().as_dict()
    cfg.update(npoin=d.npoin, nelem=d.nelem, nboun=d.nboun, ncolor=d.ncolor, bc_fixed=d.bc_fixed,
               npoin_wall=d.npoin_wall)
    ref = AdjointReference({k: setup.array(k) for k in NAMES}, cfg)
    nodes, phi_basis = ctx.inflow_basis()
    assert np.array_equal(np.sort(nodes - 1), np.sort(ref.inflow_nodes))
    ft = torch.tensor(f, requires_grad=True)
    ct = torch.tensor(c, requires_grad=True)
    log("[adjoint] CPU reverse mode")
    phi_ref = ref.run(setup.array("unkno"), ft, ct, NSTEPS, OBJ)
    phi_ref.backward()
    assert abs(phi - phi_ref.item()) <= 1e-10 * abs(phi_ref.item())
    ef = np.linalg.norm(gf - ft.grad.numpy()) / np.linalg.norm(ft.grad.numpy())
    ec = np.linalg.norm(gc - ct.grad.numpy()) / np.linalg.norm(ct.grad.numpy())
    log(f"[adjoint] rel. error vs CPU reverse mode: force {ef:.2e}, inflow {ec:.2e}")
    assert ef < 1e-8 and ec < 1e-8


def test_adjoint_matches_central_differences():
    setup, ctx = make()
    f, c = controls(ctx.npoin)
    forward(ctx, setup, f, c)
    gf, gc = ctx.adjoint_backward()
    rng = np.random.default_rng(7)
    h = 1e-6
    gnorm = np.sqrt((gf ** 2).sum() + (gc ** 2).sum())
    abs_err, rel_err = [], []
    for k in range(32):
        df = rng.standard_normal(f.shape)
        dc = rng.standard_normal(16)
        nrm = np.sqrt((df ** 2).sum() + (dc ** 2).sum())
        df, dc = df / nrm, dc / nrm
        fp = forward(ctx, setup, f + h * df, c + h * dc)
        fm = forward(ctx, setup, f - h * df, c - h * dc)
        fd = (fp - fm) / (2 * h)
        ad = float((gf * df).sum() + (gc * dc).sum())
        abs_err.append(abs(fd - ad))
        rel_err.append(abs(fd - ad) / max(abs(ad), 1e-300))
    # unit directions: |<g, d>| <= |g|, so the FD noise (forward solves
    # converged to 1e-13, amplified by 1/h) is measured against |g|
    log(f"[adjoint] 32 directions: max |fd - ad| / |g| = {max(abs_err) / gnorm:.2e}, "
        f"median |fd - ad| / |ad| = {np.median(rel_err):.2e}, max {max(rel_err):.2e}")
    assert max(abs_err) / gnorm < 1e-6
    assert np.median(rel_err) < 2e-5


def test_torch_autograd_binding():
    setup, ctx = make()
    f, c = controls(ctx.npoin, seed=9)
    ft = torch.tensor(f, requires_grad=True)
    ct = torch.tensor(c, requires_grad=True)
    phi = CbsObjective.apply(ft, ct, ctx, setup.array("unkno"), setup.array("pres"), NSTEPS, OBJ, K, TOL)
    (2.0 * phi).backward()
    phi2 = forward(ctx, setup, f, c)
    gf, gc = ctx.adjoint_backward(2.0)
    assert float(phi) == phi2
    assert np.array_equal(ft.grad.numpy(), gf) and np.array_equal(ct.grad.numpy(), gc)


def test_channel64_adjoint_matches_cpu_reverse_mode():
    """The adjoint workload at full size: channel 64x16x16x6, Re 100, 100
    steps with dt fixed at the channel's first-step stability time step, the
    objective over the last 20 steps, f ~ N(0, 0.01^2) per node (seed 4),
    inflow coefficients 0; checkpoints every 10 steps."""
    w = cases.ADJOINT_WORKLOAD
    dt = stability_dt(cases.channel(w["nx"], w["ny"], w["nz"], re=w["re"], pr=0.71, jitter=0.0, seed=4,
                                    ntime=w["ntime"]))
    log(f"[adjoint] channel64: dtfix = {dt!r}")
    case = cases.adjoint_channel(w["nx"], w["ny"], w["nz"], re=w["re"], ntime=w["ntime"], dtfix=dt,
                                 rel_toler=1e-12)
    setup = CaseSetup.from_case(case)
    ctx = HipContext(setup)
    ctx.buildElementPressureTerms()
    f = np.random.default_rng(4).normal(0.0, 0.01, (ctx.npoin, 3))
    c = np.zeros(16)
    ctx.set_body_force(f)
    ctx.set_inflow_coefficients(c)
    phi = ctx.adjoint_forward(w["ntime"], w["objective_steps"], 10, 1e-12)
    gf, gc = ctx.adjoint_backward()
    log(f"[adjoint] channel64: Phi = {phi!r}; CPU reverse mode ...")
    d = setup.mesh_desc()
    cfg = setup.config().as_dict()
    cfg.update(npoin=d.npoin, nelem=d.nelem, nboun=d.nboun, ncolor=d.ncolor, bc_fixed=d.bc_fixed,
               npoin_wall=d.npoin_wall)
    ref = AdjointReference({k: setup.array(k) for k in NAMES}, cfg)
    ft = torch.tensor(f, requires_grad=True)
    ct = torch.tensor(c, requires_grad=True)
    phi_ref = ref.run(setup.array("unkno"), ft, ct, w["ntime"], w["objective_steps"])
    phi_ref.backward()
    ef = np.linalg.norm(gf - ft.grad.numpy()) / np.linalg.norm(ft.grad.numpy())
    ec = np.linalg.norm(gc - ct.grad.numpy()) / np.linalg.norm(ct.grad.numpy())
    log(f"[adjoint] channel64: Phi rel. diff {abs(phi - phi_ref.item()) / abs(phi_ref.item()):.2e}, "
        f"grad rel. error force {ef:.2e}, inflow {ec:.2e}")
    assert abs(phi - phi_ref.item()) <= 1e-9 * abs(phi_ref.item())
    assert ef < 1e-7 and ec < 1e-7
