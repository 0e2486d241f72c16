"""torch.autograd binding of the device adjoint.

    phi = CbsObjective.apply(f, c, ctx, u0, p0, nsteps, objective_steps)
    phi.backward()            # f.grad, c.grad from the hand-written adjoint

f is the nodal body force (N, 3) and c the 16 inflow coefficients (float64
CPU tensors); ctx a HipContext of a fixed-time-step case whose element
pressure terms are built; u0 (N, 3) and p0 (N,) the raw initial fields (the
run applies the initial boundary values itself).  The forward runs
cbs_hip_adjoint_forward on the GPU and returns the objective; the backward
runs cbs_hip_adjoint_backward, which differentiates the last forward run of
that context, so each forward is to be followed by its backward before the
context runs another.
"""

from __future__ import annotations

import numpy as np
import torch


class CbsObjective(torch.autograd.Function):
    @staticmethod
    def forward(actx, f, c, hip, u0, p0, nsteps, objective_steps, checkpoint_every=10, rel_tol=1e-12):
        hip.set("unkno", np.asarray(u0, np.float64))
        hip.set("pres", np.asarray(p0, np.float64))
        hip.set_body_force(f.detach().cpu().numpy())
        hip.set_inflow_coefficients(c.detach().cpu().numpy())
        phi = hip.adjoint_forward(int(nsteps), int(objective_steps), int(checkpoint_every), float(rel_tol))
        actx.hip = hip
        return torch.tensor(phi, dtype=torch.float64)

    @staticmethod
    def backward(actx, g):
        gf, gc = actx.hip.adjoint_backward(float(g))
        return (torch.from_numpy(gf), torch.from_numpy(gc), None, None, None, None, None, None, None)
