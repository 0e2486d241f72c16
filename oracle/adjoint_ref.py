"""FP64 CPU reverse-mode reference of the fixed-time-step CBS step.  TEST INFRASTRUCTURE.

Only the tests (and adjoint checks run beside the bench) use this module, as
the checker of the device adjoint.  It restates, with torch tensors in
float64 on the CPU, the per-step map that the device adjoint differentiates:

  u*pre = u^n + (R1(u^n) + m f) w,          u* = BC(u*pre)
  p_f   = A_ff^-1 (R2(u*, u^n) - A_fc p_c),  p_c fixed values, inactive 0
  u**pre = u* + R3(p) w,                     u^{n+1} = BC(u**pre)

with w = elcoe2 = 1 / sum_e(m_ea / dt), m the lumped fluid mass, R1 the Step 1
element terms (src/assembly/MomentumAssembly.cpp:144-432), R2 the Step 2 RHS
(src/assembly/PressureAssembly.cpp:236-335), R3 the Step 3 gradient term
(src/solver/Steps.cpp:377-428), BC the velocity package (symmetry, last
prescribed value, outlet backflow clip, solid zeroing:
src/boundary/Boundary.cpp:275-723, src/solver/Steps.cpp:102-176) and A the
pressure operator dt * (gstifE, pdiagE) (src/timestep/TimeStep.cpp:873-901).
Controls: the nodal body force f (an acceleration) and 16 Fourier
coefficients c added to the BC-508 inflow profile (the build's forward-model
extensions; both zero reproduce the reference).  Objective: the sum over the
last `objective_steps` steps of sum_e nu_e V_e |grad u|^2_e.

Gradients come from torch.autograd through every element operation; the
pressure solve is an exact sparse LU solve whose backward is the solve with
the transposed factor (the implicit-function adjoint), so this reference is
independent of the device code's hand-written transposes.  Sums run in
vectorised order, not the reference's: the forward agrees with the oracle
to rounding (tests/test_adjoint_reference.py pins it).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

# nodes (0-based) of local face is (the face opposite node is): the table of
# the reference's ippn1 shifted to 0-based ids
FACE_NODES = [[2, 1, 3], [2, 3, 0], [0, 3, 1], [0, 1, 2]]
PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def rectangular_profile_u(y):
    return 6.0 * y * (1.0 - y)


class _LuSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, lu):
        ctx.lu = lu
        return torch.from_numpy(lu.solve(b.detach().numpy()))

    @staticmethod
    def backward(ctx, g):
        return torch.from_numpy(ctx.lu.solve(g.detach().numpy(), trans="T")), None


class AdjointReference:
    """Reverse-mode reference for one case; `arrays` holds the reference's
    arrays (CaseSetup.array names), `cfg` its run controls (with npoin,
    nelem, nboun, bc_fixed, npoin_wall)."""

    def __init__(self, arrays: dict, cfg: dict):
        t = lambda a: torch.as_tensor(np.asarray(a, np.float64))
        self.N = N = int(cfg["npoin"])
        self.E = E = int(cfg["nelem"])
        B = int(cfg["nboun"])
        self.conn = torch.as_tensor(np.asarray(arrays["intma"], np.int64).reshape(E, 4) - 1)
        mat = np.asarray(arrays["mat_elem"]).reshape(E)
        self.fluid = torch.as_tensor((mat == 0).astype(np.float64))
        self.dN = t(arrays["dNkdx"]).reshape(E, 3, 4)          # [e][d][a]
        self.detJ = t(arrays["detJ"]).reshape(E)
        self.ann = t(arrays["annxf"]).reshape(E, 4, 4)         # [e][face][comp]
        fedge = np.asarray(arrays["fedge"]).reshape(E, 4)
        self.fnz = torch.as_tensor((fedge != 0).astype(np.float64))
        self.fone = torch.as_tensor((fedge == 1).astype(np.float64))
        elcoe = np.asarray(arrays["elcoe_e"], np.float64).reshape(E, 4)
        self.dt = float(cfg["dtfix"])
        if int(cfg.get("dimensional_mode", 0)) > 0 and int(cfg.get("material_properties_enabled", 0)) > 0:
            self.nu = t(np.asarray(arrays["mu_e"], np.float64) / np.asarray(arrays["rho_e"], np.float64))
        else:
            self.nu = torch.full((E,), float(cfg["ani"]), dtype=torch.float64)
        conn = self.conn.numpy()
        fl = mat == 0
        mass = np.zeros(N)
        np.add.at(mass, conn[fl].reshape(-1), elcoe[fl].reshape(-1))
        m2 = np.zeros(N)
        np.add.at(m2, conn[fl].reshape(-1), (elcoe[fl] / self.dt).reshape(-1))
        touch_fluid = np.zeros(N, bool)
        touch_fluid[conn[fl].reshape(-1)] = True
        touch_solid = np.zeros(N, bool)
        touch_solid[conn[~fl].reshape(-1)] = True
        self.mass = t(mass)
        self.w = t(np.where(touch_fluid, 1.0 / np.where(m2 > 0, m2, 1.0), 0.0))
        # pressure operator and constraints
        vol = np.asarray(arrays["detJ"], np.float64) / 6.0
        dn = self.dN.numpy()
        rows, cols, vals = [], [], []
        diag = np.zeros(N)
        for a in range(4):
            pd = vol * np.einsum("ed,ed->e", dn[:, :, a], dn[:, :, a])
            np.add.at(diag, conn[fl, a], (pd[fl] * self.dt))
        for a, b in PAIRS:
            g = vol * np.einsum("ed,ed->e", dn[:, :, a], dn[:, :, b]) * self.dt
            rows += [conn[fl, a], conn[fl, b]]
            cols += [conn[fl, b], conn[fl, a]]
            vals += [g[fl], g[fl]]
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
        A = A + sp.diags(diag)
        fixed = np.zeros(N, bool)
        fixed_val = np.zeros(N)
        bcl = np.asarray(arrays["bc_list"], np.int64).reshape(-1)[: int(cfg["bc_fixed"])] - 1
        bcv = np.asarray(arrays["bc_values"], np.float64).reshape(-1)[: int(cfg["bc_fixed"])]
        for ip, v in zip(bcl, bcv):  # last writer wins
            fixed[ip] = True
            fixed_val[ip] = v
        free = touch_fluid & ~fixed
        self.free_idx = torch.as_tensor(np.nonzero(free)[0])
        self.pc = t(np.where(fixed, fixed_val, 0.0))
        Aff = A[free][:, free].tocsc()
        self.Afc = A[free]  # times the constrained vector (zeros off the fixed nodes)
        self.lu = spla.splu(Aff)
        # velocity boundary programme in the reference's order
        iside = np.asarray(arrays["iside"], np.int64).reshape(B, 6)
        fnorm = np.asarray(arrays["face_norm"], np.float64).reshape(B, 4)
        coord = np.asarray(arrays["coord"], np.float64).reshape(N, 3)
        sym = [[] for _ in range(N)]
        out = [[] for _ in range(N)]
        val = {}
        src = {}

        def prescribe(ip, u, v, w, s=0):
            val[ip] = (u, v, w)
            src[ip] = s

        for ib in range(B):
            if iside[ib, 5] == 506:
                n = fnorm[ib, :3] / fnorm[ib, 3]
                for k in range(3):
                    sym[iside[ib, k] - 1].append(n)
        iu, iv, iw = float(cfg.get("inlet_u", 0.0)), float(cfg.get("inlet_v", 0.0)), float(cfg.get("inlet_w", 0.0))
        for ib in range(B):
            bc = iside[ib, 5]
            for k in range(3):
                ip = iside[ib, k] - 1
                y = coord[ip, 1]
                if bc in (500, 510):
                    prescribe(ip, iu, iv, iw)
                elif bc in (501, 502, 530, 532, 901):
                    prescribe(ip, 0.0, 0.0, 0.0)
                elif bc == 503:
                    prescribe(ip, 1.0, 0.0, 0.0)
                elif bc == 508:
                    prescribe(ip, rectangular_profile_u(y), 0.0, 0.0, 508)
                elif bc in (507, 511):
                    raise NotImplementedError(f"BC {bc} is not used by the adjoint cases")
        for ip in np.asarray(arrays["wall_node_list"], np.int64).reshape(-1)[: int(cfg["npoin_wall"])] - 1:
            prescribe(ip, 0.0, 0.0, 0.0)
        for ib in range(B):
            if iside[ib, 5] == 500:
                for k in range(3):
                    prescribe(iside[ib, k] - 1, iu, iv, iw)
        solid = touch_solid
        for ip in np.nonzero(solid)[0]:
            prescribe(ip, 0.0, 0.0, 0.0)
        for ib in range(B):
            if iside[ib, 5] == 520:
                n = fnorm[ib, :3] / fnorm[ib, 3]
                for k in range(3):
                    out[iside[ib, k] - 1].append(n)
        pres_nodes = np.array(sorted(val), np.int64)
        self.pres_nodes = torch.as_tensor(pres_nodes)
        self.pres_base = t(np.array([val[i] for i in pres_nodes]).reshape(-1, 3))
        inflow = np.array([i for i in pres_nodes if src[i] == 508], np.int64)
        self.inflow_nodes = inflow
        pos = {int(n): k for k, n in enumerate(pres_nodes)}
        self.inflow_rows = torch.as_tensor(np.array([pos[int(i)] for i in inflow], np.int64))
        yz = coord[inflow][:, 1:3] if inflow.size else np.zeros((0, 2))
        phi = np.zeros((inflow.size, 16))
        if inflow.size:
            lo, hi = yz.min(0), yz.max(0)
            yh = (yz[:, 0] - lo[0]) / (hi[0] - lo[0]) if hi[0] > lo[0] else 0 * yz[:, 0]
            zh = (yz[:, 1] - lo[1]) / (hi[1] - lo[1]) if hi[1] > lo[1] else 0 * yz[:, 1]
            for j in range(1, 5):
                for k in range(1, 5):
                    phi[:, (j - 1) * 4 + k - 1] = np.sin(j * np.pi * yh) * np.sin(k * np.pi * zh)
        self.phi = t(phi)
        self.solid = torch.as_tensor(solid)
        self.sym = self._normals_table(sym)
        self.out = self._normals_table(out)

    @staticmethod
    def _normals_table(lists):
        """(K, N, 3) normals and (K, N) masks: the k-th normal of every node."""
        K = max((len(v) for v in lists), default=0)
        N = len(lists)
        n = np.zeros((K, N, 3))
        m = np.zeros((K, N), bool)
        for i, v in enumerate(lists):
            for k, vec in enumerate(v):
                n[k, i] = vec
                m[k, i] = True
        return torch.as_tensor(n), torch.as_tensor(m)

    # ---- pieces of the step -------------------------------------------
    def prescribed_values(self, c):
        vals = self.pres_base.clone()
        if c is not None and self.inflow_rows.numel():
            add = self.phi @ c
            vals = vals.index_put((self.inflow_rows, torch.zeros_like(self.inflow_rows)),
                                  vals[self.inflow_rows, 0] + add)
        return vals

    def bc(self, U, which, vals):
        if which == 0:
            n, m = self.sym
            for k in range(n.shape[0]):
                un = (U * n[k]).sum(1, keepdim=True)
                U = torch.where(m[k][:, None], U - un * n[k], U)
        U = U.index_put((self.pres_nodes,), vals)
        if which == 0:
            n, m = self.out
            for k in range(n.shape[0]):
                un = (U * n[k]).sum(1, keepdim=True)
                act = m[k][:, None] & (un.detach() < 0.0) & ~self.solid[:, None]
                U = torch.where(act, U - un * n[k], U)
            U = torch.where(self.solid[:, None], torch.zeros_like(U), U)
        return U

    def _scatter(self, lrhs):  # lrhs (E, 4, k) -> (N, k)
        out = torch.zeros((self.N, lrhs.shape[2]), dtype=torch.float64)
        return out.index_add(0, self.conn.reshape(-1), lrhs.reshape(-1, lrhs.shape[2]))

    def step1_rhs(self, U):
        Ue = U[self.conn]                                       # [e][a][i]
        P = Ue[:, :, :, None] * Ue[:, :, None, :]              # [e][a][i][j]
        S = P.sum(1)                                           # [e][i][j]
        dN, dj = self.dN, self.detJ
        lrhs = torch.einsum("eja,eij->eai", dN, S) * (dj / 24.0)[:, None, None]
        for f in range(4):
            mask = self.fnz[:, f][:, None]
            lsf = S - P[:, f]
            ff = self.ann[:, f, :3] / 12.0
            for a in FACE_NODES[f]:
                lrhs[:, a, :] = lrhs[:, a, :] - mask * torch.einsum("eij,ej->ei", lsf + P[:, a], ff)
        T = torch.einsum("eja,eaij->eij", dN, P)
        sumj = T.sum(2)
        umean = Ue.sum(1) / 4.0
        K = sumj[:, :, None] * umean[:, None, :]               # [e][i][k]
        ld = 0.5 * self.dt * dj / 6.0
        ld2 = 0.5 * self.dt / 3.0
        lrhs = lrhs - torch.einsum("eka,eik->eai", dN, K) * ld[:, None, None]
        for f in range(4):
            mask = self.fnz[:, f][:, None]
            add = torch.einsum("eik,ek->ei", K, self.ann[:, f, :3]) * ld2
            for a in FACE_NODES[f]:
                lrhs[:, a, :] = lrhs[:, a, :] + mask * add
        G = torch.einsum("eja,eai->eij", dN, Ue)
        vfd = self.nu * dj / 6.0
        lrhs = lrhs - torch.einsum("eja,eij->eai", dN, G) * vfd[:, None, None]
        fdf = self.nu / 3.0
        for f in range(4):
            mask = self.fnz[:, f][:, None]
            add = torch.einsum("eij,ej->ei", G, self.ann[:, f, :3]) * fdf[:, None]
            for a in FACE_NODES[f]:
                lrhs[:, a, :] = lrhs[:, a, :] + mask * add
        return self._scatter(lrhs * self.fluid[:, None, None])

    def step2_rhs(self, Us, Un):
        vol4 = self.detJ / 24.0
        us = Us[self.conn].sum(1) * vol4[:, None]             # [e][d]
        l = torch.einsum("eda,ed->ea", self.dN, us)
        for f in range(4):
            mask = self.fone[:, f]
            fus = Un[self.conn[:, FACE_NODES[f]]].sum(1)
            nf = self.ann[:, f, :3] / 12.0
            for a in FACE_NODES[f]:
                l[:, a] = l[:, a] - mask * ((fus + Un[self.conn[:, a]]) * nf).sum(1)
        l = l * self.fluid[:, None]
        return self._scatter(l[:, :, None])[:, 0]

    def step3_rhs(self, p):
        vol4 = self.detJ / 24.0
        gp = torch.einsum("eda,ea->ed", self.dN, p[self.conn]) * vol4[:, None] * self.fluid[:, None]
        return self._scatter((-gp)[:, None, :].expand(-1, 4, -1).contiguous())

    def solve(self, b):
        rhs = b[self.free_idx] - torch.as_tensor(self.Afc @ self.pc.numpy())
        pf = _LuSolve.apply(rhs, self.lu)
        return self.pc.index_put((self.free_idx,), pf)

    def objective_term(self, U):
        G = torch.einsum("eja,eai->eij", self.dN, U[self.conn])
        return (self.nu * self.detJ / 6.0 * (G * G).sum((1, 2)) * self.fluid).sum()

    # ---- the run ---------------------------------------------------------
    def run(self, u_raw, f, c, nsteps: int, objective_steps: int, return_states: bool = False):
        """Phi of the run from the raw initial velocity u_raw (N, 3); f (N, 3)
        or None, c (16,) or None (torch tensors, possibly requiring grad)."""
        vals = self.prescribed_values(c)
        U = self.bc(torch.as_tensor(np.asarray(u_raw, np.float64).reshape(self.N, 3)), 1, vals)
        phi = torch.zeros((), dtype=torch.float64)
        states = []
        for it in range(1, nsteps + 1):
            rhs = self.step1_rhs(U)
            if f is not None:
                rhs = rhs + self.mass[:, None] * f
            Us = self.bc(U + rhs * self.w[:, None], 0, vals)
            p = self.solve(self.step2_rhs(Us, U))
            U = self.bc(Us + self.step3_rhs(p) * self.w[:, None], 0, vals)
            if it > nsteps - objective_steps:
                phi = phi + self.objective_term(U)
            if return_states:
                states.append((U.detach().clone(), p.detach().clone()))
        return (phi, states) if return_states else phi
