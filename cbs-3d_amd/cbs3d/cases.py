"""Synthetic structured-split tetrahedral cases in the CBS3D++ case-file formats.

A box [0,Lx]x[0,Ly]x[0,Lz] is cut into nx*ny*nz hexahedra and every hexahedron
into the six Kuhn tetrahedra that share its (0,0,0)-(1,1,1) diagonal, one per
axis permutation.  Negatively oriented tetrahedra get local nodes 2 and 3
swapped, the convention of tools/export_cbs3d_tet.py:225-249.  Elements are
hex-major (x fastest) with six tetrahedra per hex; node (i,j,k) has the 1-based
number 1 + i + (nx+1)*(j + (ny+1)*k).

Boundary triangles carry their parent tetrahedron and a raw cube-side id
1..6 = xmin, xmax, ymin, ymax, zmin, zmax (tools/export_cbs3d_tet.py:62-83).
Interior nodes may be jittered by a uniform +-alpha*h per coordinate drawn from
SplitMix64 (u = (x >> 11) * 2**-53), so the case is bit-for-bit reproducible
from its seed on any machine.  Boundary nodes stay on their planes.

The writers produce the reference's positional .plt/.bco/.par/.material/.matprop
files (src/io/MeshIO.cpp:613-1623), with coordinates printed "%.16e" so both
codes read identical doubles.
"""

from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Dict, Optional

import numpy as np

# Local face f is opposite local node f (include/cbs/core/CBSStateSI.hpp:342-365).
IPPN1 = np.array([[3, 2, 4], [3, 4, 1], [1, 4, 2], [1, 2, 3]], dtype=np.int64)

# Solver BC ids (include/cbs/core/RunConfig.hpp:258-278).
BC_LID = 500
BC_WALL_T0 = 502
BC_PARABOLIC_INLET = 508
BC_OUTLET = 520
BC_WALL_ADIABATIC = 530

_MASK64 = (1 << 64) - 1


def splitmix64_stream(seed: int, count: int) -> np.ndarray:
    """Return `count` successive SplitMix64 outputs (uint64) for `seed`."""
    golden = np.uint64(0x9E3779B97F4A7C15)
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK64) + idx * golden
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """Doubles in [0,1): the top 53 bits of SplitMix64 times 2**-53."""
    return (splitmix64_stream(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclasses.dataclass
class CaseParams:
    """The .par controls the case writer emits (layout: src/io/MeshIO.cpp:986-1437)."""

    ntime: int = 200
    temp_calc: int = 0
    cinf: tuple = (0.0, 0.0, 0.0, 0.0, 0.0)
    dtfixed: int = 0
    dtfix: float = 1.0
    csafm: float = 0.5
    ilots: int = 0
    htype: int = 1
    # .par line 14 (MeshIO.cpp:153-188): rem_deltp, beta_opt, dtfix_end, deltr
    rem_deltp: int = 0
    beta_opt: int = 1
    dtfix_end: int = 0
    deltr: float = 1.0e20
    re: float = 100.0
    pr: float = 0.71
    pnode: int = 1
    rel_toler: float = 1.0e-8
    abs_toler: float = 1.0e-300
    cg_preconditioner: int = 1
    cg_conv_test: int = 2
    cg_max_iter: int = 100000
    dimensional_mode: int = 0
    material_properties_enabled: int = 0
    inlet_temperature: float = 568.15
    outlet_pressure_gauge: float = 0.0
    # mass-flow inlet, BC 511 (MeshIO.cpp:190-242; Preprocess.cpp:1054-1112)
    mass_flow_inlet_enabled: int = 0
    inlet_mass_flow_rate: float = 0.0
    inlet_density: float = 1.0
    inlet_u: float = 0.0
    inlet_v: float = 0.0
    inlet_w: float = 0.0
    heat_flux_bc: float = 0.0
    transient_on: int = 0
    # output controls (.par line 44, MeshIO.cpp:321-356)
    residual_log_every: int = 1
    vtu_output_enabled: int = 0
    vtu_output_every_iterations: int = 100
    vtu_output_every_sim_time: float = 0.0
    write_boundary_debug_arrays: int = 0


@dataclasses.dataclass
class Case:
    """A generated case: mesh, boundary triangles, BC map, materials, controls."""

    name: str
    nx: int
    ny: int
    nz: int
    coord: np.ndarray        # (npoin, 3) float64
    intma: np.ndarray        # (nelem, 4) int32, 1-based node ids
    bface: np.ndarray        # (nboun, 5) int32: n1 n2 n3 parent raw_side_id
    side_to_bc: Dict[int, int]
    params: CaseParams
    # Optional single fluid material for dimensional mode (rho, cp, k, mu, Qvol).
    material: Optional[tuple] = None

    @property
    def nelem(self) -> int:
        return int(self.intma.shape[0])

    @property
    def npoin(self) -> int:
        return int(self.coord.shape[0])

    @property
    def nboun(self) -> int:
        return int(self.bface.shape[0])

    def solver_bc(self) -> np.ndarray:
        """Per-face solver BC id after the .bco mapping (MeshIO.cpp:963-981)."""
        table = np.zeros(7, dtype=np.int32)
        for side, bc in self.side_to_bc.items():
            table[side] = bc
        return table[self.bface[:, 4]]


def _kuhn_connectivity(nx: int, ny: int, nz: int) -> np.ndarray:
    ex, ey, ez = 1, nx + 1, (nx + 1) * (ny + 1)
    step = {"x": ex, "y": ey, "z": ez}
    perms = ["xyz", "xzy", "yxz", "yzx", "zxy", "zyx"]
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # hex-major order with x fastest
    base = (1 + i + ey * j + ez * k).transpose(2, 1, 0).reshape(-1).astype(np.int64)
    tets = np.empty((base.size, 6, 4), dtype=np.int64)
    for t, p in enumerate(perms):
        n0 = base
        n1 = n0 + step[p[0]]
        n2 = n1 + step[p[1]]
        n3 = n2 + step[p[2]]
        odd = p in ("xzy", "yxz", "zyx")
        if odd:  # swap local nodes 2 and 3 to make the volume positive
            tets[:, t, :] = np.stack([n0, n2, n1, n3], axis=1)
        else:
            tets[:, t, :] = np.stack([n0, n1, n2, n3], axis=1)
    return tets.reshape(-1, 4)


def _boundary_faces(intma: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    ex, ey = 1, nx + 1
    ez = (nx + 1) * (ny + 1)
    node0 = intma - 1
    gi = node0 % ey
    gj = (node0 // ey) % (ny + 1)
    gk = node0 // ez
    records = []
    for f in range(4):
        loc = IPPN1[f] - 1
        fi, fj, fk = gi[:, loc], gj[:, loc], gk[:, loc]
        conds = [
            (fi == 0).all(axis=1), (fi == nx).all(axis=1),
            (fj == 0).all(axis=1), (fj == ny).all(axis=1),
            (fk == 0).all(axis=1), (fk == nz).all(axis=1),
        ]
        side = np.zeros(intma.shape[0], dtype=np.int64)
        for s, c in enumerate(conds, start=1):
            side[c] = s
        el = np.nonzero(side)[0]
        if el.size:
            nodes = intma[el][:, loc]
            records.append(np.column_stack([el, np.full(el.size, f), nodes, el + 1, side[el]]))
    rec = np.concatenate(records, axis=0)
    order = np.lexsort((rec[:, 1], rec[:, 0]))  # by parent element, then local face
    rec = rec[order]
    return rec[:, 2:].astype(np.int32)


def build_box(name: str, nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0),
              jitter: float = 0.0, seed: int = 0) -> tuple:
    """Return (coord, intma, bface) for the Kuhn-split box."""
    lx, ly, lz = lengths
    i = np.arange(nx + 1, dtype=np.float64)
    j = np.arange(ny + 1, dtype=np.float64)
    k = np.arange(nz + 1, dtype=np.float64)
    xs, ys, zs = (i * lx) / nx, (j * ly) / ny, (k * lz) / nz
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coord = np.column_stack([xs[I.ravel()], ys[J.ravel()], zs[K.ravel()]])
    if jitter > 0.0:
        interior = ((I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)).ravel()
        idx = np.nonzero(interior)[0]
        u = uniform01(seed, 3 * idx.size).reshape(idx.size, 3)
        h = np.array([lx / nx, ly / ny, lz / nz])
        coord[idx] += (2.0 * u - 1.0) * (jitter * h)
    intma = _kuhn_connectivity(nx, ny, nz)
    bface = _boundary_faces(intma, nx, ny, nz)
    return coord, intma.astype(np.int32), bface


def cavity(n: int, re: float = 100.0, jitter: float = 0.0, seed: int = 0,
           ntime: int = 200, name: Optional[str] = None) -> Case:
    """Lid-driven cavity: lid ymax -> BC 500 with u=(1,0,0), other faces -> 530."""
    coord, intma, bface = build_box(name or f"cavity{n}", n, n, n, jitter=jitter, seed=seed)
    params = CaseParams(ntime=ntime, re=re, inlet_u=1.0)
    sides = {1: BC_WALL_ADIABATIC, 2: BC_WALL_ADIABATIC, 3: BC_WALL_ADIABATIC,
             4: BC_LID, 5: BC_WALL_ADIABATIC, 6: BC_WALL_ADIABATIC}
    return Case(name or f"cavity{n}", n, n, n, coord, intma, bface, sides, params)


def channel(nx: int, ny: int, nz: int, re: float = 200.0, pr: float = 0.71,
            jitter: float = 0.0, seed: int = 1, ntime: int = 500, t_init: float = 1.0,
            name: Optional[str] = None) -> Case:
    """Channel [0,4]x[0,1]^2: inflow 508 at x=0, outlet 520 at x=4, walls 502.

    Dimensional single-fluid material mode with rho=cp=1, k=1/(Re*Pr), mu=1/Re,
    so the Step 4 diffusion uses k and the time step uses max(alpha, nu)
    (src/timestep/TimeStep.cpp:213-252).  The reference has no "T=1 inflow"
    condition (508 forces T=0: src/boundary/Boundary.cpp:569-572), so the
    scalar starts at T=t_init everywhere and is cooled by the walls and inlet.
    """
    coord, intma, bface = build_box(name or f"channel{nx}", nx, ny, nz,
                                    lengths=(4.0, 1.0, 1.0), jitter=jitter, seed=seed)
    params = CaseParams(ntime=ntime, re=re, pr=pr, temp_calc=1,
                        cinf=(0.0, 0.0, 0.0, 0.0, t_init),
                        dimensional_mode=1, material_properties_enabled=1)
    sides = {1: BC_PARABOLIC_INLET, 2: BC_OUTLET, 3: BC_WALL_T0,
             4: BC_WALL_T0, 5: BC_WALL_T0, 6: BC_WALL_T0}
    material = (1.0, 1.0, 1.0 / (re * pr), 1.0 / re, 0.0)
    return Case(name or f"channel{nx}", nx, ny, nz, coord, intma, bface, sides, params,
                material=material)


def adjoint_channel(nx: int = 64, ny: int = 16, nz: int = 16, re: float = 100.0, ntime: int = 100,
                    dtfix: float = 0.01, rel_toler: float = 1.0e-12, name: Optional[str] = None) -> Case:
    """The channel geometry with a fixed time step, for the adjoint workload:
    dtfixed = 1 (the minimum over elements of the stability time step is not
    differentiable), dtfix given (the adjoint workload uses the stability
    time step of the channel's first step), every pressure solve to
    rel_toler, and no energy equation (the temperature does not feed back
    into the velocity, so it cannot enter the velocity objective)."""
    case = channel(nx, ny, nz, re=re, pr=0.71, jitter=0.0, seed=4, ntime=ntime,
                   name=name or f"adjoint_channel{nx}")
    case.params.dtfixed = 1
    case.params.dtfix = dtfix
    case.params.temp_calc = 0
    case.params.rel_toler = rel_toler
    return case


# ---------------------------------------------------------------------------
# Reference file writers
# ---------------------------------------------------------------------------

def _fmt(v: float) -> str:
    return repr(float(v))


MESH_MAGIC = b"CBSMESH1"


def mesh_checksum(parts) -> int:
    """Checksum of a binary mesh's arrays (include/cbs_case.h): the sum of
    their concatenated bytes read as little-endian 64-bit words (the last one
    zero-padded), word i multiplied by 2i + 1, modulo 2^64."""
    buf = b"".join(np.ascontiguousarray(p).tobytes() for p in parts)
    if len(buf) % 8:
        buf += bytes(8 - len(buf) % 8)
    w = np.frombuffer(buf, "<u8")
    mult = 2 * np.arange(w.size, dtype=np.uint64) + np.uint64(1)
    return int(np.sum(w * mult, dtype=np.uint64))


def write_mesh_binary(path: str, coord, intma, bface, mat=None) -> None:
    """<base>.cbsmesh: the .plt content (and material ids) as raw arrays."""
    coord = np.ascontiguousarray(coord, "<f8")
    intma = np.ascontiguousarray(intma, "<i4")
    bface = np.ascontiguousarray(bface, "<i4")
    parts = [coord, intma, bface] + ([np.ascontiguousarray(mat, "<i4")] if mat is not None else [])
    head = MESH_MAGIC + np.array([1, 1 if mat is not None else 0], "<u4").tobytes() + \
        np.array([intma.size // 4, coord.size // 3, bface.size // 5, 0], "<i4").tobytes()
    with open(path + ".part", "wb") as f:
        f.write(head)
        for a in parts:
            f.write(a.tobytes())
        f.write(np.array([mesh_checksum(parts)], "<u8").tobytes())
    os.replace(path + ".part", path)


def write_case_files(case: Case, directory: str, binary: bool = False) -> str:
    """Write <dir>/<name>.{plt,bco,par[,material,matprop]}; return the base path.
    binary: the mesh (and material ids) go to <name>.cbsmesh instead of
    <name>.plt and <name>.material."""
    os.makedirs(directory, exist_ok=True)
    base = os.path.join(directory, case.name)
    if binary:
        mat = None
        if case.params.material_properties_enabled:
            mat = getattr(case, "mat", np.zeros(case.nelem, np.int32))
        write_mesh_binary(base + ".cbsmesh", case.coord, case.intma, case.bface, mat)
    else:
        with open(base + ".plt", "w") as f:
            f.write(f"{case.nelem:10d}{case.npoin:10d}{case.nboun:10d}\n")
            ids = np.arange(1, case.nelem + 1)
            np.savetxt(f, np.column_stack([ids, case.intma]), fmt="%10d", delimiter="")
            nids = np.arange(1, case.npoin + 1)
            for n, (x, y, z) in zip(nids, case.coord):
                f.write(f"{n:10d} {x:.16e} {y:.16e} {z:.16e}\n")
            np.savetxt(f, case.bface, fmt="%10d", delimiter="")
    sides = sorted(set(int(s) for s in case.bface[:, 4]))
    with open(base + ".bco", "w") as f:
        f.write(f"{len(sides):10d}{0:10d}\n")
        for s in sides:
            f.write(f"{s:10d}{case.side_to_bc[s]:10d}\n")
    p = case.params
    lines = [
        ("SOLVER_OPT", "1"),
        ("RESTART_OPT", "0"),
        ("TEMP_CALC", str(p.temp_calc)),
        ("UX UY UZ P T", " ".join(_fmt(v) for v in p.cinf)),
        ("NTIME TRANSIENT_ON DTFIXED DTFIX IWRITE", f"{p.ntime} {p.transient_on} {p.dtfixed} {_fmt(p.dtfix)} 1"),
        ("CSAFM THETA1", f"{_fmt(p.csafm)} 0.5"),
        ("CBS_SCHEME ILOTS HTYPE STEP2_CHECK REM_DELTP BETA_OPT DTFIX_END CSAFM2 EPSILON1 DELTR",
         f"1 {p.ilots} {p.htype} 0 {p.rem_deltp} {p.beta_opt} {p.dtfix_end} 2.0 1e-06 {_fmt(p.deltr)}"),
        ("RE PR RA RI", f"{_fmt(p.re)} {_fmt(p.pr)} 0.0 0.0"),
        ("CONVECTION_TYPE", "0"),
        ("PNODE", str(p.pnode)),
        ("WRITE_OUTPUT WRITE_TIME_OUTPUT TIME_OUTPUT_INTERVAL END_RTIME", "0 0.0 0.0 0.0"),
        ("RELTOLER ABSTOLER", f"{_fmt(p.rel_toler)} {_fmt(p.abs_toler)}"),
        # vel_check=1 with a tolerance no residual reaches: run all ntime steps
        ("VEL_CHECK L2NORM_VEL_TOL TEMP_CHECK L2NORM_TEMP_TOL", "1 1e-300 0 0.0"),
        ("PARAVIEW TECPLOT NUSSELT_CALC NUSSELT_FLAG", "0 0 0 0"),
        ("RUNTIME_MOD", "0"),
        ("ALPHA_SF K_RATIO SOURCE_SOLID HEAT_FLUX_BC", f"1.0 1.0 1.0 {_fmt(p.heat_flux_bc)}"),
        ("DIMENSIONAL_MODE MATERIAL_PROPERTIES_ENABLED MASS_FLOW_INLET_ENABLED",
         f"{p.dimensional_mode} {p.material_properties_enabled} {p.mass_flow_inlet_enabled}"),
        ("P_REF MODEL_DEPTH INLET_MFR INLET_DENSITY INLET_T OUTLET_PGAUGE",
         f"15500000.0 1.0 {_fmt(p.inlet_mass_flow_rate)} {_fmt(p.inlet_density)} "
         f"{_fmt(p.inlet_temperature)} {_fmt(p.outlet_pressure_gauge)}"),
        ("INLET_U INLET_V INLET_W", f"{_fmt(p.inlet_u)} {_fmt(p.inlet_v)} {_fmt(p.inlet_w)}"),
        ("NUSSELT_TINF NUSSELT_TREF NUSSELT_D", "25.0 100.0 1.0"),
        ("CG_PRECONDITIONER CG_CONV_TEST CG_MAX_ITER L2NORM_PRES_TOL",
         f"{p.cg_preconditioner} {p.cg_conv_test} {p.cg_max_iter} 0.0"),
        ("RESLOG RESLOG_EVERY CONSOLE_EVERY LIVE_PLOT VTU VTU_EVERY VTU_SIMTIME BDEBUG STEADY_MIN",
         f"1 {p.residual_log_every} 0 0 {p.vtu_output_enabled} {p.vtu_output_every_iterations} "
         f"{_fmt(p.vtu_output_every_sim_time)} {p.write_boundary_debug_arrays} 1"),
        ("ART_DIFF", "0"),
    ]
    with open(base + ".par", "w") as f:
        for label, data in lines:
            f.write(label + "\n" + data + "\n")
    if p.material_properties_enabled:
        if not binary:
            with open(base + ".material", "w") as f:
                f.write(f"{case.nelem}\n")
                ids = np.arange(1, case.nelem + 1)
                np.savetxt(f, np.column_stack([ids, case.intma, np.zeros(case.nelem, dtype=np.int64)]),
                           fmt="%d", delimiter=" ")
        rho, cp, k, mu, q = case.material
        with open(base + ".matprop", "w") as f:
            f.write("1\n")
            f.write(f"0 fluid fluid {_fmt(rho)} {_fmt(cp)} {_fmt(k)} {_fmt(mu)} {_fmt(q)}\n")
    return base


def ensemble_params(count: int = 64, seed: int = 3, nmin: int = 12, nmax: int = 64, re_lo: float = 50.0,
                    re_hi: float = 1000.0) -> List[tuple]:
    """(n, Re) of each member of a cavity batch: n_k = nmin + floor(u_2k (nmax -
    nmin + 1)) cells per side and Re_k = exp(log re_lo + u_2k+1 (log re_hi -
    log re_lo)), with u the SplitMix64 uniform stream of `seed`."""
    u = uniform01(seed, 2 * count)
    out = []
    for k in range(count):
        n = min(nmin + int(np.floor(u[2 * k] * (nmax - nmin + 1))), nmax)
        re = float(np.exp(np.log(re_lo) + u[2 * k + 1] * (np.log(re_hi) - np.log(re_lo))))
        out.append((n, re))
    return out


def ensemble(count: int = 64, seed: int = 3, ntime: int = 200, **kw) -> List[Case]:
    """Batch of lid-driven cavities, the cavity20 recipe with each member's own
    size and Reynolds number (ensemble_params)."""
    return [cavity(n, re=re, ntime=ntime, name=f"ensemble{seed}_{k:02d}_n{n}")
            for k, (n, re) in enumerate(ensemble_params(count, seed, **kw))]


ADJOINT_WORKLOAD = dict(nx=64, ny=16, nz=16, re=100.0, ntime=100, objective_steps=20)


def workload(name: str) -> Case:
    """The five synthetic workloads of the CBS hot path, by name.

    cavity20  : 20^3x6 lid-driven cavity, Re=100, 200 steps (48,000 tets)
    channel128: 128x32x32x6 channel, jitter 0.15h seed 1, Re=200, Pr=0.71, 500 steps
    cavity160 : 160^3x6 cavity, jitter 0.1h seed 2, Re=400, 50 steps (24,576,000 tets)
    channel64 : 64x16x16x6 channel, Re=100, 100 steps (adjoint geometry)
    """
    if name == "cavity20":
        return cavity(20, re=100.0, ntime=200, name=name)
    if name == "channel128":
        return channel(128, 32, 32, re=200.0, pr=0.71, jitter=0.15, seed=1, ntime=500, name=name)
    if name == "cavity160":
        return cavity(160, re=400.0, jitter=0.1, seed=2, ntime=50, name=name)
    if name == "channel64":
        return channel(64, 16, 16, re=100.0, pr=0.71, jitter=0.0, seed=4, ntime=100, name=name)
    raise KeyError(name)
