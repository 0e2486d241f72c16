"""Shared test helpers: golden fixtures, run-control dictionaries, comparisons."""

from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "cbs-3d_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN_CASES = ["cav4", "cav6j", "chan8", "mixbc", "cht8", "dtfix", "dtend", "lots1", "lots2", "conv1", "conv3",
                "mixbc2", "cht2"]

SETUP_ARRAYS = ["intma", "coord", "mat_elem", "rho_e", "mu_e", "k_e", "rho_cp_e", "alpha_e",
                "Qvol_e", "dNkdx", "detJ", "annxf", "fedge", "iside", "face_norm", "elcoe_e",
                "alen_e", "wall_node_list", "bc_list", "bc_values", "color_ptr", "color_elem",
                "pdiagE", "gstifE"]


class Golden:
    """A golden fixture written by tests/golden/make_golden.py."""

    def __init__(self, name: str):
        self.name = name
        with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
            self.d = {k.replace("__", "/"): z[k] for k in z.files}
        self.meta = json.loads(bytes(self.d.pop("meta")).decode())
        self.nsteps = int(self.meta["nsteps"])

    def __getitem__(self, k):
        return self.d[k]

    def __contains__(self, k):
        return k in self.d

    def scalar(self, k):
        return self.d[k].reshape(-1)[0].item()

    def cfg(self) -> dict:
        """Run controls as the reference holds them after MeshIO::readAll."""
        p = self.meta["params"]
        extra = {}
        if "inlet_u_from_massflow" in self.d:
            extra["inlet_u_from_massflow"] = self.scalar("inlet_u_from_massflow")
        return run_config(p, npoin=self.scalar("npoin"), nelem=self.scalar("nelem"),
                          nboun=self.scalar("nboun"), ncolor=self.scalar("ncolor"),
                          bc_fixed=self.scalar("bc_fixed"), npoin_wall=self.scalar("npoin_wall"),
                          ani=self.scalar("ani"), temp_calc=self.scalar("temp_calc"), **extra)

    def setup_arrays(self) -> dict:
        out = {k: self.d[k] for k in SETUP_ARRAYS}
        out["unkno"] = self.d["init/unkno"]
        out["pres"] = self.d["init/pres"]
        out["temperature"] = self.d["init/temperature"]
        return out


def run_config(p: dict, **kw) -> dict:
    """RunConfig fields of the hot path from CaseParams-style .par values."""
    cfg = dict(
        temp_calc=p["temp_calc"], cbs_scheme=1, ilots=p["ilots"], htype=p["htype"],
        dtfixed=p["dtfixed"], dtfix_end=p.get("dtfix_end", 0), transient_on=0, solver_opt=1,
        rem_deltp=p.get("rem_deltp", 0), beta_opt=p.get("beta_opt", 1), step2_check=0, dimensional_mode=p["dimensional_mode"],
        material_properties_enabled=p["material_properties_enabled"],
        cg_preconditioner=p["cg_preconditioner"], cg_conv_test=p["cg_conv_test"],
        cg_max_iter=p["cg_max_iter"], csafm=p["csafm"], dtfix=p["dtfix"], deltr=p.get("deltr", 1.0e20),
        ani=1.0 / p["re"], relToler=p["rel_toler"], absToler=p["abs_toler"], dtreal=0.0,
        rtime=0.0, csafm2=2.0, epsilon1=1.0e-6, inlet_u=p["inlet_u"], inlet_v=p["inlet_v"],
        inlet_w=p["inlet_w"], inlet_temperature=p["inlet_temperature"],
        inlet_u_from_massflow=0.0, heat_flux_bc=p["heat_flux_bc"])
    cfg.update(kw)
    return cfg


def bits_equal(a, b) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def assert_bits(name, got, want):
    got = np.asarray(got, dtype=np.float64).reshape(-1)
    want = np.asarray(want, dtype=np.float64).reshape(-1)
    assert got.shape == want.shape, f"{name}: shape {got.shape} vs {want.shape}"
    if not bits_equal(got, want):
        bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
        i = bad[0]
        raise AssertionError(f"{name}: {bad.size} of {got.size} values differ in bits; "
                             f"first at {i}: {got[i]!r} vs {want[i]!r}")


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
