"""Solution output (include/cbs_post.h, cbs-3d_amd/csrc/host/post_output.cpp)
against the reference's own Post (src/io/Post.cpp).

tests/golden/post_*.npz hold what the reference wrote when oracle/ref_dump
drove its Post routines in Solver::run's order (tests/golden/make_golden.py
--post): the residual CSV, the PVD and the ASCII VTU files, together with the
arrays and scalars they were written from.  The writer here is fed the same
arrays in the same call sequence; then

  - the CSV and PVD files must equal the reference's byte for byte;
  - the same VTU files must exist, each with the reference's arrays in the
    reference's order and types, the same piece sizes and data attributes,
    and every value equal to the reference's text for it: integers exactly,
    doubles as their "%.16g" text (the reference streams setprecision(16)).

The cavity fixture is transient with VTU output every 2 iterations and every
0.17 of simulated time at dt = 0.0884: the reference writes iteration 3
because an iteration-triggered output leaves the time threshold where it was;
the channel fixture has solid nodes (masked pressure and velocity), output by
simulated time only and every other residual row.
"""

import json
import os
import re

import numpy as np
import pytest

from cbs3d import post as cpost
from cbs3d._native import Scalars

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["pcav4", "pcht8"]

_TAG = re.compile(r"<DataArray ([^>]*)>(.*?)</DataArray>", re.S)
_ATTR = re.compile(r'(\w+)="([^"]*)"')


def load(name):
    d = np.load(os.path.join(GOLDEN, f"post_{name}.npz"))
    meta = json.loads(d["meta"].tobytes().decode())
    files = {k[len("file__"):]: d[k].tobytes() for k in d.files if k.startswith("file__")}
    return d, meta, files


def ascii_arrays(text):
    """(section, attributes, tokens) of every DataArray of a reference VTU."""
    out = []
    for m in _TAG.finditer(text):
        attrs = dict(_ATTR.findall(m.group(1)))
        head = text[:m.start()]
        section = max(("Points", "Cells", "PointData", "CellData"), key=lambda s: head.rfind("<" + s))
        out.append((section, attrs, m.group(2).split()))
    return out


def replay(d, meta, root):
    """The reference's call sequence (oracle/ref_dump.cpp in post mode)."""
    p = meta["params"]
    cfg = cpost.config(residual_log_enabled=1, residual_log_every=p["residual_log_every"],
                       vtu_output_enabled=p["vtu_output_enabled"],
                       vtu_output_every_iterations=p["vtu_output_every_iterations"],
                       vtu_output_every_sim_time=p["vtu_output_every_sim_time"], transient_on=p["transient_on"],
                       write_boundary_debug_arrays=p["write_boundary_debug_arrays"])
    mesh = {"coord": d["coord"], "intma": d["intma"], "iside": d["iside"], "mat_elem": d["mat_elem"]}
    w = cpost.Post(mesh, cfg, root=root, case_name=meta["name"])

    def fields(prefix, post_names):
        keys = ("pres", "unkno", "temperature", "velocity")
        src = [f"{prefix}{'post_' if post_names else ''}{k}" for k in keys]
        return {k: d[s] for k, s in zip(keys, src)}

    w.initialiseRunOutputs(float(d["init__rtime"][0]), fields("init__", False))
    n = meta["nsteps"]
    for it in range(1, n + 1):
        s = f"s{it}__"
        sc = Scalars()
        sc.dtreal = float(d[s + "dtreal"][0])
        sc.rtime = float(d[s + "rtime"][0])
        for k in range(15):
            sc.hb[k] = float(d[s + "hb"][k])
        sc.last_cg_iterations = int(d[s + "cg_iterations"][0])
        sc.last_cg_initial_l2 = float(d[s + "cg_initial_l2"][0])
        sc.last_cg_final_l2 = float(d[s + "cg_final_l2"][0])
        sc.last_cg_relative_l2 = float(d[s + "cg_relative_l2"][0])
        sc.last_cg_max_abs = float(d[s + "cg_max_abs"][0])
        w.writeResidualRow(it, sc)
        if w.shouldWriteVTU(it, sc.rtime):
            w.writeSolution(it, sc.rtime, fields(s, True))
    if w.last_output() != n:
        w.writeSolution(n, float(d[f"s{n}__rtime"][0]), fields(f"s{n}__", True))
    w.flush()
    w.close()
    return os.path.join(root, meta["name"])


@pytest.mark.parametrize("name", CASES)
def test_files_match_reference(name, tmp_path):
    d, meta, ref = load(name)
    out = replay(d, meta, str(tmp_path))
    assert sorted(os.listdir(out)) == sorted(ref)
    for fname, want in ref.items():
        got_path = os.path.join(out, fname)
        if not fname.endswith(".vtu"):
            with open(got_path, "rb") as f:
                assert f.read() == want, fname
            continue
        text = want.decode()
        got = cpost.read_vtu(got_path)
        with open(got_path, "rb") as f:
            head = f.read().split(b"<AppendedData", 1)[0].decode()
        for tag in ("Piece", "PointData", "CellData"):
            assert re.search(f"<{tag} [^>]*>", head).group(0) == re.search(f"<{tag} [^>]*>", text).group(0)
        arrays = ascii_arrays(text)
        ours = [m for m in re.finditer(r"<DataArray ([^>]*)/>", head)]
        assert len(ours) == len(arrays), fname
        for (section, attrs, tokens), m in zip(arrays, ours):
            mine = dict(_ATTR.findall(m.group(1)))
            key = attrs.get("Name", "points")
            assert mine.get("Name", "points") == key and mine["type"] == attrs["type"], (fname, key)
            assert mine["NumberOfComponents"] == attrs.get("NumberOfComponents", "1" if key != "points" else "3")
            vals = got[key]
            assert vals.size == len(tokens), (fname, key)
            if attrs["type"] == "Float64":
                text_vals = ["%.16g" % v for v in vals]
                bad = [i for i, (a, b) in enumerate(zip(text_vals, tokens)) if a != b]
                assert not bad, (fname, key, bad[:5], [(text_vals[i], tokens[i]) for i in bad[:5]])
            else:
                np.testing.assert_array_equal(vals, np.array(tokens, dtype=np.int64), err_msg=f"{fname} {key}")


def test_nonfinite_values_are_written_as_zero(tmp_path):
    """Post.cpp:117-122: NaN and infinities become 0 in the CSV and the VTU."""
    d, meta, _ = load("pcav4")
    mesh = {"coord": d["coord"], "intma": d["intma"], "iside": d["iside"], "mat_elem": d["mat_elem"]}
    w = cpost.Post(mesh, cpost.config(vtu_output_every_iterations=1), root=str(tmp_path), case_name="nf")
    n = d["coord"].size // 3
    pres = np.full(n, np.nan)
    pres[1] = np.inf
    f = {"pres": pres, "temperature": np.full(n, -np.inf), "unkno": np.full(3 * n, np.nan),
         "velocity": np.full(n, np.nan)}
    w.initialiseRunOutputs(0.0, f)
    sc = Scalars()
    sc.rtime = float("nan")
    sc.hb[0] = float("inf")
    w.writeResidualRow(1, sc)
    w.close()
    v = cpost.read_vtu(os.path.join(str(tmp_path), "nf", "nf_step_00000000.vtu"))
    for k in ("pressure", "temperature", "velocity", "velocity_magnitude"):
        assert np.all(v[k] == 0.0), k
    with open(os.path.join(str(tmp_path), "nf", "nf_residuals.csv")) as fh:
        row = fh.read().splitlines()[1].split(",")
    assert row[0] == "1" and row[1] == "0" and row[3] == "0"


def test_disabled_outputs_write_nothing(tmp_path):
    d, _, _ = load("pcav4")
    mesh = {"coord": d["coord"], "intma": d["intma"], "iside": d["iside"], "mat_elem": d["mat_elem"]}
    w = cpost.Post(mesh, cpost.config(residual_log_enabled=0, vtu_output_enabled=0), root=str(tmp_path),
                   case_name="off")
    w.initialiseRunOutputs(0.0)
    w.writeResidualRow(1, Scalars())
    assert not w.shouldWriteVTU(0, 0.0)
    w.close()
    assert os.listdir(os.path.join(str(tmp_path), "off")) == []


def test_connectivity_out_of_range_is_the_reference_error(tmp_path):
    d, _, _ = load("pcav4")
    intma = d["intma"].copy()
    intma[5] = d["coord"].size // 3 + 1
    mesh = {"coord": d["coord"], "intma": intma, "iside": d["iside"], "mat_elem": d["mat_elem"]}
    with pytest.raises(cpost.PostError, match="Post::writeVTU - element connectivity node out of range"):
        cpost.Post(mesh, cpost.config(), root=str(tmp_path), case_name="bad")
