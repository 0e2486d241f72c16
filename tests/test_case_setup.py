"""Host set-up (include/cbs_case.h) against the reference's own initialisation.

The C++ restatement of MeshIO::initialiseFields, Preprocess::* and
Coloring::build must reproduce the reference's arrays bit for bit on every
golden case, from in-memory data and from the reference's case files.
"""

import os
import sys

import numpy as np
import pytest

from helpers import GOLDEN, GOLDEN_CASES, Golden, assert_bits
import oracle as orc

sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

from cbs3d import CaseSetup, CbsError, cases  # noqa: E402

CASES = {c.name: c for c, _, _ in make_golden.golden_cases()}

GEOMETRY = ["dNkdx", "detJ", "annxf", "face_norm", "elcoe_e", "alen_e", "Mdiag_real"]
TOPOLOGY = ["intma", "iside", "fedge", "mat_elem", "wall_node_list", "bc_list", "color_ptr", "color_elem"]
MATERIAL = ["rho_e", "mu_e", "k_e", "rho_cp_e", "alpha_e", "Qvol_e", "bc_values", "coord"]


@pytest.fixture(scope="module", params=GOLDEN_CASES)
def pair(request):
    g = Golden(request.param)
    return g, CaseSetup.from_case(CASES[request.param])


def test_geometry_bitwise(pair):
    g, c = pair
    for k in GEOMETRY + MATERIAL:
        assert_bits(k, c.array(k), g[k])


def test_topology_and_colouring_exact(pair):
    g, c = pair
    for k in TOPOLOGY:
        assert np.array_equal(c.array(k), g[k]), k


def test_config_derivations(pair):
    g, c = pair
    cfg = c.config()
    assert cfg.ani == g.scalar("ani")
    assert cfg.temp_calc == g.scalar("temp_calc")
    d = c.mesh_desc()
    assert d.ncolor == g.scalar("ncolor")
    assert d.npoin_wall == g.scalar("npoin_wall")
    assert d.bc_fixed == g.scalar("bc_fixed")


def test_initial_fields_after_boundary_values(pair):
    """Initial fields + the oracle's Boundary::apply* equal the reference's
    state at the end of Solver::initialise."""
    g, c = pair
    arrays = {k: c.array(k) for k in GEOMETRY + TOPOLOGY + MATERIAL}
    arrays.update(unkno=c.array("unkno"), pres=c.array("pres"), temperature=c.array("temperature"))
    st = orc.OracleState(arrays, g.cfg())
    st.apply_temperature()
    st.apply_velocity()
    st.apply_pressure()
    st.update_velocity_magnitude()
    assert_bits("unkno", st["unkno"], g["init/unkno"])
    assert_bits("pres", st["pres"], g["init/pres"])
    assert_bits("temperature", st["temperature"], g["init/temperature"])
    assert_bits("velocity", st["velocity"], g["init/velocity"])


def test_reads_reference_case_files(tmp_path):
    """cbs_case_read parses the reference's .plt/.bco/.par/.material/.matprop."""
    for name in ("chan8", "cht8", "mixbc"):
        case = CASES[name]
        base = cases.write_case_files(case, str(tmp_path))
        if hasattr(case, "materials"):
            make_golden.write_multi_material(case, base)
        c = CaseSetup.read(base)
        g = Golden(name)
        for k in GEOMETRY + MATERIAL:
            assert_bits(k, c.array(k), g[k])
        for k in TOPOLOGY:
            assert np.array_equal(c.array(k), g[k]), k


def test_inverted_element_is_rejected():
    case = cases.cavity(2)
    case.intma[3, [1, 2]] = case.intma[3, [2, 1]]
    with pytest.raises(CbsError) as e:
        CaseSetup.from_case(case)
    assert "non-positive tetrahedron volume at element 4" in str(e.value)


def test_unsupported_bc_is_rejected():
    case = cases.cavity(2)
    case.side_to_bc[1] = 999
    with pytest.raises(CbsError) as e:
        CaseSetup.from_case(case)
    assert "unsupported BC_ID 999" in str(e.value)
