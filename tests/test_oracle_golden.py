"""Pins the CPU oracle against the reference's own outputs (golden fixtures).

Every fixture was produced by the compiled reference (tests/golden/make_golden.py)
at one OpenMP thread, so every stage of every step must agree bit for bit:
colouring, element pressure terms, time step, LHS diagonals, Step 1 RHS and
predictor, Step 2 RHS, the Jacobi-PCG solution and its Result fields, Step 3,
Step 4, the velocity magnitude and the 15 convergence norms.
"""

import numpy as np
import pytest

from helpers import GOLDEN_CASES, Golden, assert_bits
import oracle as orc


@pytest.fixture(scope="module", params=GOLDEN_CASES)
def golden(request):
    return Golden(request.param)


def test_coloring_matches_reference(golden):
    nc, ptr, elem = orc.coloring(golden["intma"], golden.scalar("npoin"))
    assert nc == golden.scalar("ncolor")
    assert np.array_equal(ptr, golden["color_ptr"])
    assert np.array_equal(elem, golden["color_elem"])


def test_coloring_is_valid_and_greedy(golden):
    """No two elements of one colour share a node; every element's colour is
    the smallest one absent among its lower-numbered neighbours."""
    intma = golden["intma"].reshape(-1, 4)
    nc, ptr, elem = orc.coloring(intma, golden.scalar("npoin"))
    color = np.empty(intma.shape[0], np.int64)
    for c in range(nc):
        members = elem[ptr[c]:ptr[c + 1]] - 1
        color[members] = c
        nodes = intma[members].reshape(-1)
        assert np.unique(nodes).size == nodes.size
        assert np.all(np.diff(members) > 0)
    node_elems = {}
    for e, row in enumerate(intma):
        for n in row:
            node_elems.setdefault(int(n), []).append(e)
    for e, row in enumerate(intma):
        earlier = {color[f] for n in row for f in node_elems[int(n)] if f < e}
        c = 0
        while c in earlier:
            c += 1
        assert color[e] == c


def test_element_pressure_terms(golden):
    st = orc.OracleState(golden.setup_arrays(), golden.cfg())
    st.build_element_pressure_terms()
    assert_bits("pdiagE", st["pdiagE"], golden["pdiagE"])
    assert_bits("gstifE", st["gstifE"], golden["gstifE"])


def test_apply_matches_reference(golden):
    if "s1/apply_x" not in golden:
        pytest.skip("fixture has no matvec record")
    st = orc.OracleState(golden.setup_arrays(), golden.cfg())
    st["gstif"][:] = golden["s1/gstif"]
    # the recorded matvec ran with the LHS pdiag, before CG imposed constraints
    st["pdiag"][:] = golden["s1/lhs_pdiag"]
    y = st.pressure_multiply(golden["s1/apply_x"])
    assert_bits("apply_y", y, golden["s1/apply_y"])


def test_steps_match_reference_bitwise(golden):
    st = orc.OracleState(golden.setup_arrays(), golden.cfg())
    for it in range(1, golden.nsteps + 1):
        p = f"s{it}/"
        st.begin_step()
        st.compute_timestep(it)
        assert st.dtreal == golden.scalar(p + "dtreal")
        for k in ["delte", "deltp", "deltp1", "deltp2", "beta1", "vvis", "velcp",
                  "unkno_unit_vec", "alen"]:
            assert_bits(p + k, st[k], golden[p + k])
        st.update_lhs_diagonal()
        assert_bits(p + "elcoe2", st["elcoe2"], golden[p + "elcoe2"])
        assert_bits(p + "elcoe2p", st["elcoe2p"], golden[p + "elcoe2p"])
        assert_bits(p + "pdiag", st["pdiag"], golden[p + "lhs_pdiag"])
        assert_bits(p + "gstif", st["gstif"], golden[p + "gstif"])
        st.step1()
        assert_bits(p + "step1_rhs", st["rhs"], golden[p + "step1_rhs"])
        assert_bits(p + "step1_unkno", st["unkno"], golden[p + "step1_unkno"])
        st.assemble_step2_rhs()
        assert_bits(p + "step2_rhs1", st["rhs1"], golden[p + "step2_rhs1"])
        st.step2()
        assert st.last_cg_iterations == golden.scalar(p + "cg_iterations")
        for k in ["initial_l2", "final_l2", "relative_l2", "max_abs"]:
            assert getattr(st, "last_cg_" + k) == golden.scalar(p + "cg_" + k), k
        assert_bits(p + "step2_pres", st["pres"], golden[p + "step2_pres"])
        assert_bits(p + "step2_pdiag_after", st["pdiag"], golden[p + "step2_pdiag_after"])
        st.step3()
        assert_bits(p + "step3_rhs", st["rhs"], golden[p + "step3_rhs"])
        assert_bits(p + "step3_unkno", st["unkno"], golden[p + "step3_unkno"])
        st.step4()
        assert_bits(p + "step4_rhs1", st["rhs1"], golden[p + "step4_rhs1"])
        assert_bits(p + "step4_temperature", st["temperature"], golden[p + "step4_temperature"])
        st.update_velocity_magnitude()
        assert_bits(p + "velocity", st["velocity"], golden[p + "velocity"])
        st.convergence()
        assert_bits(p + "hb", st["hb"], golden[p + "hb"])


def test_advance_one_step_matches_staged_run(golden):
    """orc_advance_one_step is the same sequence as the staged calls."""
    st = orc.OracleState(golden.setup_arrays(), golden.cfg())
    for it in range(1, golden.nsteps + 1):
        st.advance_one_step(it)
    p = f"s{golden.nsteps}/"
    assert_bits("unkno", st["unkno"], golden[p + "step3_unkno"])
    assert_bits("pres", st["pres"], golden[p + "step2_pres"])
    assert_bits("temperature", st["temperature"], golden[p + "step4_temperature"])
    assert_bits("hb", st["hb"], golden[p + "hb"])


@pytest.mark.parametrize("name", ["cav4", "conv1"])
def test_cg_residual_history(name):
    """Per-iteration CG residual norms equal the reference's over 40
    iterations: Jacobi-PCG (cav4) and the unpreconditioned CG
    (cg_preconditioner 0) stopped by max |r| (cg_conv_test 1) (conv1)."""
    g = Golden(name)
    trace = g["s1/cg_trace_l2"]
    st = orc.OracleState(g.setup_arrays(), g.cfg())
    st.begin_step()
    st.compute_timestep(1)
    st.update_lhs_diagonal()
    st.step1()
    st.assemble_step2_rhs()
    saved = {k: st[k].copy() for k in ["pres", "rhs1", "pdiag"]}
    for k in range(1, trace.size + 1):
        for n, v in saved.items():
            st[n][:] = v
        st._s.cg_max_iter = k
        r = st.solve_pressure()
        assert r.final_l2 == trace[k - 1], k
        assert r.final_max_abs == g["s1/cg_trace_max"][k - 1], k


def test_run_control_fixtures_exercise_their_branches():
    """The run-control fixtures reach the branches they are named for, by
    what the reference's definitions fix: a fixed dt gives dtreal = dtfix
    and uniform delte with beta1 = 1 (TimeStep.cpp:492-509), dtfix_end
    switches to the stability step after its last step, ilots 1/2 keep
    non-uniform element steps with ilots 2 capped at deltr, and the conv3
    fixture has a solve stopped by each criterion."""
    g = Golden("dtfix")
    dtfix = g.meta["params"]["dtfix"]
    for it in range(1, g.nsteps + 1):
        assert g.scalar(f"s{it}/dtreal") == dtfix
        assert np.all(g[f"s{it}/delte"] == dtfix) and np.all(g[f"s{it}/beta1"] == 1.0)
    g = Golden("dtend")
    assert g.scalar("s1/dtreal") == g.scalar("s2/dtreal") == g.meta["params"]["dtfix"]
    assert g.scalar("s3/dtreal") != g.meta["params"]["dtfix"]
    for name in ("lots1", "lots2"):
        g = Golden(name)
        de = g["s1/delte"]
        assert de.max() > 2 * de.min()
        assert g.scalar("s1/dtreal") == de.min()
    g = Golden("lots2")
    assert g["s1/delte"].max() == g.meta["params"]["deltr"]
    g = Golden("conv3")
    p = g.meta["params"]
    stops = set()
    for it in range(1, g.nsteps + 1):
        rel_ok = g.scalar(f"s{it}/cg_relative_l2") <= p["rel_toler"]
        abs_ok = g.scalar(f"s{it}/cg_max_abs") <= p["abs_toler"]
        assert rel_ok or abs_ok
        stops.add("rel" if rel_ok and not abs_ok else "abs" if abs_ok and not rel_ok else "both")
    assert {"rel", "abs"} <= stops


def test_mass_flow_inlet_speed_closed_form():
    """U_in = m_dot / (rho_in A_in) over the BC-511 faces' areas
    (Preprocess.cpp:1054-1112), and the 511 nodes carry -U_in n."""
    g = Golden("mixbc2")
    p = g.meta["params"]
    iside = g["iside"].reshape(-1, 6)
    fn = g["face_norm"].reshape(-1, 4)
    area = fn[iside[:, 5] == 511, 3].sum()
    u_in = p["inlet_mass_flow_rate"] / (p["inlet_density"] * area)
    assert abs(g.scalar("inlet_u_from_massflow") - u_in) <= 1e-15 * u_in
    # inlet at x = 0 with outward normal -x: interior nodes of that face move at +U_in in x
    st = orc.OracleState(g.setup_arrays(), g.cfg())
    st.apply_velocity()
    coord = g["coord"].reshape(-1, 3)
    u = st["unkno"].reshape(-1, 3)
    inner = (coord[:, 0] == 0.0) & (coord[:, 1] > 0) & (coord[:, 1] < 1) & (coord[:, 2] > 0) & (coord[:, 2] < 1)
    assert inner.any()
    assert np.allclose(u[inner, 0], g.scalar("inlet_u_from_massflow"), rtol=1e-15, atol=0)
    assert np.all(u[inner, 1:] == 0.0)


@pytest.mark.parametrize("name", ["cav6j", "chan8", "cht2"])
def test_openmp_build_of_the_oracle(name):
    """The CPU baseline's timing variant (the oracle with the reference's
    OpenMP pragmas) gives the reference's bits at one thread, and at four
    threads differs only by the dot products' reduction order."""
    import ctypes

    from helpers import rel_l2

    gomp = ctypes.CDLL("libgomp.so.1")
    saved = gomp.omp_get_max_threads()
    g = Golden(name)
    p = f"s{g.nsteps}/"
    want = {"unkno": g[p + "step3_unkno"], "pres": g[p + "step2_pres"], "temperature": g[p + "step4_temperature"]}
    try:
        for threads in (1, 4):
            gomp.omp_set_num_threads(threads)
            st = orc.OracleState(g.setup_arrays(), g.cfg(), openmp=True)
            for it in range(1, g.nsteps + 1):
                st.advance_one_step(it)
            for k, v in want.items():
                if threads == 1:
                    assert_bits(k, st[k], v)
                else:
                    assert rel_l2(st[k], v) <= 1e-12, (k, rel_l2(st[k], v))
    finally:
        gomp.omp_set_num_threads(saved)
