"""Device-side set-up: Coloring::build on the GPU (cbs_hip_coloring).

The device colours elements level by level (an element once all its lower
neighbours are coloured), which yields the reference's greedy first-fit
colouring exactly.  Checked bit for bit against:
  - the reference's own colour arrays in the golden fixtures (five meshes:
    cavities, a jittered channel, the mixed-BC box, the two-material
    channel);
  - the host colouring (itself pinned against the reference) on a jittered
    cavity and on a channel whose elements are renumbered at random, which
    changes the dependency structure the levels follow;
and through CaseSetup(device_coloring=True), whose colour arrays must equal
the host set-up's.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["cav4", "cav6j", "chan8", "mixbc", "cht8"])
def test_device_coloring_matches_reference(name):
    from cbs3d import device_coloring_of
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    nc, ptr, elem, levels = device_coloring_of(d["intma"], int(d["npoin"][0]))
    assert nc == int(d["ncolor"][0])
    np.testing.assert_array_equal(ptr, d["color_ptr"])
    np.testing.assert_array_equal(elem, d["color_elem"])
    assert levels >= nc


def _host_coloring(case):
    from cbs3d import CaseSetup
    s = CaseSetup.from_case(case)
    return s.mesh_desc().ncolor, s.array("color_ptr"), s.array("color_elem")


def _renumbered(case, seed):
    perm = np.random.default_rng(seed).permutation(case.nelem)
    case.intma = np.ascontiguousarray(case.intma[perm])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(case.nelem)
    b = case.bface.copy()
    b[:, 3] = inv[b[:, 3] - 1] + 1  # parent element ids follow the renumbering
    case.bface = b
    return case


@pytest.mark.parametrize("kind", ["cavity", "renumbered_channel"])
def test_device_coloring_matches_host(kind):
    from cbs3d import cases, device_coloring_of
    if kind == "cavity":
        case = cases.cavity(24, re=100.0, jitter=0.1, seed=5, name="dc24")
    else:
        case = _renumbered(cases.channel(16, 8, 8, re=100.0, pr=0.71, jitter=0.1, seed=2, name="dcr"), 11)
    nc_h, ptr_h, elem_h = _host_coloring(case)
    nc, ptr, elem, _ = device_coloring_of(case.intma.reshape(-1), case.npoin)
    assert nc == nc_h
    np.testing.assert_array_equal(ptr, ptr_h[:nc + 1])
    np.testing.assert_array_equal(elem, elem_h)


def test_case_setup_with_device_coloring():
    from cbs3d import CaseSetup, cases
    case = cases.channel(8, 4, 4, re=200.0, pr=0.71, jitter=0.15, seed=1, name="dcs")
    host = CaseSetup.from_case(case)
    dev = CaseSetup.from_case(case, device_coloring=True)
    assert dev.mesh_desc().ncolor == host.mesh_desc().ncolor
    for name in ("color_ptr", "color_elem", "dNkdx", "annxf", "elcoe_e", "fedge"):
        np.testing.assert_array_equal(dev.array(name), host.array(name), err_msg=name)


def test_device_coloring_rejects_bad_connectivity():
    from cbs3d import CbsError, device_coloring_of
    intma = np.array([[1, 2, 3, 4], [2, 3, 4, 9]], np.int32)
    with pytest.raises(CbsError, match="connectivity node out of range"):
        device_coloring_of(intma, 5)
