"""Binary mesh files (<base>.cbsmesh, include/cbs_case.h).

A case read from its binary mesh must set up exactly as the same case read
from the reference's ASCII .plt (and .material): every array of the set-up,
bit for bit.  The file format is written twice, independently: by the C++
converter from the ASCII files and by the generator's numpy writer; the two
files must be identical.  A damaged file (a flipped byte, a truncation, a
wrong magic) is rejected with the reader's message.
"""

import ctypes as C
import os

import numpy as np
import pytest

from cbs3d import CaseSetup, CbsError, cases, host_lib

ARRAYS = ("intma", "iside", "mat_elem", "fedge", "wall_node_list", "bc_list", "color_ptr", "color_elem", "coord",
          "unkno", "pres", "temperature", "rho_e", "k_e", "mu_e", "dNkdx", "detJ", "annxf", "face_norm",
          "elcoe_e", "alen_e", "Mdiag_real", "bc_values")


def _lib():
    lib = host_lib()
    lib.cbs_case_write_mesh_binary.argtypes = [C.c_char_p]
    return lib


def _multi_material():
    """The two-material channel of the golden fixtures."""
    c = cases.channel(8, 4, 4, re=100.0, pr=0.71, jitter=0.0, name="bcht")
    hexid = np.arange(c.nelem) // 6
    hx, hy = hexid % 8, (hexid // 8) % 4
    c.mat = np.where((hx >= 3) & (hx <= 4) & (hy <= 1), 1, 0).astype(np.int32)
    c.materials = [(0, "fluid", "fluid", 1.0, 1.0, 1.0 / 71.0, 0.01, 0.0),
                   (1, "steel", "solid", 2.0, 0.5, 0.05, 0.0, 0.3)]
    return c


def _write(case, directory, binary):
    base = cases.write_case_files(case, directory, binary=binary)
    if hasattr(case, "materials"):
        with open(base + ".matprop", "w") as f:
            f.write(f"{len(case.materials)}\n")
            for m in case.materials:
                f.write(" ".join(str(v) if isinstance(v, str) else repr(v) for v in m) + "\n")
        if not binary:
            with open(base + ".material", "w") as f:
                f.write(f"{case.nelem}\n")
                for e in range(case.nelem):
                    n = case.intma[e]
                    f.write(f"{e + 1} {n[0]} {n[1]} {n[2]} {n[3]} {case.mat[e]}\n")
    return base


@pytest.mark.parametrize("make", [lambda: cases.cavity(6, re=400.0, jitter=0.15, seed=7, name="bcav"),
                                  lambda: cases.channel(8, 4, 4, re=200.0, jitter=0.15, seed=1, name="bchan"),
                                  _multi_material], ids=["cavity", "channel", "two_materials"])
def test_binary_mesh_sets_up_like_ascii(make, tmp_path):
    case = make()
    ascii_dir, bin_dir = str(tmp_path / "ascii"), str(tmp_path / "bin")
    a = _write(case, ascii_dir, binary=False)
    b = _write(case, bin_dir, binary=True)
    assert not os.path.exists(b + ".plt")
    # the converter's file and the generator's are the same bytes
    assert _lib().cbs_case_write_mesh_binary(a.encode()) == 0
    with open(a + ".cbsmesh", "rb") as f1, open(b + ".cbsmesh", "rb") as f2:
        assert f1.read() == f2.read()
    os.remove(a + ".cbsmesh")
    sa, sb = CaseSetup.read(a), CaseSetup.read(b)
    for name in ARRAYS:
        np.testing.assert_array_equal(sb.array(name), sa.array(name), err_msg=name)


def test_converter_refuses_to_overwrite(tmp_path):
    base = _write(cases.cavity(3, name="bow"), str(tmp_path), binary=False)
    lib = _lib()
    assert lib.cbs_case_write_mesh_binary(base.encode()) == 0
    assert lib.cbs_case_write_mesh_binary(base.encode()) != 0
    assert b"exists" in lib.cbs_case_last_error()


@pytest.mark.parametrize("damage,message", [("flip", "checksum mismatch"), ("truncate", "truncated"),
                                            ("magic", "not a version-1 binary mesh")])
def test_damaged_binary_mesh_is_rejected(damage, message, tmp_path):
    base = _write(cases.cavity(4, jitter=0.1, seed=2, name="bdm"), str(tmp_path), binary=True)
    path = base + ".cbsmesh"
    data = bytearray(open(path, "rb").read())
    if damage == "flip":
        data[len(data) // 2] ^= 0x10
    elif damage == "truncate":
        data = data[:len(data) - 40]
    else:
        data[0:8] = b"CBSMESH2"
    with open(path, "wb") as f:
        f.write(bytes(data))
    with pytest.raises(CbsError, match=message):
        CaseSetup.read(base)


def test_checksum_covers_odd_word_boundaries(tmp_path):
    """An odd number of 4-byte words before a double array and at the end:
    the checksum reads the bytes as one stream, the last word zero-padded
    (checked against a direct sum in Python integers)."""
    parts = [np.arange(3, dtype="<i4"), np.arange(5, dtype="<f8") + 0.5, np.arange(2, dtype="<i4") - 7]
    joined = b"".join(p.tobytes() for p in parts)
    joined += bytes((-len(joined)) % 8)  # the last 4-byte word padded with zeros
    w = np.frombuffer(joined, "<u8").astype(object)
    want = sum(int(x) * (2 * i + 1) for i, x in enumerate(w)) % (1 << 64)
    assert cases.mesh_checksum(parts) == want


def test_cpp_and_numpy_writers_agree_on_odd_counts(tmp_path):
    """Odd numbers of 4-byte words inside the arrays (3 elements, 3 faces):
    the C++ writer's checksum carries across array boundaries like the
    numpy one."""
    rng = np.random.default_rng(1)
    coord = rng.standard_normal((7, 3))
    intma = rng.integers(1, 8, (3, 4)).astype(np.int32)
    bface = rng.integers(1, 8, (3, 5)).astype(np.int32)
    mat = np.array([0, 1, 0], np.int32)
    for m in (None, mat):
        p1, p2 = str(tmp_path / "c.cbsmesh"), str(tmp_path / "n.cbsmesh")
        lib = host_lib()
        PI, PD = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        lib.cbs_case_write_mesh_arrays_binary.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, PD, PI, PI,
                                                          PI]
        cc, ci, cb = (np.ascontiguousarray(a) for a in (coord, intma, bface))
        rc = lib.cbs_case_write_mesh_arrays_binary(p1.encode(), 3, 7, 3, cc.ctypes.data_as(PD), ci.ctypes.data_as(PI),
                                                   cb.ctypes.data_as(PI),
                                                   None if m is None else m.ctypes.data_as(PI))
        assert rc == 0
        cases.write_mesh_binary(p2, coord, intma, bface, m)
        with open(p1, "rb") as f1, open(p2, "rb") as f2:
            assert f1.read() == f2.read()


def test_stale_binary_mesh_is_rejected(tmp_path):
    """A .cbsmesh converted before its .plt (or .material) was edited would set
    up the old mesh: the reader refuses it and names the newer file."""
    base = _write(cases.cavity(3, name="bst"), str(tmp_path), binary=False)
    assert _lib().cbs_case_write_mesh_binary(base.encode()) == 0
    CaseSetup.read(base)
    t = os.stat(base + ".cbsmesh").st_mtime_ns
    os.utime(base + ".plt", ns=(t + 10**9, t + 10**9))
    with pytest.raises(CbsError, match="older than .*bst.plt"):
        CaseSetup.read(base)
