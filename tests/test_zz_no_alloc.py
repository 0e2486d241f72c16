"""The colouring entry point that allocates nothing (cbs_hip_coloring_noalloc)
against cbs_hip_coloring, through the C ABI.

Two calls in a row on the same GPU, the second on a smaller mesh, so the
second runs in a workspace that still holds the first call's contents; both
must give the host entry point's colour arrays exactly (integer data).  A
mesh beyond the fixed workspace is rejected before any device work.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from cbs3d import CBS_OK, cases, device_coloring_of  # noqa: E402
from cbs3d._native import hip_lib  # noqa: E402

MAX_COLOURS = 256
CBS_ERR_BAD_ARGUMENT = 1


def log(msg):
    print(msg, flush=True)


def hip_runtime():
    """The HIP runtime libcbs3d_hip.so runs on (loaded first, so the soname
    resolves to that instance): the test's device buffers come from it."""
    hip_lib()
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, rt, host):
        self.rt, self.host = rt, np.ascontiguousarray(host)
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(self.host.nbytes, 4)) == 0
        assert rt.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        assert self.rt.hipFree(self.p) == 0


def noalloc_coloring(intma, npoin):
    lib, rt = hip_lib(), hip_runtime()
    nelem = intma.shape[0]
    d_intma = DeviceArray(rt, np.ascontiguousarray(intma, np.int32).reshape(-1))
    d_ptr = DeviceArray(rt, np.full(MAX_COLOURS + 1, -7, np.int32))
    d_elem = DeviceArray(rt, np.full(nelem, -7, np.int32))
    try:
        nc, lv = C.c_int32(), C.c_int32()
        rc = lib.cbs_hip_coloring_noalloc(nelem, npoin, d_intma.p, C.byref(nc), d_ptr.p, d_elem.p, C.byref(lv))
        assert rc == CBS_OK, lib.cbs_hip_last_error(None).decode()
        return nc.value, d_ptr.get(), d_elem.get(), lv.value
    finally:
        for a in (d_intma, d_ptr, d_elem):
            a.free()


def shuffled_channel():
    """A jittered channel with its elements renumbered at random: another
    dependency structure for the first-fit order than the hex-major mesh."""
    c = cases.channel(24, 8, 8, re=200.0, jitter=0.15, seed=1, name="chan24")
    perm = np.random.default_rng(17).permutation(c.nelem)
    return c.intma[perm], c.npoin


def test_noalloc_coloring_matches_coloring_twice():
    lib = hip_lib()
    log(f"[no-alloc] workspace {lib.cbs_hip_coloring_noalloc_scratch_bytes()} bytes")
    big = shuffled_channel()
    small = cases.cavity(6, re=400.0, jitter=0.15, seed=7, name="cav6j")
    for name, (intma, npoin) in (("chan24 shuffled", big), ("cav6j", (small.intma, small.npoin))):
        log(f"[no-alloc] {name}: {intma.shape[0]} elements, {npoin} nodes")
        nc, ptr, elem, lv = noalloc_coloring(intma, npoin)
        want_nc, want_ptr, want_elem, want_lv = device_coloring_of(intma.reshape(-1), npoin)
        log(f"[no-alloc] {name}: {nc} colours, {lv} levels (cbs_hip_coloring: {want_nc}, {want_lv})")
        assert nc == want_nc and lv == want_lv
        assert np.array_equal(ptr[:nc + 1], want_ptr)
        assert np.all(ptr[nc + 1:] == intma.shape[0])
        assert np.array_equal(elem, want_elem)


def test_noalloc_coloring_rejects_meshes_beyond_its_workspace():
    lib, rt = hip_lib(), hip_runtime()
    d = DeviceArray(rt, np.zeros(4, np.int32))
    try:
        nc = C.c_int32()
        for nelem, npoin in (((1 << 20) + 1, 8), (8, (1 << 19) + 1)):
            rc = lib.cbs_hip_coloring_noalloc(nelem, npoin, d.p, C.byref(nc), d.p, d.p, None)
            assert rc == CBS_ERR_BAD_ARGUMENT
            assert "fixed workspace" in lib.cbs_hip_last_error(None).decode()
    finally:
        d.free()
