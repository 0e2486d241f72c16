"""The C-ABI libraries load without a GPU and export every function that
include/cbs_hip.h, include/cbs_case.h and include/cbs_post.h declare (no
device call is made)."""

import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cbs-3d_amd", "cbs3d")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_]\w*[\s\*]+\**\s*(cbs_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("cbs_hip.h", "libcbs3d_hip.so"), ("cbs_case.h", "libcbs3d_host.so"),
                                        ("cbs_post.h", "libcbs3d_host.so")])
def test_every_declared_function_is_exported(header, lib):
    names = declared(header)
    assert len(names) > 5, names
    path = os.path.join(PKG, lib)
    if not os.path.exists(path):
        pytest.fail(f"{lib} is not built: run __graft_entry__.build()")
    so = C.CDLL(path)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
