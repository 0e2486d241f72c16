"""The reference-side binding shown in INTEGRATION.md compiles and links.

The C++ block headed "src/hip/CbsHipBinding.cpp" is extracted from
INTEGRATION.md as written, compiled with the reference's own headers
(include/cbs/..., C++20 like the reference's build) and linked into a shared
object against libcbs3d_hip.so, so that every field it copies between
CBSStateSI, ConjugateGradient::Result and the C ABI exists with a compatible
type.  Nothing runs.  Skipped where the reference tree is absent.
"""

import os
import re
import subprocess

import pytest

from helpers import ROOT

REF_INCLUDE = "/root/reference/include"
HIP_LIB_DIR = os.path.join(ROOT, "cbs-3d_amd", "cbs3d")


def binding_source() -> str:
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        text = f.read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    src = [b for b in blocks if b.startswith("// src/hip/CbsHipBinding.cpp")]
    assert len(src) == 1, "INTEGRATION.md must hold exactly one CbsHipBinding.cpp block"
    return src[0]


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference headers absent")
def test_binding_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "CbsHipBinding.cpp"
    src.write_text(binding_source())
    out = tmp_path / "libbinding.so"
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror", "-fPIC", "-shared",
           "-I", REF_INCLUDE, "-I", os.path.join(ROOT, "include"), str(src),
           "-L", HIP_LIB_DIR, "-lcbs3d_hip", "-L", "/opt/rocm/lib", "-Wl,--no-undefined",
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    syms = subprocess.run(["nm", "-D", "--defined-only", str(out)], capture_output=True, text=True).stdout
    assert "solvePressure" in syms and "create" in syms
