"""Host checks of the PCG's assembled operator (cbs-3d_amd/csrc/hip/k_sell.hip).

tests/native/sell_layout_check.hip links the library's own layout builder
(build_sell_host) into a host program that replays the device kernels'
indexing with bounds checks and compares the apply with an element-by-element
apply of the same coefficients.  Cases cover tiles split by a small LDS
budget, one tile holding the whole mesh, and two segments."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "sell_layout_check.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("sell") / "sell_layout_check")
    subprocess.run([HIPCC if os.path.exists(HIPCC) else "hipcc", "-O2", "-std=c++20", "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", out], check=True, capture_output=True)
    return out


@pytest.mark.parametrize("n,chunk,lds,nseg", [(7, 150, 3200, 1), (9, 400, 2400, 1), (5, 5000, 65536, 1),
                                              (12, 900, 4000, 1), (6, 200, 2800, 2)])
def test_layout_and_apply(checker, n, chunk, lds, nseg):
    r = subprocess.run([checker, str(n), str(chunk), str(lds), str(nseg)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
