"""Host checks of the smoothed-aggregation multigrid set-up
(cbs-3d_amd/csrc/hip/amg_setup.hpp), compiled with g++ into
tests/native/amg_check.cpp: on the P1 Laplacian of jittered Kuhn-tet cubes
with one face of fixed nodes, the aggregates cover the active nodes, every
coarse operator equals the dense P^T A P, the V-cycle is symmetric positive
definite, and PCG with it needs at most a third of Jacobi-PCG's iterations
to a 1e-8 relative residual (a mesh-independent count)."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "amg_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("amg") / "amg_check")
    subprocess.run([gxx, "-O2", "-std=c++20", "-pthread", SRC, "-o", out], check=True, capture_output=True)
    return out


def test_multigrid_setup(checker):
    counts = []
    for n in (12, 20, 32):
        r = subprocess.run([checker, str(n)], capture_output=True, text=True, timeout=300)
        print(r.stdout)
        assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
        counts.append(int(re.search(r"PCG to 1e-8: (\d+) iterations with the V-cycle", r.stdout).group(1)))
    assert max(counts) - min(counts) <= 4, counts  # the iteration count does not grow with the mesh
