"""Builds the native libraries of the CBS hot path in-tree.

    python cbs-3d_amd/build.py            # host set-up + HIP library (gfx950)

  cbs3d/libcbs3d_host.so  C++20 + OpenMP host set-up and solution output
                          (include/cbs_case.h, include/cbs_post.h)
  cbs3d/libcbs3d_hip.so   HIP kernels for gfx950 + C ABI  (include/cbs_hip.h)

HIP code is compiled with -ffp-contract=off (no multiply-add fusion: the
reference's x86-64 arithmetic has none, and bit-exact parity depends on it)
and with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "cbs3d")

HOST_SOURCES = [os.path.join(CSRC, "host", f) for f in ("case_setup.cpp", "post_output.cpp")]
HIP_SOURCES = [os.path.join(CSRC, "hip", f) for f in ("ctx.hip", "k_pressure.hip", "k_steps.hip", "k_multi.hip", "k_adjoint.hip", "k_sell.hip", "k_amg.hip", "k_setup.hip")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host() -> str:
    out = os.path.join(OUT, "libcbs3d_host.so")
    run(["g++", "-std=c++20", "-O2", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
         "-Wall", "-Wextra", "-I", INCLUDE, *HOST_SOURCES, "-o", out])
    return out


def build_hip() -> str:
    out = os.path.join(OUT, "libcbs3d_hip.so")
    run([HIPCC, "--offload-arch=gfx950", "-std=c++20", "-O3", "-fPIC", "-shared",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-I", INCLUDE, *HIP_SOURCES, "-o", out])
    return out


def main(argv) -> int:
    what = argv[1:] or ["host", "hip"]
    if "host" in what:
        build_host()
    if "hip" in what:
        build_hip()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
