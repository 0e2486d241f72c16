"""Short fixed runs of the CBS hot path for profilers (rocprofv3 counters).

    python3 cbs-3d_amd/cbs3d/probe.py [--workload cavity160] [--precision mixed|fp64|amg]
                                      [--steps 1] [--copy-bytes 1073741824] [--no-graphs]
                                      [--multi-rhs 0]

Builds the workload's context, advances `steps` CBS steps with the chosen
pressure-solver form (the first solve also runs the apply autotuner; amg is
the multigrid-preconditioned solve, whose first solve also builds the
hierarchy), then runs the plain-copy kernel (k_copy) over `copy-bytes`: its
counters calibrate the HBM-byte counters against a kernel of known traffic
(read + write of copy-bytes each).  Prints one line per stage.
--no-graphs launches the PCG kernels one by one instead of replaying
captured HIP graphs (same kernels; for profilers that do not follow graph
launches).
--multi-rhs N then solves N right-hand sides together (b ~ N(0,1) per node
minus its mean, seed 3) on the last step's operator with the multi-RHS PCG
(fp64, or the mixed form for --precision mixed).
"""

import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from cbs3d import AMG, FP64, MIXED, CaseSetup, HipContext, cases  # noqa: E402


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cavity160")
    ap.add_argument("--precision", choices=("fp64", "mixed", "amg"), default="mixed")
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--multi-rhs", type=int, default=0)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    setup = CaseSetup.from_case(cases.workload(args.workload))
    ctx = HipContext(setup)
    if args.no_graphs:
        ctx.set_graphs(False)
    ctx.initialise()
    ctx.buildElementPressureTerms()
    print(f"[probe] {args.workload}: context ready in {time.perf_counter() - t0:.1f} s", flush=True)
    prec = {"mixed": MIXED, "fp64": FP64, "amg": AMG}[args.precision]
    for it in range(1, args.steps + 1):
        t0 = time.perf_counter()
        r = ctx.advanceOneStep(it, precision=prec)
        ctx.synchronize()
        print(f"[probe] step {it} ({args.precision}): {r.iterations} PCG iterations, "
              f"{time.perf_counter() - t0:.3f} s", flush=True)
    if args.multi_rhs > 0:
        import numpy as np
        B = np.random.default_rng(3).standard_normal((args.multi_rhs, ctx.npoin))
        B -= B.mean(axis=1, keepdims=True)
        mprec = MIXED if args.precision == "mixed" else FP64
        t0 = time.perf_counter()
        _, res = ctx.solvePressureMulti(B, precision=mprec)
        its = [r.iterations for r in res]
        print(f"[probe] {args.multi_rhs}-RHS PCG ({'mixed' if mprec == MIXED else 'fp64'}): "
              f"{time.perf_counter() - t0:.2f} s, iterations {min(its)}..{max(its)}", flush=True)
    if args.copy_bytes > 0:
        gbs = ctx.bandwidth_probe(args.copy_bytes, reps=1)
        print(f"[probe] k_copy of {args.copy_bytes} B (read + write each): {gbs:.0f} GB/s", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
