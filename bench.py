"""Benchmark of the CBS time-step hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cavity160]

Workload (default): the 160^3 x 6 lid-driven cavity (24,576,000 tets,
4,173,281 nodes, interior nodes jittered +-0.1h from seed 2, Re=400), the
configuration the mixed-vs-FP64 target is stated on.  A "step" is one full
CBS step (Solver::advanceOneStep): time step, LHS diagonals, Step 1, Step 2
RHS + Jacobi-PCG to 1e-8 relative residual, Step 3, velocity magnitude and
convergence norms.  All inputs are resident in HBM before the timed region.

Both pressure-solver forms run from the same initial state: FP64 (the
reference algorithm) and mixed (FP32 apply and Jacobi inside FP64 iterative
refinement, same 1e-8 true relative residual).  W untimed warm-up steps are
followed by exactly K timed steps of each form.  `value` is the mixed form's
CBS steps/s; the FP64 figures are reported beside it.

roofline.traffic: the HBM bytes per apply measured with rocprofv3 counters
(profiles/hbm_traffic_<workload>.json, null for workloads not profiled).
roofline: the PCG's operator apply (k_apply_sell: the assembled pressure
operator as sliced ELLPACK, one workgroup per row tile with x staged in LDS;
fused Ap = free ? acc + pdiag p : 0 and (p, Ap) partials), timed with HIP
events on the context's stream inside the timed region; algorithmic bytes
per apply = nnz*(s+2) + N*3s (an s-byte value and a 16-bit column per
off-diagonal nonzero; x, pdiag and y once per node).  Where the context's
first solve found it faster (cbs_hip_direction_fusion), the apply also
forms the iteration's direction p = z + beta p while it stages it; that
kernel's bytes add r read and p written per node (FP64: N*5s in all) and,
in FP32, the inverse diagonal read (N*6s).  The element form the reference
streams, F*(6s+16) + N*3s, is reported beside it
(bytes_per_apply_element_form).
Legs beside the headline (--legs, all by default), each with its own
steps/s (the PCG forms also with their apply roofline):
  amg        the multigrid-preconditioned form (FP64 CG with one FP32
             smoothed-aggregation V-cycle per iteration, build-only) on the
             same workload: CBS steps/s, iterations, the hierarchy's set-up,
             and its fields' distance from the FP64 run;
  multi_rhs  32 right-hand sides solved together on the cavity160 operator,
             FP64 and mixed, against single solves;
  c0         cavity20 (48,000 tets, Re 100): its 200 steps on the GPU (FP64
             and mixed) and on the CPU, and the E2 of the GPU fields after
             200 steps against the CPU's;
  c1         channel128 with energy (786,432 tets): its 500 steps on the GPU,
             the first 3 on the CPU, E2 after those 3;
  ensemble   64 cavities (n in 12..64, Re in [50, 1000]) in one batch
             context, against the same members one context at a time;
  adjoint    the channel64 adjoint (100 steps, objective over the last 20):
             forward and backward times, central differences on a few random
             directions (--adjoint-fd; 32 for the full check).
--setup adds the set-up measured on the same workload: the host set-up's
phases (Solver::initialise: Preprocess and Coloring::build, C++), and
Coloring::build on the device (cbs_hip_coloring, host arrays in and out),
whose colour arrays must equal the host's.
--output K adds the solution output measured on the same workload (mixed
form): CBS steps/s with the residual CSV every step and a VTU every K steps
written asynchronously from device snapshots (include/cbs_post.h), the
snapshot's cost on the compute stream and the writer's throughput.
cpu_baseline: the CPU oracle built with the reference's OpenMP loops
(oracle/libcbs_oracle_omp.so), on all the host CPUs this process may use
(OMP_NUM_THREADS when set) with OMP_PROC_BIND=close and OMP_PLACES=cores,
in a child process that does no GPU work, and the host's CPU model; on cavity160 a bounded sample (the non-solver
stages of one step once plus 8 PCG iterations, scaled to a step with the
GPU's mean iterations per step), on the c0 and c1 legs whole steps.
--probe runs a short fixed run for profilers instead (see run_probe).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

# The CPU baseline runs the oracle's OpenMP build on the host's cores with the
# reference's binding (tools/scaling_study.sh:36-37).  It runs in a child
# process (--cpu-leg): OMP_PROC_BIND pins the thread that enters the first
# parallel region, and every thread started from it afterwards, so in the
# GPU process it would serialise the host threads of the device set-up.
CPU_LEG_ENV = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}
# the CPUs this process may run on, taken before any OpenMP region pins it
AFFINITY_CPUS = len(os.sched_getaffinity(0))

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cbs-3d_amd"))

HBM_PEAK_GBS = 8000.0

# measurements beside the headline (the default line carries them all)
ALL_LEGS = {
    "amg": "multigrid-preconditioned form on cavity160",
    "multi_rhs": "32 right-hand sides together on the cavity160 operator, FP64 and mixed",
    "c0": "cavity20, its 200 steps, GPU and CPU",
    "c1": "channel128 with energy, its 500 steps on the GPU, the first 3 on the CPU",
    "ensemble": "64 cavities of n in 12..64 in one batch context",
    "adjoint": "adjoint of channel64 (100 steps) with a central-difference check",
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(workload):
    """HBM bytes per apply from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/ for this workload (copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d if d.get("workload") == workload else None


def run_form(ctx, init, precision, warmup, steps, quiet=False, profile_steps=-1):
    """From the initial fields: `warmup` untimed CBS steps, then `steps`
    timed ones.  profile_steps < 0: the apply timer (HIP events on the
    context's stream) is on during the timed steps.  profile_steps > 0: the
    timed steps run with every timer off, and the warm-up and the first
    profile_steps of them are then replayed from the same initial fields
    with the timer on (on small meshes, where a step is a few ms, the
    timer's eager iterations between graph replays would weigh on the
    steps/s).  steps = 0 only returns the fields after the warm-up."""
    from cbs3d import FP64

    def start():
        for name, v in init.items():
            ctx.set(name, v)
        ctx.initialise()
        for it in range(1, warmup + 1):
            ctx.advanceOneStep(it, precision=precision)

    ctx.set_profile(False)
    start()
    if steps == 0:
        return {"fields_after_warmup": {k: ctx.get(k) for k in ("unkno", "pres", "temperature")}}
    if profile_steps < 0:
        ctx.set_profile(True)
        ctx.reset_apply_timer()
    ctx.synchronize()
    t0 = time.perf_counter()
    results = []
    for it in range(warmup + 1, warmup + steps + 1):
        results.append(ctx.advanceOneStep(it, precision=precision))
    ctx.synchronize()
    wall = time.perf_counter() - t0
    fields = {k: ctx.get(k) for k in ("pres", "unkno")}
    graphs = ctx.graphs_active()[0 if precision == FP64 else 1]
    apply_ms, apply_n = 0.0, 0
    if profile_steps < 0:
        apply_ms, apply_n = ctx.apply_timer(precision)
        ctx.set_profile(False)
    elif profile_steps > 0:
        ctx.set_profile(True)
        start()
        ctx.reset_apply_timer()
        for it in range(warmup + 1, warmup + min(steps, profile_steps) + 1):
            ctx.advanceOneStep(it, precision=precision)
        ctx.synchronize()
        apply_ms, apply_n = ctx.apply_timer(precision)
        ctx.set_profile(False)
    threads, batch, cand = ctx.apply_config(precision)
    iters = [r.iterations for r in results]
    cg_bytes = sum(r.algorithmic_bytes for r in results)
    cg_secs = sum(r.total_seconds for r in results)
    tag = "fp64" if precision == FP64 else "mixed"
    if not quiet:
        log(f"[bench] {tag}: {steps} steps in {wall:.3f} s, CG iterations {iters}, "
            f"refinements {[r.refinements for r in results]}")
    return dict(wall=wall, steps_per_s=steps / wall, iters=iters, cg_gbs=cg_bytes / cg_secs / 1e9,
                apply_ms=apply_ms, apply_n=apply_n, graphs=graphs,
                tile=dict(threads=threads, pair_batch=batch, autotune_ms=cand), fields=fields)


def run_ensemble(count, steps, sequential=True, forms=("fp64", "mixed")):
    """The seeded cavity batch (cases.ensemble: n ~ U{12..64}, Re log-uniform
    in [50, 1000], seed 3): one batch context per solver form, timing `steps`
    CBS steps after one untimed step (which also runs the apply autotuner and
    captures the PCG graphs), against the same members run one context at a
    time (FP64)."""
    from cbs3d import FP64, MIXED, BatchContext, CaseSetup, HipContext, cases

    t0 = time.perf_counter()
    members = cases.ensemble(count=count)
    setups = [CaseSetup.from_case(c) for c in members]
    log(f"[bench] ensemble: {count} cavities, {sum(c.nelem for c in members)} tets, "
        f"{sum(c.npoin for c in members)} nodes ({time.perf_counter() - t0:.1f} s to set up)")
    init = {k: np.concatenate([st.array(k) for st in setups]) for k in ("unkno", "pres", "temperature")}
    t0 = time.perf_counter()
    batch = BatchContext(setups)
    batch.buildElementPressureTerms()
    out = {"members": count, "tets": int(sum(c.nelem for c in members)),
           "nodes": int(sum(c.npoin for c in members)), "steps": steps,
           "n_range": [int(min(c.nx for c in members)), int(max(c.nx for c in members))],
           "batch_context_s": round(time.perf_counter() - t0, 2)}
    for tag in forms:
        prec = {"fp64": FP64, "mixed": MIXED}[tag]
        for k, v in init.items():
            batch.set(k, v)
        batch.initialise()
        batch.advanceOneStep(1, precision=prec)
        batch.synchronize()
        t0 = time.perf_counter()
        for it in range(2, steps + 2):
            batch.advanceOneStep(it, precision=prec)
        batch.synchronize()
        wall = time.perf_counter() - t0
        iters = [batch.member_cg_result(k).iterations for k in range(count)]
        log(f"[bench] ensemble batched {tag}: {steps} steps in {wall:.3f} s, last-step iterations "
            f"min {min(iters)} max {max(iters)}")
        out[tag] = {"batched_member_steps_per_s": round(count * steps / wall, 3),
                    "ms_per_step": round(1e3 * wall / steps, 3),
                    "last_step_cg_iterations": {"min": min(iters), "max": max(iters)},
                    "hip_graphs": batch.graphs_active()[0 if tag == "fp64" else 1]}
    batch.close()
    if sequential:
        seq = 0.0
        for setup in setups:
            ctx = HipContext(setup)
            ctx.initialise()
            ctx.buildElementPressureTerms()
            ctx.advanceOneStep(1)
            ctx.synchronize()
            t0 = time.perf_counter()
            for it in range(2, steps + 2):
                ctx.advanceOneStep(it)
            ctx.synchronize()
            seq += time.perf_counter() - t0
            ctx.close()
        log(f"[bench] ensemble one context at a time: {steps} steps in {seq:.3f} s")
        out["fp64_sequential_member_steps_per_s"] = round(count * steps / seq, 3)
        out["fp64_batched_over_sequential"] = round(out["fp64"]["batched_member_steps_per_s"] /
                                                    out["fp64_sequential_member_steps_per_s"], 3)
    return out


def run_amg(ctx, init, warmup, steps, fp64_fields):
    """The multigrid-preconditioned form (build-only) on the same workload
    and steps: CBS steps/s, PCG iterations, the hierarchy's set-up, and the
    fields' distance from the FP64 Jacobi run."""
    from cbs3d import AMG

    for name, v in init.items():
        ctx.set(name, v)
    ctx.initialise()
    t0 = time.perf_counter()
    for it in range(1, warmup + 1):
        ctx.advanceOneStep(it, precision=AMG)
    ctx.synchronize()
    warm = time.perf_counter() - t0
    info = ctx.amg_info()
    t0 = time.perf_counter()
    results = [ctx.advanceOneStep(it, precision=AMG) for it in range(warmup + 1, warmup + steps + 1)]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    iters = [r.iterations for r in results]
    err = {k: float(np.linalg.norm(ctx.get(k) - fp64_fields[k]) / max(np.linalg.norm(fp64_fields[k]), 1e-300))
           for k in ("pres", "unkno")}
    log(f"[bench] amg: {steps} steps in {wall:.3f} s, CG iterations {iters}; hierarchy {info} "
        f"(warm-up incl. set-up {warm:.1f} s)")
    return {"steps_per_s": round(steps / wall, 4), "ms_per_step": round(1e3 * wall / steps, 3),
            "cg_iterations": iters, "hierarchy": info, "graphs": ctx.graphs_active(),
            "vs_fp64_rel_l2": err,
            "note": "FP64 CG with one FP32 smoothed-aggregation V-cycle as preconditioner (build-only; "
                    "the reference's preconditioner is Jacobi); same workload, steps and 1e-8 criterion"}


def run_output(ctx, setup, init, steps, every, no_output_steps_per_s):
    """Solution output on the workload: mixed steps with the residual CSV
    every step and a VTU every `every` steps written by the writer's thread
    from device snapshots."""
    import ctypes as C
    import shutil
    import tempfile

    from cbs3d import MIXED
    from cbs3d import post as cpost
    from cbs3d._native import hip_lib

    for name, v in init.items():
        ctx.set(name, v)
    ctx.initialise()
    root = tempfile.mkdtemp(prefix="cbs_bench_out_")
    try:
        t0 = time.perf_counter()
        w = cpost.Post(setup, cpost.config(vtu_output_every_iterations=every, write_boundary_debug_arrays=1),
                       root=root, case_name="bench")
        create_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        w.initialiseRunOutputs(ctx=ctx)  # iteration 0, synchronous
        init_s = time.perf_counter() - t0
        # the snapshot's cost on the compute stream, with both slots
        # allocated (their first use allocates device and pinned memory)
        lib = hip_lib()
        first = [C.c_void_p(), C.c_void_p()]
        t0 = time.perf_counter()
        for sn in first:
            ctx._check(lib.cbs_hip_snapshot_begin(ctx._h, C.byref(sn)))
        for sn in first:
            lib.cbs_hip_snapshot_wait(sn)
            lib.cbs_hip_snapshot_release(sn)
        alloc_s = time.perf_counter() - t0
        snap = C.c_void_p()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx._check(lib.cbs_hip_snapshot_begin(ctx._h, C.byref(snap)))
        ctx.synchronize()
        snap_ms = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        lib.cbs_hip_snapshot_wait(snap)
        download_ms = snap_ms + 1e3 * (time.perf_counter() - t0)
        lib.cbs_hip_snapshot_release(snap)
        t0 = time.perf_counter()
        for it in range(1, steps + 1):
            ctx.advanceOneStep(it, precision=MIXED)
            sc = ctx.scalars()
            w.writeResidualRow(it, sc)
            if w.shouldWriteVTU(it, sc.rtime):
                w.writeSolution(it, sc.rtime, ctx=ctx)
        ctx.synchronize()
        loop_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        w.flush()
        drain_s = time.perf_counter() - t0
        st = w.stats()
        w.close()
    finally:
        shutil.rmtree(root, ignore_errors=True)
    per_file = st["vtu_bytes"] / max(st["vtu_files"], 1)
    out = {"vtu_every": every, "steps": steps, "steps_per_s": round(steps / loop_s, 4),
           "steps_per_s_without_output": round(no_output_steps_per_s, 4),
           "drain_after_loop_s": round(drain_s, 3), "snapshot_ms_on_compute_stream": round(snap_ms, 3),
           "snapshot_ms_to_host": round(download_ms, 3), "snapshot_slots_first_use_s": round(alloc_s, 3),
           "vtu_files": st["vtu_files"], "vtu_bytes_per_file": int(per_file),
           "writer_gbs": round(st["vtu_bytes"] / max(st["write_seconds"], 1e-9) / 1e9, 3),
           "writer_seconds_per_vtu": round(st["write_seconds"] / max(st["vtu_files"], 1), 3),
           "static_part_encode_s": round(create_s, 3), "initial_vtu_sync_s": round(init_s, 3)}
    log(f"[bench] output: {out}")
    return out


def run_setup(setup, ctx, ctx_seconds):
    """Host set-up phases against the device colouring on the workload, and
    the phases of the device context's construction."""
    from cbs3d import cases, device_coloring_of

    names = ["validate_bcs", "shape_functions", "boundary_faces", "normals", "mass_matrix", "face_edges",
             "element_size", "walls", "mass_flow_inlet", "pressure_nodes", "coloring"]
    phases = {n: round(float(v), 4) for n, v in zip(names, setup.array("setup_seconds"))}
    d = setup.mesh_desc()
    intma = setup.array("intma")
    small = cases.cavity(8, name="warm")
    device_coloring_of(small.intma.reshape(-1), small.npoin)  # warm-up: module load, first allocations
    t0 = time.perf_counter()
    nc, ptr, elem, levels = device_coloring_of(intma, d.npoin)
    dev_s = time.perf_counter() - t0
    same = bool(nc == d.ncolor and np.array_equal(ptr, setup.array("color_ptr")[:nc + 1]) and
                np.array_equal(elem, setup.array("color_elem")))
    # case files: the reference's ASCII .plt against the binary mesh, on a
    # bounded sample
    import shutil
    import tempfile

    from cbs3d import CaseSetup
    sample = cases.cavity(64, re=100.0, jitter=0.1, seed=2, name="read64")
    tmp = tempfile.mkdtemp(prefix="cbs_bench_read_")
    try:
        reads = {"sample": "cavity 64^3 x 6 (1,572,864 tets)"}
        for tag, binary in (("ascii_plt_s", False), ("binary_cbsmesh_s", True)):
            base = cases.write_case_files(sample, os.path.join(tmp, tag), binary=binary)
            t0 = time.perf_counter()
            CaseSetup.read(base, setup=False)
            reads[tag] = round(time.perf_counter() - t0, 3)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out = {"host_phases_s": phases, "host_total_s": round(sum(phases.values()), 3), "case_read": reads,
           "host_coloring_threads": 1, "device_coloring_s": round(dev_s, 4), "device_coloring_levels": levels,
           "colours": nc, "coloring_speedup": round(phases["coloring"] / dev_s, 1),
           "device_equals_host": same, "context_s": round(ctx_seconds, 3),
           "context_phases_s": ctx.create_seconds()}
    log(f"[bench] setup: {out}")
    return out


def run_multi_rhs(ctx, nrhs, single_sample=2):
    """nrhs right-hand sides b ~ N(0,1) per node minus their mean (pinned rows
    take their BC value inside the solve), solved together on the context's
    current operator, against single solves of the first `single_sample`
    columns (scaled to nrhs), in FP64 and in the mixed form."""
    from cbs3d import FP64, MIXED

    N = ctx.npoin
    rng = np.random.default_rng(3)
    B = rng.standard_normal((nrhs, N))
    B -= B.mean(axis=1, keepdims=True)
    out = {"nrhs": nrhs, "single_sample": single_sample}
    for prec, tag in ((FP64, "fp64"), (MIXED, "mixed")):
        ctx.synchronize()
        t0 = time.perf_counter()
        X, res = ctx.solvePressureMulti(B, precision=prec)
        t_multi = time.perf_counter() - t0
        its = [r.iterations for r in res]
        assert all(r.converged for r in res)
        log(f"[bench] {nrhs}-RHS {tag} PCG: {t_multi:.2f} s, iterations min {min(its)} max {max(its)}")
        t_single = 0.0
        single_its = []
        for j in range(single_sample):
            ctx.set("rhs1", B[j])
            ctx.set("pres", np.zeros(N))
            ctx.synchronize()
            t0 = time.perf_counter()
            r = ctx.solvePressure(prec)
            t_single += time.perf_counter() - t0
            single_its.append(r.iterations)
        seq = t_single / single_sample * nrhs
        log(f"[bench] {tag} single solves: {single_sample} in {t_single:.2f} s (iterations {single_its})")
        out[tag] = {"seconds": round(t_multi, 3), "iterations": {"min": min(its), "max": max(its)},
                    "column_solves_per_s": round(nrhs / t_multi, 3),
                    "sequential_estimate_seconds": round(seq, 3),
                    "multi_over_sequential": round(seq / t_multi, 3)}
    return out


def run_adjoint(fd_directions=4, h=1e-6):
    """The adjoint workload (channel 64x16x16x6, Re 100, 100 steps, dt fixed
    at the channel's first-step stability time step, objective over the last
    20 steps, f ~ N(0, 0.01^2) per node seed 4, inflow coefficients 0):
    device forward and backward times, and central differences of the device
    forward on random unit directions of (f, c) against the adjoint."""
    from cbs3d import CaseSetup, HipContext, cases, stability_dt

    w = cases.ADJOINT_WORKLOAD
    dt = stability_dt(cases.channel(w["nx"], w["ny"], w["nz"], re=w["re"], pr=0.71, jitter=0.0, seed=4,
                                    ntime=w["ntime"]))
    case = cases.adjoint_channel(w["nx"], w["ny"], w["nz"], re=w["re"], ntime=w["ntime"], dtfix=dt,
                                 rel_toler=1e-12)
    setup = CaseSetup.from_case(case)
    ctx = HipContext(setup)
    ctx.buildElementPressureTerms()
    f = np.random.default_rng(4).normal(0.0, 0.01, (ctx.npoin, 3))
    c = np.zeros(16)
    u0, p0 = setup.array("unkno"), setup.array("pres")

    def forward(ff, cc):
        ctx.set("unkno", u0)
        ctx.set("pres", p0)
        ctx.set_body_force(ff)
        ctx.set_inflow_coefficients(cc)
        return ctx.adjoint_forward(w["ntime"], w["objective_steps"], 10, 1e-12)

    forward(f, c)  # warm-up (autotuner)
    t0 = time.perf_counter()
    phi = forward(f, c)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    gf, gc = ctx.adjoint_backward()
    t_bwd = time.perf_counter() - t0
    log(f"[bench] adjoint {case.name}: dtfix {dt:.6g}, Phi {phi:.12g}, forward {t_fwd:.2f} s, backward {t_bwd:.2f} s")
    rng = np.random.default_rng(7)
    gnorm = float(np.sqrt((gf ** 2).sum() + (gc ** 2).sum()))
    abs_err, rel_err = [], []
    for k in range(fd_directions):
        df = rng.standard_normal(f.shape)
        dc = rng.standard_normal(16)
        nrm = np.sqrt((df ** 2).sum() + (dc ** 2).sum())
        df, dc = df / nrm, dc / nrm
        fd = (forward(f + h * df, c + h * dc) - forward(f - h * df, c - h * dc)) / (2 * h)
        ad = float((gf * df).sum() + (gc * dc).sum())
        abs_err.append(abs(fd - ad))
        rel_err.append(abs(fd - ad) / max(abs(ad), 1e-300))
        if (k + 1) % 4 == 0:
            log(f"[bench] adjoint FD: {k + 1} of {fd_directions} directions")
    log(f"[bench] adjoint FD over {fd_directions} directions: max |fd-ad|/|g| {max(abs_err) / gnorm:.2e}, "
        f"median rel {np.median(rel_err):.2e}")
    ctx.close()
    return {"workload": f"channel {w['nx']}x{w['ny']}x{w['nz']}x6, Re {w['re']}, {w['ntime']} steps, "
                        f"objective over the last {w['objective_steps']}",
            "tets": int(case.nelem), "nodes": int(case.npoin), "dtfix": dt, "objective": phi,
            "forward_s": round(t_fwd, 3), "backward_s": round(t_bwd, 3),
            "backward_over_forward": round(t_bwd / t_fwd, 3), "checkpoint_every": 10,
            "fd": {"directions": fd_directions, "h": h, "max_abs_err_over_grad_norm": max(abs_err) / gnorm,
                   "median_rel_err": float(np.median(rel_err)), "max_rel_err": max(rel_err)}}


ORACLE_ARRAYS = ["intma", "coord", "mat_elem", "rho_e", "mu_e", "k_e", "rho_cp_e", "alpha_e", "Qvol_e",
                 "dNkdx", "detJ", "annxf", "fedge", "iside", "face_norm", "elcoe_e", "alen_e",
                 "wall_node_list", "bc_list", "bc_values", "color_ptr", "color_elem",
                 "unkno", "pres", "temperature"]


def cpu_threads() -> int:
    """Threads of the CPU baseline: OMP_NUM_THREADS, else the CPUs this
    process may run on."""
    n = os.environ.get("OMP_NUM_THREADS", "")
    return int(n) if n.isdigit() and int(n) > 0 else AFFINITY_CPUS


def host_cpu() -> dict:
    """The host's CPU as lscpu reports it, and the threads the baseline uses."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    model = info.get("Model name")
    if not model:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "unknown")

    def num(key):
        v = info.get(key, "")
        return int(v) if v.isdigit() else None

    sockets, cores = num("Socket(s)"), num("Core(s) per socket")
    return {"model": model, "sockets": sockets, "physical_cores": sockets * cores if sockets and cores else None,
            "threads_per_core": num("Thread(s) per core"), "cpus_online": num("CPU(s)"),
            "cpus_in_affinity": AFFINITY_CPUS, "omp_threads": cpu_threads(),
            "omp_proc_bind": CPU_LEG_ENV["OMP_PROC_BIND"], "omp_places": CPU_LEG_ENV["OMP_PLACES"]}


def oracle_state(setup, openmp=True):
    """The oracle (its OpenMP build by default) seeded with a set-up's
    arrays, after Solver::initialise's boundary values and pressure terms."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    d = setup.mesh_desc()
    cfg = setup.config().as_dict()
    cfg.update(npoin=d.npoin, nelem=d.nelem, nboun=d.nboun, ncolor=d.ncolor, bc_fixed=d.bc_fixed,
               npoin_wall=d.npoin_wall)
    st = orc.OracleState({k: setup.array(k) for k in ORACLE_ARRAYS}, cfg, openmp=openmp)
    st.apply_temperature()
    st.apply_velocity()
    st.apply_pressure()
    st.build_element_pressure_terms()
    return st


def run_cpu_leg(workload, steps=0, sample_iters=0, iters_per_step=0.0):
    """Runs one CPU measurement in a child process (bench.py --cpu-leg) with
    the reference's OpenMP binding: `steps` whole steps of the workload
    (cpu_full_steps), or the sampled step (cpu_baseline) when steps is 0.
    Returns its record; the fields of a whole-step run come back as arrays."""
    import subprocess
    import tempfile

    fd, out = tempfile.mkstemp(prefix="cbs_cpu_leg_", suffix=".npz")
    os.close(fd)
    try:
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", workload, "--cpu-leg-steps", str(steps),
               "--cpu-sample-iters", str(sample_iters), "--cpu-iters-per-step", repr(iters_per_step),
               "--cpu-leg-out", out]
        r = subprocess.run(cmd, env=dict(os.environ, **CPU_LEG_ENV), capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CPU leg {workload} failed ({r.returncode}): {r.stderr[-2000:]}")
        rec = json.loads(r.stdout.strip().splitlines()[-1])
        if steps > 0:
            with np.load(out, allow_pickle=False) as z:
                rec["fields"] = {k: z[k] for k in z.files}
        return rec
    finally:
        os.unlink(out)


def cpu_leg_main(args) -> int:
    """--cpu-leg: the child process of run_cpu_leg (no GPU work)."""
    from cbs3d import CaseSetup, cases

    setup = CaseSetup.from_case(cases.workload(args.cpu_leg))
    if args.cpu_leg_steps > 0:
        rec = cpu_full_steps(setup, args.cpu_leg_steps)
        np.savez(args.cpu_leg_out, **rec.pop("fields"))
    else:
        rec = cpu_baseline(setup, args.cpu_iters_per_step, args.cpu_sample_iters)
    print(json.dumps(rec), flush=True)
    return 0


def cpu_full_steps(setup, nsteps):
    """The OpenMP oracle over the first nsteps whole CBS steps: steps/s with
    nothing extrapolated, PCG iterations, and the fields it ends with."""
    st = oracle_state(setup)
    t0 = time.perf_counter()
    its = []
    for it in range(1, nsteps + 1):
        st.advance_one_step(it)
        its.append(int(st.last_cg_iterations))
    wall = time.perf_counter() - t0
    return {"value": nsteps / wall, "unit": "CBS steps/s", "steps": nsteps, "seconds": wall, "cores": cpu_threads(),
            "cg_iterations": {"min": min(its), "max": max(its), "mean": float(np.mean(its))},
            "fields": {k: st[k].copy() for k in ("unkno", "pres", "temperature")}}


def cpu_baseline(setup, iters_per_step, sample_iters):
    """The cavity160 step on the host's cores (OpenMP oracle), on a bounded
    sample: the non-solver stages of one step once, plus sample_iters PCG
    iterations, scaled to a step with the GPU's mean PCG iterations per step
    (one whole step is ~1,500 iterations: minutes of CPU time)."""
    st = oracle_state(setup)
    d = setup.mesh_desc()
    st.begin_step()
    t0 = time.perf_counter()
    st.compute_timestep(1)
    st.update_lhs_diagonal()
    st.step1()
    st.assemble_step2_rhs()
    t_pre = time.perf_counter() - t0
    st._s.cg_max_iter = sample_iters
    t0 = time.perf_counter()
    r = st.solve_pressure()
    t_cg = time.perf_counter() - t0
    per_iter = t_cg / max(r.iterations, 1)
    t0 = time.perf_counter()
    st.step3()
    st.update_velocity_magnitude()
    st.convergence()
    t_post = time.perf_counter() - t0
    step_s = t_pre + t_post + iters_per_step * per_iter
    bytes_iter = d.nelem * 64.0 + d.npoin * 24.0 + d.npoin * 65.0
    threads = cpu_threads()
    return dict(value=1.0 / step_s, unit="CBS steps/s", cores=threads, kind="port",
                pcg_gbs=bytes_iter / per_iter / 1e9, cpu=host_cpu(),
                sample=(f"oracle with the reference's OpenMP loops, {threads} threads: time step + LHS + Step 1 + "
                        f"Step 2 RHS + Step 3 + norms once ({t_pre + t_post:.2f} s) and {r.iterations} FP64 PCG "
                        f"iterations ({t_cg:.2f} s, {per_iter * 1e3:.1f} ms/iteration) on the cavity160 mesh, "
                        f"scaled to one step by {iters_per_step:.1f} PCG iterations per step (GPU FP64 mean)"))


def apply_bytes(ctx, precision):
    """Algorithmic bytes of one PCG apply of a solver form and whether it
    forms the direction: nnz*(s+2) + N*3s, plus N*2s (FP64: r read, p
    written) or N*3s (FP32: r and the inverse diagonal read, p written) when
    the direction update is fused into it."""
    from cbs3d import FP64

    op = ctx.operator_info()
    s = 8 if precision == FP64 else 4
    fused = ctx.direction_fusion(precision)[0] == 1
    extra = (2 if precision == FP64 else 3) * s if fused else 0
    return op["nnz"] * (s + 2) + ctx.npoin * (3 * s + extra), fused


def apply_roofline(ctx, form, precision):
    """The PCG apply of one form: mean event time per apply and GB/s of its
    algorithmic bytes (apply_bytes) against the 8 TB/s peak."""
    b, fused = apply_bytes(ctx, precision)
    if not form["apply_n"]:
        return {"bytes_per_apply": b, "direction_fused": fused, "ms_per_apply": None, "achieved_gbs": None,
                "frac": None}
    ms = form["apply_ms"] / form["apply_n"]
    gbs = b / (ms * 1e-3) / 1e9
    return {"bytes_per_apply": b, "direction_fused": fused, "ms_per_apply": round(ms, 5),
            "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "timed_applies": form["apply_n"]}


def run_config_leg(name, steps, cpu_steps, note):
    """One of the smaller configurations as its own measurement: the FP64
    and mixed forms for `steps` timed CBS steps each from the initial state
    (after one untimed priming step that runs the apply autotuner), their
    apply rooflines, the OpenMP oracle over its first `cpu_steps` whole
    steps, and the per-field relative L2 (E2) of the GPU fields after
    cpu_steps steps against the oracle's."""
    from cbs3d import FP64, MIXED, CaseSetup, HipContext, cases

    t0 = time.perf_counter()
    case = cases.workload(name)
    setup = CaseSetup.from_case(case)
    ctx = HipContext(setup)
    ctx.buildElementPressureTerms()
    init = {k: setup.array(k) for k in ("unkno", "pres", "temperature")}
    log(f"[bench] {name}: {case.nelem} tets, {case.npoin} nodes, ready in {time.perf_counter() - t0:.1f} s")
    out = {"workload": name, "tets": case.nelem, "nodes": case.npoin, "note": note}
    fields_at = {}
    for prec, tag in ((FP64, "fp64"), (MIXED, "mixed")):
        run_form(ctx, init, prec, 1, 0)  # priming: autotuner, graph capture
        # the first cpu_steps steps, for the comparison with the oracle
        first = run_form(ctx, init, prec, cpu_steps, 0)
        fields_at[tag] = first["fields_after_warmup"]
        f = run_form(ctx, init, prec, 0, steps, quiet=True, profile_steps=max(1, steps // 10))
        out[tag] = {"steps": steps, "steps_per_s": round(f["steps_per_s"], 3),
                    "ms_per_step": round(1e3 * f["wall"] / steps, 4),
                    "cg_iterations": {"min": min(f["iters"]), "max": max(f["iters"]),
                                      "mean": round(float(np.mean(f["iters"])), 1)},
                    "pcg_gbs": round(f["cg_gbs"], 1), "hip_graphs": f["graphs"],
                    "apply": apply_roofline(ctx, f, prec)}
    ctx.close()
    out["mixed_over_fp64"] = round(out["mixed"]["steps_per_s"] / out["fp64"]["steps_per_s"], 4)
    t0 = time.perf_counter()
    cpu = run_cpu_leg(name, steps=cpu_steps)
    log(f"[bench] {name}: CPU {cpu_steps} steps in {time.perf_counter() - t0:.1f} s")
    e2 = {}
    for tag, flds in fields_at.items():
        e2[tag] = {k: float(np.linalg.norm(flds[k] - cpu["fields"][k]) /
                            max(np.linalg.norm(cpu["fields"][k]), 1e-300)) for k in cpu["fields"]}
    out["cpu_baseline"] = {"value": round(cpu["value"], 4), "unit": "CBS steps/s", "cores": cpu["cores"],
                           "kind": "port", "steps": cpu_steps, "seconds": round(cpu["seconds"], 3),
                           "cg_iterations": cpu["cg_iterations"],
                           "sample": f"oracle with the reference's OpenMP loops over the first {cpu_steps} whole "
                                     f"steps (nothing extrapolated)"}
    out["gpu_over_cpu_fp64"] = round(out["fp64"]["steps_per_s"] / cpu["value"], 1)
    out["e2_vs_oracle_after_cpu_steps"] = e2
    log(f"[bench] {name}: {json.dumps({k: v for k, v in out.items() if k not in ('note',)})}")
    return out


def run_probe(args) -> int:
    """--probe: a short fixed run for profilers (rocprofv3 counter passes),
    with no JSON line.  Builds the workload's context, advances --steps CBS
    steps with the --precision form (the first solve also runs the apply
    autotuner; amg's first solve also builds the hierarchy), then optionally
    solves --probe-multi-rhs right-hand sides together (b ~ N(0,1) per node
    minus its mean, seed 3) on the last step's operator, then runs the plain-copy
    kernel (k_copy) over --copy-probe-bytes: its counters calibrate the
    HBM-byte counters against a kernel of known traffic (read + write of that
    many bytes each).  Prints one line per stage."""
    from cbs3d import AMG, FP64, MIXED, CaseSetup, HipContext, cases

    t0 = time.perf_counter()
    setup = CaseSetup.from_case(cases.workload(args.workload))
    ctx = HipContext(setup)
    if args.no_graphs:
        ctx.set_graphs(False)
    if args.fusion != "auto":
        ctx.set_direction_fusion(FP64, int(args.fusion))
        ctx.set_direction_fusion(MIXED, int(args.fusion))
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
    if args.probe_multi_rhs > 0:
        B = np.random.default_rng(3).standard_normal((args.probe_multi_rhs, ctx.npoin))
        B -= B.mean(axis=1, keepdims=True)
        mprec = MIXED if args.precision == "mixed" else FP64
        t0 = time.perf_counter()
        _, res = ctx.solvePressureMulti(B, precision=mprec)
        its = [r.iterations for r in res]
        print(f"[probe] {args.probe_multi_rhs}-RHS PCG ({'mixed' if mprec == MIXED else 'fp64'}): "
              f"{time.perf_counter() - t0:.2f} s, iterations {min(its)}..{max(its)}", flush=True)
    if args.copy_probe_bytes > 0:
        gbs = ctx.bandwidth_probe(args.copy_probe_bytes, reps=1)
        print(f"[probe] k_copy of {args.copy_probe_bytes} B (read + write each): {gbs:.0f} GB/s", flush=True)
    ctx.close()
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cavity160")
    ap.add_argument("--cpu-sample-iters", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--legs", default=",".join(ALL_LEGS),
                    help="measurements beside the headline, comma-separated (empty: none): " + ", ".join(
                        f"{k} ({v})" for k, v in ALL_LEGS.items()))
    ap.add_argument("--ensemble", type=int, default=64,
                    help="members of the ensemble leg (64: the ensemble workload)")
    ap.add_argument("--ensemble-steps", type=int, default=3)
    ap.add_argument("--ensemble-sequential", type=int, default=1,
                    help="also run the ensemble's members one context at a time (0: skip)")
    ap.add_argument("--adjoint-fd", type=int, default=4,
                    help="central-difference directions of the adjoint leg (32: the full check)")
    ap.add_argument("--multi-rhs", type=int, default=32,
                    help="right-hand sides of the multi-RHS leg")
    ap.add_argument("--c0-steps", type=int, default=200, help="GPU steps of the cavity20 leg (its full run)")
    ap.add_argument("--c1-steps", type=int, default=500, help="GPU steps of the channel128 leg (its full run)")
    ap.add_argument("--c1-cpu-steps", type=int, default=3, help="CPU steps of the channel128 leg")
    ap.add_argument("--setup", action="store_true",
                    help="also measure the host set-up phases against the device colouring")
    ap.add_argument("--output", type=int, default=0,
                    help="also measure solution output with a VTU every K steps (mixed form)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the PCG kernels one by one instead of replaying captured HIP graphs")
    ap.add_argument("--copy-probe-bytes", type=int, default=1 << 30,
                    help="size of the plain-copy calibration kernel (0 to skip)")
    ap.add_argument("--probe", action="store_true",
                    help="short fixed run for profilers: --steps steps of --precision, then --probe-multi-rhs "
                         "and the copy kernel; prints stage lines and no JSON")
    ap.add_argument("--precision", choices=("fp64", "mixed", "amg"), default="mixed",
                    help="solver form of --probe")
    ap.add_argument("--fusion", choices=("auto", "0", "1"), default="auto",
                    help="the PCG direction update of the headline context in its own kernel (0) or in the "
                         "apply (1), e.g. for reproducible profiler runs; auto (default): timed at the first solve")
    ap.add_argument("--probe-multi-rhs", type=int, default=0,
                    help="right-hand sides --probe solves together after its steps (0: none)")
    ap.add_argument("--cpu-leg", default="", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-leg-steps", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-iters-per-step", type=float, default=0.0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-leg-out", default="", help=argparse.SUPPRESS)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              f"(the hot path runs on a single GPU; how it would shard is design only)", file=sys.stderr)
        return 2
    if args.cpu_leg:
        return cpu_leg_main(args)
    if args.probe:
        return run_probe(args)
    legs = [x for x in args.legs.split(",") if x]
    unknown = [x for x in legs if x not in ALL_LEGS]
    if unknown:
        print(f"bench.py: unknown legs {unknown}; known: {list(ALL_LEGS)}", file=sys.stderr)
        return 2

    from cbs3d import FP64, MIXED, CaseSetup, HipContext, cases

    t0 = time.perf_counter()
    case = cases.workload(args.workload)
    log(f"[bench] workload {args.workload}: {case.nelem} tets, {case.npoin} nodes "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    t0 = time.perf_counter()
    setup = CaseSetup.from_case(case)
    d = setup.mesh_desc()
    log(f"[bench] host set-up {time.perf_counter() - t0:.1f} s, {d.ncolor} colours")
    t0 = time.perf_counter()
    ctx = HipContext(setup)
    ctx_seconds = time.perf_counter() - t0
    if args.no_graphs:
        ctx.set_graphs(False)
    if args.fusion != "auto":
        ctx.set_direction_fusion(FP64, int(args.fusion))
        ctx.set_direction_fusion(MIXED, int(args.fusion))
    ctx.buildElementPressureTerms()
    log(f"[bench] device context {time.perf_counter() - t0:.1f} s")
    init = {k: setup.array(k) for k in ("unkno", "pres", "temperature")}

    fp64 = run_form(ctx, init, FP64, args.warmup, args.steps)
    mixed = run_form(ctx, init, MIXED, args.warmup, args.steps)
    e_pres = float(np.linalg.norm(mixed["fields"]["pres"] - fp64["fields"]["pres"]) /
                   max(np.linalg.norm(fp64["fields"]["pres"]), 1e-300))
    e_vel = float(np.linalg.norm(mixed["fields"]["unkno"] - fp64["fields"]["unkno"]) /
                  max(np.linalg.norm(fp64["fields"]["unkno"]), 1e-300))

    F = int(ctx.nfluid)
    N = d.npoin
    op = ctx.operator_info()
    b64, fused64 = apply_bytes(ctx, FP64)
    b32, fused32 = apply_bytes(ctx, MIXED)
    apply64 = fp64["apply_ms"] / max(fp64["apply_n"], 1)
    apply32 = mixed["apply_ms"] / max(mixed["apply_n"], 1)
    ach64 = b64 / (apply64 * 1e-3) / 1e9
    traffic = measured_traffic(args.workload)
    roofline = {"bound": "hbm", "achieved": round(ach64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach64 / HBM_PEAK_GBS, 4),
                "traffic": traffic.get("fp64_apply_fused_hbm_bytes" if fused64 else "fp64_apply_hbm_bytes")
                if traffic else None,
                "traffic_source": traffic.get("source") if traffic else None,
                "kernel": "k_apply_sell<double,true,B,%s> (FP64 PCG apply, assembled operator%s)" % (
                    "true" if fused64 else "false", ", direction update fused" if fused64 else ""),
                "direction_fused": fused64, "direction_fusion_ms": list(ctx.direction_fusion(FP64)[1:]),
                "ms_per_apply": round(apply64, 5), "bytes_per_apply": b64,
                "bytes_per_apply_element_form": F * (6 * 8 + 16) + N * 3 * 8, "timed_applies": fp64["apply_n"],
                "launch": fp64["tile"], "operator": op}
    fp32_roof = {"kernel": "k_apply_sell<float,true,B,%s> (FP32 apply of the mixed form, assembled operator%s)" % (
                     "true" if fused32 else "false", ", direction update fused" if fused32 else ""),
                 "direction_fused": fused32, "direction_fusion_ms": list(ctx.direction_fusion(MIXED)[1:]),
                 "ms_per_apply": round(apply32, 5) if mixed["apply_n"] else None,
                 "bytes_per_apply": b32, "timed_applies": mixed["apply_n"], "launch": mixed["tile"],
                 "traffic": traffic.get("fp32_apply_fused_hbm_bytes" if fused32 else "fp32_apply_hbm_bytes")
                 if traffic else None,
                 "achieved_gbs": round(b32 / (apply32 * 1e-3) / 1e9, 1) if mixed["apply_n"] else None}
    fp32_roof["frac"] = round(fp32_roof["achieved_gbs"] / HBM_PEAK_GBS, 4) if fp32_roof["achieved_gbs"] else None

    amg = run_amg(ctx, init, args.warmup, args.steps, fp64["fields"]) if "amg" in legs else None
    multi = run_multi_rhs(ctx, args.multi_rhs) if "multi_rhs" in legs else None
    setup_line = run_setup(setup, ctx, ctx_seconds) if args.setup else None
    output = (run_output(ctx, setup, init, args.steps, args.output, mixed["steps_per_s"])
              if args.output > 0 else None)
    copy_gbs = ctx.bandwidth_probe(args.copy_probe_bytes) if args.copy_probe_bytes > 0 else None
    if copy_gbs:
        log(f"[bench] plain copy of {args.copy_probe_bytes} B: {copy_gbs:.0f} GB/s")
    cpu = None
    if not args.no_cpu_baseline:
        try:
            t0 = time.perf_counter()
            cpu = run_cpu_leg(args.workload, sample_iters=args.cpu_sample_iters,
                              iters_per_step=float(np.mean(fp64["iters"])))
            log(f"[bench] cpu baseline {time.perf_counter() - t0:.1f} s: {cpu['sample']}")
        except Exception as exc:  # the CPU leg must not hide the GPU result
            log(f"[bench] cpu baseline failed: {exc}")

    line = {
        "metric": "CBS steps/sec + PCG GB/s (% HBM peak), mixed vs FP64 at 1e-8 rel residual",
        "value": round(mixed["steps_per_s"], 4),
        "unit": "CBS steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["wall"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 apply / f64 reductions (mixed)",
        "data": "synthetic",
        "config": {"workload": args.workload, "tets": int(d.nelem), "nodes": int(N),
                   "colours": int(d.ncolor), "rel_residual": 1e-8, "parallelism": "single GPU"},
        "mixed": {"steps_per_s": round(mixed["steps_per_s"], 4), "pcg_gbs": round(mixed["cg_gbs"], 1),
                  "pcg_frac_hbm": round(mixed["cg_gbs"] / HBM_PEAK_GBS, 4),
                  "cg_iterations": mixed["iters"], "hip_graphs": mixed["graphs"]},
        "fp64": {"steps_per_s": round(fp64["steps_per_s"], 4), "pcg_gbs": round(fp64["cg_gbs"], 1),
                 "pcg_frac_hbm": round(fp64["cg_gbs"] / HBM_PEAK_GBS, 4), "cg_iterations": fp64["iters"],
                 "ms_per_step": round(1e3 * fp64["wall"] / args.steps, 3), "hip_graphs": fp64["graphs"]},
        "mixed_over_fp64": round(mixed["steps_per_s"] / fp64["steps_per_s"], 4),
        "mixed_vs_fp64_rel_l2": {"pres": e_pres, "unkno": e_vel},
        "roofline": roofline,
        "roofline_fp32_apply": fp32_roof,
        "copy_probe_gbs": round(copy_gbs, 1) if copy_gbs else None,
        "cpu_baseline": cpu,
    }
    if amg:
        line["amg"] = amg
    if multi:
        line["multi_rhs"] = multi
    if output:
        line["output"] = output
    if setup_line:
        line["setup"] = setup_line
    ctx.close()
    del setup
    if "c0" in legs:
        line["c0_cavity20"] = run_config_leg(
            "cavity20", args.c0_steps, args.c0_steps,
            "cavity 20^3 x 6, Re 100: 200 steps on the GPU and on the CPU; cache-resident, so a latency and "
            "launch measurement, not an HBM one")
    if "c1" in legs:
        line["c1_channel128"] = run_config_leg(
            "channel128", args.c1_steps, args.c1_cpu_steps,
            f"channel 128x32x32x6 with energy (Steps 1-4), jitter 0.15h seed 1, Re 200, Pr 0.71: "
            f"{args.c1_steps} GPU steps, the first {args.c1_cpu_steps} on the CPU")
    if "ensemble" in legs:
        line["c3_ensemble"] = run_ensemble(args.ensemble, args.ensemble_steps, bool(args.ensemble_sequential))
    if "adjoint" in legs:
        line["c4_adjoint"] = run_adjoint(args.adjoint_fd)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
