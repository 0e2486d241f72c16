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
off-diagonal nonzero; x, pdiag and y once per node).  The element form the
reference streams, F*(6s+16) + N*3s, is reported beside it
(bytes_per_apply_element_form).
--amg adds the multigrid-preconditioned form (FP64 CG with one FP32
smoothed-aggregation V-cycle per iteration, build-only) on the same
workload: CBS steps/s, iterations, the hierarchy's set-up, and its fields'
distance from the FP64 run.
--setup adds the set-up measured on the same workload: the host set-up's
phases (Solver::initialise: Preprocess and Coloring::build, C++), and
Coloring::build on the device (cbs_hip_coloring, host arrays in and out),
whose colour arrays must equal the host's.
--output K adds the solution output measured on the same workload (mixed
form): CBS steps/s with the residual CSV every step and a VTU every K steps
written asynchronously from device snapshots (include/cbs_post.h), the
snapshot's cost on the compute stream and the writer's throughput.
cpu_baseline: the CPU oracle (single-threaded C restatement of the
reference) on a bounded sample of the same workload: one evaluation of the
non-solver stages plus a fixed number of PCG iterations, scaled to steps/s
with the mean PCG iteration count per step measured on the GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cbs-3d_amd"))

HBM_PEAK_GBS = 8000.0


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


def run_form(ctx, init, precision, warmup, steps):
    from cbs3d import FP64

    for name, v in init.items():
        ctx.set(name, v)
    ctx.initialise()
    ctx.set_profile(True)
    for it in range(1, warmup + 1):
        ctx.advanceOneStep(it, precision=precision)
    ctx.reset_apply_timer()
    t0 = time.perf_counter()
    results = []
    for it in range(warmup + 1, warmup + steps + 1):
        results.append(ctx.advanceOneStep(it, precision=precision))
    ctx.synchronize()
    wall = time.perf_counter() - t0
    apply_ms, apply_n = ctx.apply_timer(precision)
    ctx.set_profile(False)
    threads, batch, cand = ctx.apply_config(precision)
    graphs = ctx.graphs_active()[0 if precision == FP64 else 1]
    iters = [r.iterations for r in results]
    cg_bytes = sum(r.algorithmic_bytes for r in results)
    cg_secs = sum(r.total_seconds for r in results)
    tag = "fp64" if precision == FP64 else "mixed"
    log(f"[bench] {tag}: {steps} steps in {wall:.3f} s, CG iterations {iters}, "
        f"refinements {[r.refinements for r in results]}")
    return dict(wall=wall, steps_per_s=steps / wall, iters=iters, cg_gbs=cg_bytes / cg_secs / 1e9,
                apply_ms=apply_ms, apply_n=apply_n, graphs=graphs,
                tile=dict(threads=threads, pair_batch=batch, autotune_ms=cand),
                fields={k: ctx.get(k) for k in ("pres", "unkno")})


def run_ensemble(count, steps):
    """The seeded cavity batch (cases.ensemble): one batch context against the
    same members run one context at a time; both time `steps` CBS steps
    after one untimed step (which also runs the apply autotuner)."""
    from cbs3d import BatchContext, CaseSetup, HipContext, cases

    t0 = time.perf_counter()
    members = cases.ensemble(count=count)
    setups = [CaseSetup.from_case(c) for c in members]
    log(f"[bench] ensemble: {count} cavities, {sum(c.nelem for c in members)} tets, "
        f"{sum(c.npoin for c in members)} nodes ({time.perf_counter() - t0:.1f} s to set up)")
    batch = BatchContext(setups)
    batch.initialise()
    batch.buildElementPressureTerms()
    batch.advanceOneStep(1)
    batch.synchronize()
    t0 = time.perf_counter()
    for it in range(2, steps + 2):
        batch.advanceOneStep(it)
    batch.synchronize()
    wall = time.perf_counter() - t0
    iters = [batch.member_cg_result(k).iterations for k in range(count)]
    batch.close()
    log(f"[bench] ensemble batched: {steps} steps in {wall:.3f} s, last-step iterations "
        f"min {min(iters)} max {max(iters)}")
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
    return {"members": count, "tets": int(sum(c.nelem for c in members)),
            "nodes": int(sum(c.npoin for c in members)), "steps": steps, "precision": "fp64",
            "batched_member_steps_per_s": round(count * steps / wall, 3),
            "sequential_member_steps_per_s": round(count * steps / seq, 3),
            "batched_over_sequential": round(seq / wall, 3),
            "last_step_cg_iterations": {"min": min(iters), "max": max(iters)}}


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


def run_adjoint(fd_directions=32, h=1e-6):
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


def cpu_baseline(setup, ctx, iters_per_step, sample_iters):
    """Oracle timing on a bounded sample of the workload (single thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    names = ["intma", "coord", "mat_elem", "rho_e", "mu_e", "k_e", "rho_cp_e", "alpha_e", "Qvol_e",
             "dNkdx", "detJ", "annxf", "fedge", "iside", "face_norm", "elcoe_e", "alen_e",
             "wall_node_list", "bc_list", "bc_values", "color_ptr", "color_elem",
             "unkno", "pres", "temperature"]
    d = setup.mesh_desc()
    cfg = setup.config().as_dict()
    cfg.update(npoin=d.npoin, nelem=d.nelem, nboun=d.nboun, ncolor=d.ncolor, bc_fixed=d.bc_fixed,
               npoin_wall=d.npoin_wall)
    st = orc.OracleState({k: setup.array(k) for k in names}, cfg)
    st.apply_temperature()
    st.apply_velocity()
    st.apply_pressure()
    st.build_element_pressure_terms()
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
    return dict(value=1.0 / step_s, unit="CBS steps/s", cores=1, kind="port",
                pcg_gbs=bytes_iter / per_iter / 1e9,
                sample=(f"oracle, 1 thread: time step + LHS + Step 1 + Step 2 RHS + Step 3 + norms once "
                        f"({t_pre + t_post:.2f} s) and {r.iterations} FP64 PCG iterations "
                        f"({t_cg:.2f} s, {per_iter * 1e3:.1f} ms/iteration) on the same mesh, "
                        f"scaled by {iters_per_step:.1f} PCG iterations per step (GPU mean)"))


def run_probe(args) -> int:
    """--probe: a short fixed run for profilers (rocprofv3 counter passes),
    with no JSON line.  Builds the workload's context, advances --steps CBS
    steps with the --precision form (the first solve also runs the apply
    autotuner; amg's first solve also builds the hierarchy), then optionally
    solves --multi-rhs right-hand sides together (b ~ N(0,1) per node minus
    its mean, seed 3) on the last step's operator, then runs the plain-copy
    kernel (k_copy) over --copy-probe-bytes: its counters calibrate the
    HBM-byte counters against a kernel of known traffic (read + write of that
    many bytes each).  Prints one line per stage."""
    from cbs3d import AMG, FP64, MIXED, CaseSetup, HipContext, cases

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
        B = np.random.default_rng(3).standard_normal((args.multi_rhs, ctx.npoin))
        B -= B.mean(axis=1, keepdims=True)
        mprec = MIXED if args.precision == "mixed" else FP64
        t0 = time.perf_counter()
        _, res = ctx.solvePressureMulti(B, precision=mprec)
        its = [r.iterations for r in res]
        print(f"[probe] {args.multi_rhs}-RHS PCG ({'mixed' if mprec == MIXED else 'fp64'}): "
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
    ap.add_argument("--ensemble", type=int, default=0,
                    help="also time a batch of this many seeded cavities against one context at a time (64: "
                         "the ensemble workload)")
    ap.add_argument("--ensemble-steps", type=int, default=3)
    ap.add_argument("--adjoint", action="store_true",
                    help="also time the adjoint workload and check it against central differences")
    ap.add_argument("--amg", action="store_true",
                    help="also time the multigrid-preconditioned form on the same workload")
    ap.add_argument("--multi-rhs", type=int, default=0,
                    help="also solve this many random right-hand sides together on the workload's operator")
    ap.add_argument("--setup", action="store_true",
                    help="also measure the host set-up phases against the device colouring")
    ap.add_argument("--output", type=int, default=0,
                    help="also measure solution output with a VTU every K steps (mixed form)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the PCG kernels one by one instead of replaying captured HIP graphs")
    ap.add_argument("--copy-probe-bytes", type=int, default=1 << 30,
                    help="size of the plain-copy calibration kernel (0 to skip)")
    ap.add_argument("--probe", action="store_true",
                    help="short fixed run for profilers: --steps steps of --precision, then --multi-rhs and the "
                         "copy kernel; prints stage lines and no JSON")
    ap.add_argument("--precision", choices=("fp64", "mixed", "amg"), default="mixed",
                    help="solver form of --probe")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              f"(the hot path is single-GPU in this build)", file=sys.stderr)
        return 2
    if args.probe:
        return run_probe(args)

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
    b64 = op["nnz"] * (8 + 2) + N * 3 * 8
    b32 = op["nnz"] * (4 + 2) + N * 3 * 4
    apply64 = fp64["apply_ms"] / max(fp64["apply_n"], 1)
    apply32 = mixed["apply_ms"] / max(mixed["apply_n"], 1)
    ach64 = b64 / (apply64 * 1e-3) / 1e9
    traffic = measured_traffic(args.workload)
    roofline = {"bound": "hbm", "achieved": round(ach64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach64 / HBM_PEAK_GBS, 4),
                "traffic": traffic.get("fp64_apply_hbm_bytes") if traffic else None,
                "traffic_source": traffic.get("source") if traffic else None,
                "kernel": "k_apply_sell<double,true,B> (FP64 PCG apply, assembled operator)",
                "ms_per_apply": round(apply64, 5), "bytes_per_apply": b64,
                "bytes_per_apply_element_form": F * (6 * 8 + 16) + N * 3 * 8, "timed_applies": fp64["apply_n"],
                "launch": fp64["tile"], "operator": op}
    fp32_roof = {"kernel": "k_apply_sell<float,true,B> (FP32 apply of the mixed form, assembled operator)",
                 "ms_per_apply": round(apply32, 5) if mixed["apply_n"] else None,
                 "bytes_per_apply": b32, "timed_applies": mixed["apply_n"], "launch": mixed["tile"],
                 "traffic": traffic.get("fp32_apply_hbm_bytes") if traffic else None,
                 "achieved_gbs": round(b32 / (apply32 * 1e-3) / 1e9, 1) if mixed["apply_n"] else None}
    fp32_roof["frac"] = round(fp32_roof["achieved_gbs"] / HBM_PEAK_GBS, 4) if fp32_roof["achieved_gbs"] else None

    amg = run_amg(ctx, init, args.warmup, args.steps, fp64["fields"]) if args.amg else None
    multi = run_multi_rhs(ctx, args.multi_rhs) if args.multi_rhs > 0 else None
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
            cpu = cpu_baseline(setup, ctx, float(np.mean(fp64["iters"])), args.cpu_sample_iters)
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
    if args.ensemble > 0:
        ctx.close()
        line["ensemble"] = run_ensemble(args.ensemble, args.ensemble_steps)
    if args.adjoint:
        ctx.close()
        line["adjoint"] = run_adjoint()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
