#!/usr/bin/env python3
"""Benchmark: non-rigid GN + block-PCG solve, ms per frame, on one MI355X.

Workload (default): the three-view body scene -- ~150k surfels, ~4k nodes,
3 x 640x480 depth + foreground, ~300 feature pairs per view, ~2000 cross-view
pairs -- solved with 5 Gauss-Newton x 10 PCG iterations in the mixed-precision
mode (FP32 Jacobians / J^T J blocks, FP64 dots, alpha/beta and 6x6 inverses,
one residual-refinement step per GN iteration).  A "step" is one batched solve
of --batch independent frames of that workload (distinct seeds), one launch per
stage for all of them: index build + all GN iterations + warp updates, with
every input already resident in HBM.  `value` is ms per frame (step time /
frames); the latency of one frame solved alone is reported beside it.

The JSON line carries the contract fields plus:
  roofline      the dominant kernel's algorithmic bytes / launch time vs 8 TB/s
  cpu_baseline  the CPU restatement of the solve (oracle/, FP32, 1 thread)
                timed on a bounded sample of frames of the same workload
  twist_rel_err max over GN iterations of |dx - dx_fp64| / |dx_fp64| against the
                FP64 host oracle from the same linearisation point
  plain_fp32    the plain FP32 GPU path (reference precision, no hipGraph) on
                the same frame: ms/frame and its twist error, for comparison
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOAD_NAMES = {
    0: "sphere_1view_320x240_1k_nodes",
    1: "body_3view_640x480_4k_nodes",
    3: "body_3view_640x480_8k_nodes_matrix_free",
    4: "uneven_16_scenes_1k-8k_nodes_1-3_views",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", type=int, default=1, choices=[0, 1, 3, 4],
                    help="1: the headline frame (default); 3: ~8k nodes, 15 PCG, matrix-free J^T J; "
                         "4: a batch of 16 uneven scenes (aggregate frames/s, p50 / p99 latency)")
    ap.add_argument("--precision", default="mixed", choices=["fp32", "mixed"])
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--cpu-frames", type=int, default=8, help="frames in the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stage-reps", type=int, default=50)
    ap.add_argument("--batch", type=int, default=32, help="frames solved per step (one launch per stage)")
    ap.add_argument("--lanes", type=int, default=0,
                    help="solve lanes of the batch (0: the library's default, or the autotuner's choice)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes (the autotuner changes no result, only workgroup counts)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the extra legs of the default run: the uneven 16-scene batch and the matrix-free "
                         "8k-node workload")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first: the known-size kernel that "
                         "calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def algorithmic_bytes(sz, mixed, matrix_free=False):
    """Minimum compulsory bytes per launch of each stage (format independent)."""
    N, P = sz["num_nodes"], sz["num_node_pairs"]
    Td, Tfg, E = sz["num_dense_terms"], sz["num_foreground_terms"], sz["num_smooth_terms"]
    Tf, Tc = sz["num_feature_terms"], sz["num_cross_terms"]
    kv1 = sz["num_node2term"]
    # term records: knn ids + weights + J + r (dense / foreground), edge ids + valid
    # + warped point + residual (smooth), knn + weights + target + warped (p2p)
    r_terms = 52 * (Td + Tfg) + 29 * E + 56 * (Tf + Tc)
    # preconditioner out: FP32 block inverses (plain), the FP64 Cholesky
    # factors L^-1 the fused CG scales by (mixed), FP32 + FP64 inverses
    # (matrix-free, mixed)
    inv = 144 * N + (288 * N if mixed else 0)
    if matrix_free:
        return {
            # diagonal blocks and J^T r only
            "assemble": r_terms + 4 * kv1 + 8 * (N + 1) + 16 * N + 144 * N + inv + 24 * N,
            # J^T (J d): term records, Node2Term lists, d in, q out
            "spmv": r_terms + 4 * kv1 + 4 * (N + 1) + 24 * N + 24 * N,
        }
    return {
        # Node2Term lists + offsets, pair keys + offsets, node positions in;
        # diagonal blocks, preconditioner, J^T r and off-diagonal blocks out
        "assemble": r_terms + 4 * kv1 + 8 * (N + 1) + 4 * P + 16 * N + 144 * N + (168 * N if mixed else 144 * N)
        + 24 * N + 144 * P,
        # blocks + block-row structure + the direction vector in, A d out
        "spmv": 144 * (N + P) + 4 * (N + P) + 8 * (N + 1) + 8 * P + 24 * N + 24 * N,
        # fused CG iteration: off-diagonal blocks of L^-1 A L^-T and the block-row
        # structure; r, w, s, p, x read once, r, w, s, p, x written once
        "cg_iteration": 144 * P + 12 * P + 8 * (N + 1) + 5 * 24 * N + 5 * 24 * N,
    }


STAGE_IDS = {"eval_terms": 0, "assemble": 1, "pcg_update": 2, "spmv": 3, "index": 4, "cg_iteration": 5}


def stage_time_ms(solver, stage, reps):
    import ctypes as C

    from multiviewdynrecon_amd.solver import _stream_handle, lib

    out = C.c_float()
    solver._check(lib().mvdr_solver_bench_stage(solver.h, STAGE_IDS[stage], reps, _stream_handle(None), C.byref(out)))
    return float(out.value)


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    return Oracle


def cpu_baseline_leg(scene, workload, frames):
    """CPU restatement (oracle/, FP32, one thread) timed on a bounded sample.  The
    oracle is test infrastructure: here it is only the timed baseline."""
    Oracle = _oracle()
    from multiviewdynrecon_amd.config import workload_config
    from multiviewdynrecon_amd.frame import HostFrame

    hf = HostFrame(scene)
    cfg32 = workload_config(workload, precision=0, refine_steps=0)
    times = []
    t_all = time.perf_counter()
    for _ in range(frames):
        o = Oracle(cfg32, hf, "fp32")
        t0 = time.perf_counter()
        o.solve()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all > 30.0:
            break
    cpu = {"value": float(np.median(times) * 1e3), "unit": "ms/frame", "cores": 1, "kind": "port",
           "sample": f"{len(times)} full solves (index + GN x PCG) of the benchmark frame, FP32, single thread, median",
           "cpu_model": cpu_model()}
    return cpu


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def baseline_threads():
    """Every CPU this process may run on, capped by OMP_NUM_THREADS when the
    environment sets a CPU share for the job."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(avail, int(share))) if share and share.isdigit() else avail


def cpu_baseline_threads(scene, workload, threads):
    """The same CPU restatement with one frame per thread (frames are independent;
    the oracle's calls release the GIL): aggregate ms per frame over 2 frames per
    thread."""
    from concurrent.futures import ThreadPoolExecutor

    Oracle = _oracle()
    from multiviewdynrecon_amd.config import workload_config
    from multiviewdynrecon_amd.frame import HostFrame

    hf = HostFrame(scene)
    cfg32 = workload_config(workload, precision=0, refine_steps=0)
    oracles = [Oracle(cfg32, hf, "fp32") for _ in range(2 * threads)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda o: o.solve(), oracles))
    wall = time.perf_counter() - t0
    return {"value": float(wall * 1e3 / len(oracles)), "unit": "ms/frame", "cores": threads, "kind": "port",
            "sample": f"{len(oracles)} full solves of the benchmark frame on {threads} threads, FP32, aggregate",
            "host_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": cpu_model()}


def twist_errors(scene, cfg, frame_dev):
    """Per-GN-iteration twist error of the GPU path against the FP64 oracle from
    the same linearisation point (the oracle is the checker only)."""
    Oracle = _oracle()
    from multiviewdynrecon_amd.frame import HostFrame
    from multiviewdynrecon_amd.solver import NonRigidSolver

    solver = NonRigidSolver(cfg)
    solver.SetSolverInput(frame_dev)
    solver.BuildSolverIndex()
    o64 = Oracle(cfg, HostFrame(scene), "fp64")
    o64.build_index()
    errs = []
    for k in range(cfg.gn_iterations):
        se3 = solver.solved_node_se3_host().astype(np.float64)
        solver.SolverIteration()
        tw = solver.debug_get("twist").astype(np.float64).ravel()
        o64.set("node_se3", se3)
        o64.set_iteration(k)
        o64.eval_terms()
        o64.assemble()
        o64.pcg(cfg.pcg_iterations)
        ref = o64.get("twist").ravel()
        errs.append(float(np.linalg.norm(tw - ref) / max(np.linalg.norm(ref), 1e-30)))
    return errs


def time_solves(solver, steps, warmup):
    import torch

    for _ in range(warmup):
        solver.SolveNonRigidAlignment()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.SolveNonRigidAlignment()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def measured_traffic(kernel, workload, frames):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated) for this many frames per
    launch, or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{WORKLOAD_NAMES[workload]}_x{frames}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def uneven_measure(batches, warmup, graph=1, solo_reps=5, autotune=True):
    """16 independent scenes, N ~ U[1000, 8000] nodes and 1-3 views each, solved
    as one batch (mixed precision, one hipGraph per batch): aggregate frames/s
    over `batches` batches, the per-scene latency (a scene completes with its
    batch) at p50 / p99, and each scene solved alone."""
    import torch

    from multiviewdynrecon_amd.config import workload_config
    from multiviewdynrecon_amd.frame import DeviceFrame
    from multiviewdynrecon_amd.scene import make_uneven_scenes, uneven_scene_params
    from multiviewdynrecon_amd.solver import NonRigidBatch, NonRigidSolver

    count = 16
    params = uneven_scene_params(count)
    nodes = [p["target_nodes"] for p in params]
    views = [p["num_views"] for p in params]
    scenes = make_uneven_scenes(count)
    cfg = workload_config(4, use_graph=graph, refine_steps=0)
    frames = [DeviceFrame(sc) for sc in scenes]
    solvers = []
    for fr in frames:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(fr)
        solvers.append(s)
    batch = NonRigidBatch(solvers)
    print(f"[bench] {count} uneven scenes resident: nodes {sorted(nodes)}, views {views}", file=sys.stderr, flush=True)
    batch.SolveNonRigidAlignment()
    tuned = dict(batch.autotune(reps=10), lanes=batch.lanes) if autotune else None
    for _ in range(warmup):
        batch.SolveNonRigidAlignment()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(batches)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record()
        batch.SolveNonRigidAlignment()
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    solo = [time_solves(s, solo_reps, 2) for s in solvers]
    out = {"frames_per_s": round(count * batches / wall, 2), "batches": batches, "ms_per_batch": round(wall * 1e3 / batches, 4),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)},
           "solo_latency_ms": {"p50": round(float(np.percentile(solo, 50)), 4),
                               "p99": round(float(np.percentile(solo, 99)), 4)},
           "nodes": nodes, "views": views, "gn_iterations": cfg.gn_iterations, "pcg_iterations": cfg.pcg_iterations,
           "launch_params": tuned}
    print(f"[bench] uneven batch: {out['frames_per_s']} frames/s, latency p50 {out['latency_ms']['p50']} "
          f"p99 {out['latency_ms']['p99']} ms", file=sys.stderr, flush=True)
    del batch, solvers, frames
    return out


def bench_uneven(args):
    """The uneven 16-scene batch as the benchmark line (aggregate frames/s)."""
    m = uneven_measure(max(args.steps, 100), args.warmup, args.graph, autotune=not args.no_autotune)
    out = {
        "metric": "aggregate frames/s, 16 uneven scenes per batch (1k-8k nodes, 1-3 views)",
        "value": m["frames_per_s"], "unit": "frames/s", "n_gpus": 1, "steps": m["batches"], "warmup": args.warmup,
        "ms_per_step": m["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 reductions / 6x6 factors)", "data": "synthetic (seeded scene generator)",
        "config": {"workload": WORKLOAD_NAMES[4], "scenes": 16, "nodes": m["nodes"], "views": m["views"],
                   "gn_iterations": m["gn_iterations"], "pcg_iterations": m["pcg_iterations"], "precision": "mixed",
                   "hip_graph": bool(args.graph)},
        "latency_ms": m["latency_ms"], "solo_latency_ms": m["solo_latency_ms"],
    }
    print(json.dumps(out), flush=True)


def matrix_free_measure(frames, steps, autotune=True):
    """The ~8k-node, 15-PCG matrix-free workload, mixed precision: ms per frame of
    a batched step of `frames` frames, and one frame alone."""
    import torch

    from multiviewdynrecon_amd.config import workload_config
    from multiviewdynrecon_amd.frame import DeviceFrame
    from multiviewdynrecon_amd.scene import make_scene
    from multiviewdynrecon_amd.solver import NonRigidBatch, NonRigidSolver

    cfg = workload_config(3, precision=1, refine_steps=0, use_graph=1)
    scenes = [make_scene(3)] + [make_scene(3, seed=2000 + k) for k in range(1, frames)]
    solvers = []
    for sc in scenes:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(DeviceFrame(sc))
        solvers.append(s)
    batch = NonRigidBatch(solvers)
    batch.SolveNonRigidAlignment()
    tuned = dict(batch.autotune(reps=5), lanes=batch.lanes) if autotune else None
    batch.SolveNonRigidAlignment()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.SolveNonRigidAlignment()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    single = time_solves(solvers[0], steps, 1)
    out = {"workload": WORKLOAD_NAMES[3], "frames_per_step": frames, "ms_per_frame": round(ms / frames, 4),
           "latency_ms_single_frame": round(single, 4), "nodes": scenes[0].num_nodes,
           "surfels": scenes[0].num_surfels, "gn_iterations": cfg.gn_iterations, "pcg_iterations": cfg.pcg_iterations,
           "launch_params": tuned}
    print(f"[bench] matrix-free 8k: {out['ms_per_frame']} ms/frame in batches of {frames}, "
          f"{out['latency_ms_single_frame']} ms alone", file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build solves on a single GPU; --gpus {args.gpus} is not supported "
              "(frames are independent and would be sharded one solver per GPU)", file=sys.stderr)
        sys.exit(2)
    import torch

    from multiviewdynrecon_amd.config import workload_config
    from multiviewdynrecon_amd.frame import DeviceFrame
    from multiviewdynrecon_amd.scene import make_scene
    from multiviewdynrecon_amd.solver import NonRigidBatch, NonRigidSolver

    if args.workload == 4:
        return bench_uneven(args)
    overrides = dict(use_graph=args.graph)
    mixed = args.precision == "mixed"
    if mixed:
        # the reference's schedule: GN x PCG, no residual refinement after the
        # PCG (the option changes the twist by < 3e-7 here and costs two
        # launches per GN iteration)
        overrides.update(precision=1, refine_steps=0)
    else:
        overrides.update(precision=0, refine_steps=0)
    cfg = workload_config(args.workload, **overrides)

    if args.copy_mib:
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy: {args.copy_mib} MiB read + {args.copy_mib} MiB written", file=sys.stderr)
        del src, dst

    # B independent frames of the workload (distinct seeds); frame 0 is the
    # default-seed scene that the latency, error and CPU legs use
    B = max(1, args.batch)
    scenes = [make_scene(args.workload)] + [make_scene(args.workload, seed=1000 + k) for k in range(1, B)]
    frames = [DeviceFrame(sc) for sc in scenes]
    scene, frame = scenes[0], frames[0]
    solvers = []
    for fr in frames:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(fr)
        solvers.append(s)
    batch = NonRigidBatch(solvers)
    if args.lanes:
        batch.set_lanes(args.lanes)
    print(f"[bench] {B} frames resident", file=sys.stderr, flush=True)
    tuned = {}
    if not args.no_autotune:
        # launch shapes for this batch and for a lone frame, before any timing
        batch.SolveNonRigidAlignment()
        tuned["batch"] = dict(batch.autotune(reps=10), lanes=batch.lanes)
        if args.lanes:
            batch.set_lanes(args.lanes)
            tuned["batch"]["lanes"] = args.lanes
        solvers[0].SolveNonRigidAlignment()
        tuned["single"] = solvers[0].autotune(reps=10)
        print(f"[bench] autotuned launch shapes: {json.dumps(tuned)}", file=sys.stderr, flush=True)

    # timed region: K steps, one step = one batched solve of the B frames
    for _ in range(args.warmup):
        batch.SolveNonRigidAlignment()
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        batch.SolveNonRigidAlignment()
        e1.record()
        per.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    event_ms = [a.elapsed_time(b) for a, b in per]
    ms_per_step = wall * 1e3 / args.steps
    ms_per_frame = ms_per_step / B
    print(f"[bench] {args.steps} steps of {B} frames: {ms_per_step:.3f} ms/step, {ms_per_frame:.4f} ms/frame "
          f"(event median {np.median(event_ms):.3f} ms/step)", file=sys.stderr, flush=True)

    # one frame alone: latency of a single solve
    single = solvers[0]
    latency_ms = time_solves(single, args.steps, 2)
    print(f"[bench] single-frame latency {latency_ms:.3f} ms", file=sys.stderr, flush=True)

    # per-stage kernel timing of the batched launches (hipEvents on the stream)
    mf = cfg.jtj_mode in (1, 2)
    if mf:
        per_solve = {"eval_terms": cfg.gn_iterations, "assemble": cfg.gn_iterations,
                     "spmv": cfg.gn_iterations * (cfg.pcg_iterations + 2 * cfg.refine_steps),
                     "pcg_update": cfg.gn_iterations * cfg.pcg_iterations, "index": 1}
    elif mixed:
        per_solve = {"eval_terms": cfg.gn_iterations, "assemble": cfg.gn_iterations,
                     "cg_iteration": cfg.gn_iterations * cfg.pcg_iterations, "index": 1}
    else:
        per_solve = {"eval_terms": cfg.gn_iterations, "assemble": cfg.gn_iterations,
                     "pcg_update": cfg.gn_iterations * cfg.pcg_iterations,
                     "spmv": cfg.gn_iterations * cfg.pcg_iterations, "index": 1}
    batch.SolveNonRigidAlignment()
    stages = {st: batch.bench_stage(STAGE_IDS[st], args.stage_reps) for st in per_solve}
    batch.SolveNonRigidAlignment()  # restore a consistent state after the stage loops
    torch.cuda.synchronize()
    all_sizes = [s.sizes() for s in solvers]
    sizes = all_sizes[0]
    ab = {}
    for sz in all_sizes:
        for k, v in algorithmic_bytes(sz, mixed, mf).items():
            ab[k] = ab.get(k, 0) + v
    share = {k: stages[k] * per_solve[k] for k in stages}
    print("[bench] stage ms/launch " + json.dumps({k: round(v, 4) for k, v in stages.items()}), file=sys.stderr)
    print("[bench] stage ms/step   " + json.dumps({k: round(v, 4) for k, v in share.items()}), file=sys.stderr)
    apply_stage = "cg_iteration" if mixed and not mf else "spmv"
    dominant = max(("assemble", apply_stage), key=lambda k: share[k])
    achieved = ab[dominant] / (stages[dominant] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(dominant, args.workload, B),
                "bytes_per_launch": int(ab[dominant]), "launch_ms": round(stages[dominant], 5), "frames_per_launch": B}
    apply_gbs = ab[apply_stage] / (stages[apply_stage] * 1e-3) / 1e9
    jtj_gbs = ab["assemble"] / (stages["assemble"] * 1e-3) / 1e9

    # the plain FP32 GPU path on frame 0 (reference precision, one frame, no graph)
    plain_cfg = workload_config(args.workload, precision=0, refine_steps=0, use_graph=0)
    plain = NonRigidSolver(plain_cfg)
    plain.SetSolverInput(frame)
    plain_ms = time_solves(plain, max(args.steps // 2, 1), 2)
    del plain
    print(f"[bench] plain FP32 path {plain_ms:.3f} ms/frame", file=sys.stderr, flush=True)

    cpu = cpu_mt = None
    errs = plain_errs = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_leg(scene, args.workload, args.cpu_frames)
        print(f"[bench] cpu baseline {cpu['value']:.1f} ms/frame", file=sys.stderr, flush=True)
        cpu_mt = cpu_baseline_threads(scene, args.workload, baseline_threads())
        print(f"[bench] cpu baseline, {cpu_mt['cores']} threads: {cpu_mt['value']:.1f} ms/frame", file=sys.stderr,
              flush=True)
        errs = twist_errors(scene, cfg, frame)
        plain_errs = twist_errors(scene, plain_cfg, frame)
        print(f"[bench] twist rel err per GN iteration: {args.precision} {errs} plain {plain_errs}",
              file=sys.stderr, flush=True)

    out = {
        "metric": ("non-rigid solve ms/frame (3 views, ~4k nodes) at <=1e-4 rel err vs FP64" if args.workload == 1 else
                   f"non-rigid solve ms/frame ({WORKLOAD_NAMES[args.workload]})"),
        "value": round(ms_per_frame, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 reductions / 6x6 factors)" if mixed else "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": WORKLOAD_NAMES[args.workload], "frames_per_step": B, "views": scene.num_views,
                   "width": scene.width, "height": scene.height, "nodes": scene.num_nodes,
                   "surfels": scene.num_surfels, "gn_iterations": cfg.gn_iterations,
                   "pcg_iterations": cfg.pcg_iterations, "precision": args.precision, "hip_graph": bool(args.graph),
                   "dense_terms": sizes["num_dense_terms"], "node_pairs": sizes["num_node_pairs"],
                   "parallelism": "single GPU, frames batched per launch"},
        "event_ms_median_per_step": round(float(np.median(event_ms)), 4),
        "latency_ms_single_frame": round(latency_ms, 4),
        "roofline": roofline,
        "apply_jtj_gbs": round(apply_gbs, 2),
        "jtj_assembly_gbs": round(jtj_gbs, 2),
        "stage_ms": {k: round(v, 5) for k, v in stages.items()},
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_mt,
        "twist_rel_err": max(errs) if errs else None,
        "launch_params": {"autotuned": not args.no_autotune, **tuned},
        "plain_fp32": {"ms_per_frame": round(plain_ms, 4), "speedup": round(plain_ms / ms_per_frame, 3),
                       "twist_rel_err": max(plain_errs) if plain_errs else None},
    }
    if args.workload == 1 and not args.no_extras:
        # the other two solve workloads at their real sizes, beside the headline
        del batch, solvers, frames
        out["uneven_batch"] = uneven_measure(30, 3, args.graph, solo_reps=3, autotune=not args.no_autotune)
        out["matrix_free_8k"] = matrix_free_measure(4, 5, autotune=not args.no_autotune)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
