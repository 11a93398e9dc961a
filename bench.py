"""Benchmark of the predicate and voxelization path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-sdf]

A step is one pass of the hot path over one batch: exact orient3d signs for the
2^24-tuple batch (90 % uniform in [-1,1]^3, 10 % exactly coplanar or 1-ulp off), inputs
resident in HBM as 12 float64 planes. `value` is whole-job signs/s over the K timed
steps. The JSON line also carries:
  roofline      the filter kernel (the dominant one) against the 8 TB/s HBM peak,
                97 algorithmic bytes per tuple (96 B of FP64 coordinates in, 1 B out),
                timed with HIP events on its own stream; `traffic` is the HBM bytes per
                launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE (scaled by a 1 GiB
                copy), read from the summary committed under profiles/
  cpu_baseline  the CPU oracle (a restatement of the reference's scalar predicates)
                over the same batch on one host core
  extra         insphere (2^22 tuples) signs/s, per-stage lane fractions, the SDF
                workloads (forward+backward voxels/s, batched vs plain launches), the
                PLC-carve classification (points/s and tets/s, both modes), binary
                STL input decoded and merged on the GPU, the mesh quality report,
                batched point location and Delaunay cavities
All other output goes to stderr; stdout holds exactly one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ORIENT3D_BYTES = 97   # 12 FP64 coordinates in + 1 int8 sign out
INSPHERE_BYTES = 121  # 15 FP64 coordinates in + 1 int8 sign out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_steps(fn, steps, warmup, sync):
    import torch
    for _ in range(warmup):
        fn()
    sync()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    sync()
    return start.elapsed_time(stop) / steps  # ms per step


def time_steps_streams(fn, steps, warmup, streams):
    """time_steps for work spread over several streams: the timed region starts on
    streams[0] before any of them runs a timed step and ends after all of them finish."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    main = streams[0]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(main)
    for s in streams[1:]:
        s.wait_stream(main)
    for _ in range(steps):
        fn()
    for s in streams[1:]:
        main.wait_stream(s)
    stop.record(main)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def bench_predicates(cm, W, steps, warmup):
    import torch
    n = 1 << 24
    log(f"[bench] generating {n} orient3d tuples")
    soa_np, kind = W.tuples_orient3d(n, seed=2)
    soa = torch.from_numpy(soa_np).cuda()
    out = torch.empty(n, dtype=torch.int8, device="cuda")
    ws = torch.empty(cm.pred_workspace_size(n, 4), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        cm.signs_soa(soa, "exact", out=out, workspace=ws, stream=stream)

    # A stream of batches: consecutive batches alternate between two default-priority
    # streams, each with its own workspace and output, so one batch's exact passes (a few
    # latency-bound waves) run beside the next batch's bandwidth-bound filter.
    stream2 = torch.cuda.Stream()
    out2 = torch.empty_like(out)
    ws2 = torch.empty_like(ws)
    lanes = [(stream, out, ws), (stream2, out2, ws2)]
    turn = [0]

    def pipelined_step():
        st, o, w = lanes[turn[0] & 1]
        turn[0] += 1
        cm.signs_soa(soa, "exact", out=o, workspace=w, stream=st)

    ms = time_steps_streams(pipelined_step, steps, warmup, [stream, stream2])
    single_ms = time_steps(step, steps, warmup, torch.cuda.synchronize)
    if not torch.equal(out, out2):
        raise RuntimeError("orient3d: the two streams' signs differ")
    del out2, ws2
    # dominant kernel, timed live with HIP events on the launch stream
    cm.kernel_timer("orient3d_filter", True)
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    kms, cnt = cm.kernel_time("orient3d_filter")
    cm.kernel_timer("orient3d_filter", False)
    filt_ms = kms / max(cnt, 1)
    stats = cm.pred_stats(n, 4, workspace=ws, stream=stream)
    signs = out.cpu().numpy()
    # host-resident batch (pinned): copy in, signs, copy out, per batch (PCIe included)
    host = torch.from_numpy(soa_np).pin_memory()
    host_out = torch.empty(n, dtype=torch.int8).pin_memory()

    def host_step():
        soa.copy_(host, non_blocking=True)
        step()
        host_out.copy_(out, non_blocking=True)

    pcie_ms = time_steps(host_step, 3, 1, torch.cuda.synchronize)
    del host, host_out
    res = {"n": n, "ms": ms, "single_ms": single_ms, "filter_ms": filt_ms, "stats": stats,
           "soa_np": soa_np,
           "signs": signs, "pcie_ms": pcie_ms}
    del soa, ws
    # insphere
    n5 = 1 << 22
    s5_np, _ = W.tuples_insphere(n5, seed=22)
    s5 = torch.from_numpy(s5_np).cuda()
    o5 = torch.empty(n5, dtype=torch.int8, device="cuda")
    ws5 = torch.empty(cm.pred_workspace_size(n5, 5), dtype=torch.uint8, device="cuda")

    def step5():
        cm.signs_soa(s5, "exact", out=o5, workspace=ws5, stream=stream)

    o5b = torch.empty_like(o5)
    ws5b = torch.empty_like(ws5)
    lanes5 = [(stream, o5, ws5), (stream2, o5b, ws5b)]

    def pipelined_step5():
        st, o, w = lanes5[turn[0] & 1]
        turn[0] += 1
        cm.signs_soa(s5, "exact", out=o, workspace=w, stream=st)

    ms5 = time_steps_streams(pipelined_step5, steps, warmup, [stream, stream2])
    single5 = time_steps(step5, steps, warmup, torch.cuda.synchronize)
    if not torch.equal(o5, o5b):
        raise RuntimeError("insphere: the two streams' signs differ")
    del o5b, ws5b
    cm.kernel_timer("insphere_filter", True)
    for _ in range(steps):
        step5()
    torch.cuda.synchronize()
    k5, c5 = cm.kernel_time("insphere_filter")
    cm.kernel_timer("insphere_filter", False)
    res["insphere"] = {"n": n5, "ms": ms5, "single_ms": single5, "filter_ms": k5 / max(c5, 1),
                       "stats": cm.pred_stats(n5, 5, workspace=ws5, stream=stream)}
    return res


def measured_traffic(n):
    """HBM bytes per launch of the filter kernel from the rocprofv3 PMC summary committed
    under profiles/ (FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy), when it covers
    this workload; {} otherwise."""
    path = os.path.join(ROOT, "profiles", "orient3d_2e24_filter_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    if t.get("algorithmic_bytes_per_launch") != ORIENT3D_BYTES * n:
        return {}
    return t


def cpu_baseline_predicates(soa_np, gpu_signs):
    import oracle
    t = time.perf_counter()
    want = oracle.orient3d_batch(soa_np, oracle.EXACT, nthreads=1)
    dt = time.perf_counter() - t
    if not (want == gpu_signs).all():
        raise SystemExit("bench: GPU signs differ from the CPU oracle")
    n = soa_np.shape[1]
    return {"value": n / dt, "unit": "signs/s", "cores": 1, "kind": "port",
            "sample": f"the full {n}-tuple exact orient3d batch, one thread "
                      f"({cpu_model()}); the GPU signs were checked equal to it",
            "seconds": dt}


def harness_threads() -> int:
    """Host threads for the harness-level parallel CPU baselines: the job's CPU share
    (OMP_NUM_THREADS when set), else every core."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def cpu_harness_predicates(soa_np, gpu_signs):
    """C2 with the CPU oracle over batch slices on `harness_threads()` threads (a harness
    around the reference's single-threaded predicate; its calls are independent)."""
    import oracle
    nt = harness_threads()
    t = time.perf_counter()
    want = oracle.orient3d_batch(soa_np, oracle.EXACT, nthreads=nt)
    dt = time.perf_counter() - t
    if not (want == gpu_signs).all():
        raise SystemExit("bench: GPU signs differ from the CPU oracle")
    n = soa_np.shape[1]
    return {"value": n / dt, "unit": "signs/s", "cores": nt, "kind": "port",
            "harness": True, "seconds": dt,
            "sample": f"the full {n}-tuple exact orient3d batch split over {nt} threads "
                      f"(harness-level parallel run; {cpu_model()})"}


def cpu_harness_c5(W, cases, budget_s=20.0):
    """C5 on the CPU: the reference's SDF search (3-D buckets + Chebyshev shells, restated
    in the oracle) over an evenly strided sample of every mesh's cells, one mesh per task
    on `harness_threads()` threads (a harness: the reference is single-threaded and each
    mesh is independent). Every task rebuilds its mesh's occupancy and buckets; a pass
    over one cell per mesh times those builds, and the sample is sized so the search
    takes about budget_s seconds."""
    import concurrent.futures as cf
    import oracle
    nt = harness_threads()
    grids = [oracle.make_grid(c.points, c.resolution, c.pad) for c in cases]
    total = sum(g.ncell for g in grids)

    def task(m, frac):
        c, g = cases[m], grids[m]
        occ = oracle.occupancy(c.points, c.tris, g, nthreads=1)
        n = max(1, int(g.ncell * frac))
        stride = max(1, g.ncell // n)
        oracle.sdf_shell(c.points, c.tris, g, occ, first=stride // 2, count=n, stride=stride)
        return n

    def run_all(frac):
        t = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=nt) as ex:
            n = sum(ex.map(lambda m: task(m, frac), range(len(cases))))
        return n, time.perf_counter() - t

    _, t_build = run_all(0.0)                 # one cell per mesh: the builds
    n_probe, t_probe = run_all(4e-6)
    per_cell = max(t_probe - t_build, 1e-3) / n_probe
    frac = min(0.01, budget_s / per_cell / total)
    done, dt = run_all(frac)
    return {"value": done / dt, "unit": "voxels/s (forward SDF only)", "cores": nt,
            "kind": "port", "harness": True, "seconds": dt, "cells_timed": done,
            "build_seconds": t_build,
            "search_only_voxels_per_s": done / max(dt - t_build, 1e-9),
            "sample": f"{done} of the {total} C5 cells ({frac:.2e} of each mesh's, evenly "
                      f"strided), reference shell search, one mesh per task on {nt} "
                      f"threads, occupancy and bucket builds included ({t_build:.1f} s of "
                      f"the {dt:.1f} s; {cpu_model()})"}


def cpu_sdf_sample(W, budget_s=20.0):
    """The reference's own SDF search (3-D buckets + Chebyshev shells, restated in the
    oracle) on one core, over cells spread evenly through the C3 grid (256^3 torus),
    sized to about budget_s seconds. The bucket build is timed separately (a run over
    one cell) and taken out of the per-voxel rate."""
    import oracle
    case = W.torus_c3(seed=3)
    g = oracle.make_grid(case.points, case.resolution, case.pad)
    occ = oracle.occupancy(case.points, case.tris, g, nthreads=1)

    def timed(count, stride):
        t = time.perf_counter()
        oracle.sdf_shell(case.points, case.tris, g, occ, first=stride // 2, count=count,
                         stride=stride)
        return time.perf_counter() - t

    t_build = min(timed(1, g.ncell), timed(1, g.ncell))
    probe = 256
    t_probe = max(timed(probe, g.ncell // probe) - t_build, 1e-6)
    n = int(min(g.ncell // 4, max(probe, probe * budget_s / t_probe)))
    stride = g.ncell // n
    t_run = timed(n, stride)
    rate = n / max(t_run - t_build, 1e-9)
    return {"voxels_per_s": rate, "cells_timed": n, "seconds": t_run,
            "bucket_build_s": t_build, "cores": 1, "kind": "port",
            "sample": f"{n} cells of the 256^3 C3 grid at stride {stride} (evenly spread), "
                      f"reference shell search, one thread, bucket build "
                      f"{t_build:.2f} s excluded ({cpu_model()})"}


# FP32 operations per lane of the forward voxel pass's three kinds of work (FMA = 2): an
# oriented-box lower bound (tree node) ~26, an oriented-box lower bound (triangle) ~30, a
# point-triangle distance ~80.
FLOP_BOX, FLOP_DISK, FLOP_TRI = 26, 30, 80
FP32_PEAK_TFLOPS = 157.3
# Algorithmic FP32 work of the SDF (the survey's count): 48 flops per point-triangle pair
# evaluated, plus 20 per cell in the backward pass.
ALG_FLOP_PAIR, ALG_FLOP_BWD_CELL = 48, 20


def voxel_work(counters, cells, vpl, fwd_ms, bwd_ms):
    """Work of the forward voxel pass from the engine's counters (each counted once per
    wavefront of 64 lanes x vpl voxels), the FP32 rate it implies, and the SDF roofline
    by the algorithmic count (48 flops per evaluated point-triangle pair + 20 per cell in
    the backward) over the forward + backward voxel kernels' time."""
    lanes = 64 * vpl
    box = counters["node_tests"] * lanes
    disk = counters["tri_tests"] * lanes
    tri = counters["tri_evals"] * lanes
    flops = box * FLOP_BOX + disk * FLOP_DISK + tri * FLOP_TRI
    out = {"node_tests_per_voxel": box / cells,
           "triangle_bound_tests_per_voxel": disk / cells,
           "triangle_distances_per_voxel": tri / cells,
           "counted_fp32_gflop_per_step": flops / 1e9}
    if fwd_ms > 0:
        tf = flops / (fwd_ms / 1e3) / 1e12
        out["fwd_voxel_kernel_fp32_tflops"] = tf
        out["fwd_voxel_kernel_fp32_frac_of_peak"] = tf / FP32_PEAK_TFLOPS
    alg = tri * ALG_FLOP_PAIR + cells * ALG_FLOP_BWD_CELL
    ms = fwd_ms + bwd_ms
    if ms > 0:
        a = alg / (ms / 1e3) / 1e12
        out["roofline"] = {"bound": "fp32", "achieved": a, "peak": FP32_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": a / FP32_PEAK_TFLOPS,
                           "algorithmic_flop_per_step": alg,
                           "count": "48 x point-triangle pairs evaluated (engine counter) "
                                    "+ 20 x cells (backward)",
                           "kernels": "vox_kernel (forward) + adjoint_kernel (backward)",
                           "kernel_ms": ms}
    return out


def bench_sdf(cm, W, steps, warmup, variant=None, full=True):
    """C5: 64 uneven meshes, SDF forward+backward per step, batched (one graph of
    persistent kernels) vs plain (one mesh at a time, synchronised per mesh, same kernels).
    Also C3/C4 single-mesh forward throughput. variant=(vpl, depth) skips the autotuner;
    full=False times the batched path only (for profiling its kernels)."""
    import torch
    from cybermeshgenerator_amd.batch import SdfBatch
    log("[bench] generating C5 meshes")
    cases = W.batch_c5(seed=5)
    meshes = [(c.points, c.tris) for c in cases]
    res = [c.resolution for c in cases]
    pads = [c.pad for c in cases]
    t0 = time.perf_counter()
    batch = SdfBatch(meshes, res, pads)
    plan_s = time.perf_counter() - t0
    cells = batch.cells
    tris = int(sum(c.tris.shape[0] for c in cases))
    gs = torch.from_numpy(W.upstream_grad(cells, seed=50)).cuda()
    sdf = torch.empty(cells, dtype=torch.float32, device="cuda")
    grad = torch.empty((batch.points.shape[0], 3), dtype=torch.float32, device="cuda")

    if variant is None:
        log("[bench] autotuning the voxel passes")
        tune = batch.autotune(gs, sdf=sdf, grad=grad)
    else:
        batch.set_variant(*variant)
        tune = {"vpl": variant[0], "depth": variant[1], "ms": None}

    def bstep():
        batch.step(gs, sdf=sdf, grad=grad, graph=True)

    reps = max(1, steps // 4)
    ms_b = time_steps(bstep, reps, max(1, warmup), torch.cuda.synchronize)
    st = batch.counters()
    # dominant kernel: the forward voxel pass
    timers = ("batch_sdf", "batch_sdf_backward", "batch_occupancy")
    for t in timers:
        cm.kernel_timer(t, True)
    batch.step(gs, sdf=sdf, grad=grad, graph=False)
    torch.cuda.synchronize()
    fwd_ms, _ = cm.kernel_time("batch_sdf")
    bwd_ms, _ = cm.kernel_time("batch_sdf_backward")
    occ_ms, _ = cm.kernel_time("batch_occupancy")
    for t in timers:
        cm.kernel_timer(t, False)
    if not full:
        out = {"meshes": len(cases), "cells": cells, "triangles": tris, "batched_ms": ms_b,
               "batched_voxels_per_s": cells / (ms_b / 1e3), "fwd_voxel_kernel_ms": fwd_ms,
               "bwd_voxel_kernel_ms": bwd_ms, "occupancy_kernel_ms": occ_ms,
               "variant": batch.variant_info()}
        out.update(voxel_work(st, cells, batch.vpl(), fwd_ms, bwd_ms))
        batch.close()
        return out
    # plain: one engine per mesh, one mesh at a time
    log("[bench] planning 64 single-mesh engines")
    singles = [SdfBatch([m], [r], [p]) for m, r, p in zip(meshes, res, pads)]
    for b in singles:  # the same voxel-pass variant as the batch
        b.set_variant(tune["vpl"], tune["depth"])
    sg = [gs[bg.offset:bg.offset + bg.ncell].contiguous() for bg in batch.grids]
    souts = [(torch.empty(b.cells, dtype=torch.float32, device="cuda"),
              torch.empty((b.points.shape[0], 3), dtype=torch.float32, device="cuda"))
             for b in singles]

    def pstep():
        for b, g, (o, gr) in zip(singles, sg, souts):
            b.step(g, sdf=o, grad=gr, graph=False)
            torch.cuda.synchronize()

    ms_p = time_steps(pstep, max(1, reps // 2), 1, torch.cuda.synchronize)
    # the two paths agree
    plain = torch.cat([o for o, _ in souts])
    same = bool(torch.equal(plain, sdf))
    for b in singles:
        b.close()
    out = {"meshes": len(cases), "cells": cells, "triangles": tris,
           "batched_ms": ms_b, "plain_ms": ms_p,
           "batched_voxels_per_s": cells / (ms_b / 1e3),
           "plain_voxels_per_s": cells / (ms_p / 1e3),
           "speedup_batched_vs_plain": ms_p / ms_b,
           "batched_equals_plain": same, "plan_s": plan_s,
           "fwd_voxel_kernel_ms": fwd_ms, "bwd_voxel_kernel_ms": bwd_ms,
           "occupancy_kernel_ms": occ_ms,
           "variant": batch.variant_info(), "autotune_ms": tune["ms"],
           "overflow": st["overflow"]}
    out.update(voxel_work(st, cells, batch.vpl(), fwd_ms, bwd_ms))
    batch.close()
    # single-mesh C3 and C4 forward
    for name, case in (("c3_torus_256", W.torus_c3(seed=3)), ("c4_harmonic_128", W.harmonic_c4(seed=4))):
        b = SdfBatch([(case.points, case.tris)], [case.resolution], [case.pad])
        o = torch.empty(b.cells, dtype=torch.float32, device="cuda")
        ms = time_steps(lambda: b.forward(out=o, graph=True), max(1, reps), 1,
                        torch.cuda.synchronize)
        g1 = torch.from_numpy(W.upstream_grad(b.cells, seed=40)).cuda()
        gr1 = torch.empty((b.points.shape[0], 3), dtype=torch.float32, device="cuda")
        ms2 = time_steps(lambda: b.step(g1, sdf=o, grad=gr1, graph=True), max(1, reps), 1,
                         torch.cuda.synchronize)
        out[name] = {"cells": b.cells, "fwd_ms": ms, "fwd_voxels_per_s": b.cells / (ms / 1e3),
                     "fwd_bwd_ms": ms2, "fwd_bwd_voxels_per_s": b.cells / (ms2 / 1e3),
                     "counters": b.counters()}
        b.close()
    return out


def bench_carve(cm, W, steps, warmup, cpu=True):
    """PLC-carve classification on the C3 torus (65,536 triangles, 256 x 256 projection
    grids): 2^22 query points uniform in its bounding box padded by 10 % (rays and parity
    modes) and 2^22 tets whose centroids are classified (rays), all resident in HBM."""
    import torch
    from cybermeshgenerator_amd.carve import CarveGrid
    case = W.torus_c3(seed=3)
    pts = torch.tensor(case.points, device="cuda")
    tris = torch.tensor(case.tris, device="cuda")
    rng = np.random.default_rng(31)
    lo, hi = case.points.min(0), case.points.max(0)
    span = hi - lo
    n = 1 << 22
    q_np = rng.uniform(lo - 0.1 * span, hi + 0.1 * span, (n, 3))
    q = torch.tensor(q_np, device="cuda")
    tets = torch.tensor(rng.integers(0, len(case.points), (n, 4)).astype(np.int32),
                        device="cuda")
    out = torch.empty(n, dtype=torch.uint8, device="cuda")
    res = {"mesh": "C3 torus, 65,536 triangles", "queries": n}
    for mode in ("rays", "parity"):
        CarveGrid(pts, tris, mode=mode)  # the first launches of the process load code
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cg = CarveGrid(pts, tris, mode=mode)
        setup_ms = (time.perf_counter() - t0) * 1e3
        cm.kernel_timer("carve_query", True)
        ms = time_steps(lambda: cg.points_inside(q, out), steps, warmup,
                        torch.cuda.synchronize)
        k_ms, k_n = cm.kernel_time("carve_query")
        cm.kernel_timer("carve_query", False)
        r = {"setup_ms": setup_ms, "ms_per_batch": ms, "points_per_s": n / (ms / 1e3),
             "query_kernel_ms": k_ms / max(k_n, 1),
             "inside_fraction": float(out.float().mean().item())}
        if mode == "rays":
            tms = time_steps(lambda: cg.tets_inside(tets, out), steps, warmup,
                             torch.cuda.synchronize)
            r["tets_ms_per_batch"] = tms
            r["tets_per_s"] = n / (tms / 1e3)
        res[mode] = r
        log(f"[bench] carve {mode}: {n / (ms / 1e3) / 1e9:.2f} G points/s "
            f"({ms:.3f} ms per 2^22), set-up {setup_ms:.1f} ms")
        del cg
    if cpu:
        import oracle
        cg = oracle.Carve(case.points, case.tris)
        m = 1 << 16
        t0 = time.perf_counter()
        want = cg.points(q_np[:m], nthreads=1)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": m / dt, "unit": "points/s", "cores": 1,
                               "kind": "port", "sample": f"{m} of the query points, rays"}
        got = CarveGrid(pts, tris, "rays").points_inside(q[:m]).cpu().numpy()
        res["cpu_baseline"]["identical"] = bool(np.array_equal(got, want))
    return res


def bench_quality(cm, W, steps, warmup, cpu=True):
    """Mesh quality report over a jittered Kuhn lattice of the unit cube: 128^3 cells, 6
    tets each (12,582,912 tets, 2,146,689 points) resident in HBM; the report and its 10
    worst tets per step."""
    import torch
    from cybermeshgenerator_amd import quality
    pts_np, tets_np = W.lattice_tets(128, jitter=0.3, seed=8)
    pts = torch.tensor(pts_np, device="cuda")
    tets = torch.tensor(tets_np, device="cuda")
    n = len(tets_np)
    cm.kernel_timer("quality", True)
    ms = time_steps(lambda: quality.report(pts, tets, 10), steps, warmup,
                    torch.cuda.synchronize)
    k_ms, k_n = cm.kernel_time("quality")
    cm.kernel_timer("quality", False)
    res = {"mesh": "Kuhn lattice 128^3 x 6, jitter 0.3", "tets": n, "ms_per_report": ms,
           "tets_per_s": n / (ms / 1e3), "scan_kernel_ms": k_ms / max(k_n, 1)}
    log(f"[bench] quality: {n / (ms / 1e3) / 1e9:.3f} G tets/s ({ms:.2f} ms per report, "
        f"scan kernel {res['scan_kernel_ms']:.3f} ms)")
    if cpu:
        import oracle
        m = 1 << 18
        t0 = time.perf_counter()
        c = oracle.quality_report(pts_np, tets_np[:m], 10)
        dt = time.perf_counter() - t0
        g = quality.report(pts, tets[:m], 10)
        res["cpu_baseline"] = {"value": m / dt, "unit": "tets/s", "cores": 1, "kind": "port",
                               "sample": "first 262,144 tets of the lattice",
                               "same_worst": g["worst_tets"] == c["worst_tets"]}
    return res


def bench_locate(cm, W, steps, warmup, cpu=True):
    """Batched point location: 2^20 points uniform in the unit cube located in a jittered
    Kuhn lattice of it (64^3 cells x 6 = 1,572,864 tets, jitter 0.1 cell), every walk
    starting at tet 0 (long walks: ~170 steps of up to four orientation tests)."""
    import torch
    from cybermeshgenerator_amd.locate import locate
    pts_np, tets_np = W.lattice_tets(64, jitter=0.1, seed=9)
    tets_np = W.positive_tets(pts_np, tets_np)
    nbrs_np = W.tet_neighbors(tets_np)
    n = 1 << 20
    q_np = np.random.default_rng(10).uniform(0, 1, (n, 3))
    pts, tets, nbrs, q = (torch.tensor(a, device="cuda") for a in (pts_np, tets_np, nbrs_np,
                                                                 q_np))
    last = {}

    def step():
        last["r"] = locate(pts, tets, nbrs, q, 0)

    ms = time_steps(step, steps, warmup, torch.cuda.synchronize)
    walk = last["r"][1].float()
    res = {"mesh": "Kuhn lattice 64^3 x 6, jitter 0.1", "tets": len(tets_np), "queries": n,
           "start": "tet 0 for every query", "ms_per_batch": ms,
           "queries_per_s": n / (ms / 1e3), "mean_walk_steps": float(walk.mean().item()),
           "brute_force_walks": last["r"][2]}
    log(f"[bench] locate: {n / (ms / 1e3) / 1e6:.1f} M queries/s ({ms:.2f} ms per 2^20, "
        f"{res['mean_walk_steps']:.0f} steps per walk)")
    if cpu:
        import oracle
        m = 1 << 15
        t0 = time.perf_counter()
        want, _ = oracle.locate(pts_np, tets_np, nbrs_np, q_np[:m], 0, nthreads=1)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": m / dt, "unit": "queries/s", "cores": 1,
                               "kind": "port", "sample": "first 32,768 queries",
                               "identical": bool(np.array_equal(
                                   last["r"][0][:m].cpu().numpy(), want))}
    res["cavity"] = bench_cavity(W, steps, warmup, cpu)
    return res


def bench_cavity(W, steps, warmup, cpu=True):
    """Delaunay cavities (the insertion's flood, incremental.cpp:148-161): 2^20 points
    uniform in the unit cube, located in a Delaunay tetrahedralization of 100,000 uniform
    points (scipy's Qhull), then every point's cavity from its tet, one batched call."""
    import torch
    from cybermeshgenerator_amd.locate import cavity, locate
    pts_np, tets_np, nbrs_np = W.delaunay_tets(100_000, seed=12)
    n = 1 << 20
    q_np = np.random.default_rng(13).uniform(0.001, 0.999, (n, 3))
    pts, tets, nbrs, q = (torch.tensor(a, device="cuda") for a in (pts_np, tets_np, nbrs_np,
                                                                 q_np))
    loc, _, _ = locate(pts, tets, nbrs, q, 0)
    cnt = torch.empty(n, dtype=torch.int32, device="cuda")
    cav = torch.full((n, 128), -1, dtype=torch.int32, device="cuda")
    last = {}

    def step():
        last["r"] = cavity(pts, tets, nbrs, q, loc, cap=128, count=cnt, out=cav)

    ms = time_steps(step, steps, warmup, torch.cuda.synchronize)
    count = last["r"][0]
    res = {"mesh": "Delaunay of 100,000 uniform points", "tets": len(tets_np), "queries": n,
           "ms_per_batch": ms, "cavities_per_s": n / (ms / 1e3),
           "mean_cavity_tets": float(count.float().mean().item()),
           "overflowed": int((count < 0).sum().item())}
    log(f"[bench] cavity: {n / (ms / 1e3) / 1e6:.1f} M cavities/s ({ms:.2f} ms per 2^20, "
        f"{res['mean_cavity_tets']:.1f} tets each)")
    if cpu:
        import oracle
        m = 1 << 15
        loc_np = loc[:m].cpu().numpy()
        t0 = time.perf_counter()
        want, want_cav = oracle.cavity(pts_np, tets_np, nbrs_np, q_np[:m], loc_np, cap=128,
                                       nthreads=1)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": m / dt, "unit": "cavities/s", "cores": 1,
                               "kind": "port", "sample": "first 32,768 queries",
                               "identical": bool(np.array_equal(count[:m].cpu().numpy(), want)
                                                 and np.array_equal(
                                                     last["r"][1][:m].cpu().numpy(),
                                                     want_cav))}
    return res


def bench_stl(cm, W, steps, warmup, cpu=True):
    """Binary STL to device arrays: a 4,194,304-triangle torus file (200 MiB) resident in
    HBM, decoded and merged (coincident vertices, first-occurrence numbering) into the
    float64 points and int32 triangles the voxelizer takes. One step = one call of
    cmg_hip_stl_decode (its work buffers are allocated and freed inside the call)."""
    import ctypes
    import torch
    from cybermeshgenerator_amd import files
    files._bind()
    data = W.binary_stl_bytes()
    n = (len(data) - 84) // 50
    dev = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
    xyz = torch.empty((3 * n, 3), dtype=torch.float64, device="cuda")
    tris = torch.empty((n, 3), dtype=torch.int32, device="cuda")
    npts = ctypes.c_int64(0)
    err = cm._err()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def step():
        cm._check(cm.lib().cmg_hip_stl_decode(
            ctypes.c_void_p(dev.data_ptr()), len(data), ctypes.c_void_p(xyz.data_ptr()),
            ctypes.c_void_p(tris.data_ptr()), ctypes.byref(npts), st, err, len(err)), err)

    ms = time_steps(step, steps, warmup, torch.cuda.synchronize)
    res = {"file": "binary STL, torus 2048 x 1024, 4,194,304 triangles",
           "bytes": len(data), "points": int(npts.value), "ms_per_file": ms,
           "triangles_per_s": n / (ms / 1e3), "input_gb_per_s": len(data) / (ms / 1e3) / 1e9}
    log(f"[bench] stl: {n / (ms / 1e3) / 1e9:.3f} G triangles/s ({ms:.2f} ms per file)")
    if cpu:
        import tempfile
        import oracle
        m = 1 << 18
        sample = W.binary_stl_bytes(512, 256)
        with tempfile.NamedTemporaryFile(suffix=".stl", delete=False) as f:
            f.write(sample)
        t0 = time.perf_counter()
        pts, otris, _ = oracle.read_stl(f.name)
        dt = time.perf_counter() - t0
        os.unlink(f.name)
        sd = torch.frombuffer(bytearray(sample), dtype=torch.uint8).cuda()
        gx, gt = files.stl_decode(sd)
        res["cpu_baseline"] = {
            "value": m / dt, "unit": "triangles/s", "cores": 1, "kind": "port",
            "sample": "262,144-triangle torus file read from disk and merged by the oracle",
            "identical": bool(np.array_equal(gx.cpu().numpy(), pts)
                              and np.array_equal(gt.cpu().numpy(), otris))}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sdf", action="store_true")
    ap.add_argument("--no-carve", action="store_true",
                    help="skip the carve, STL, quality and locate workloads")
    ap.add_argument("--carve-only", action="store_true",
                    help="time only the carve, STL, quality and locate workloads")
    ap.add_argument("--sdf-only", action="store_true",
                    help="time only the batched 64-mesh SDF step (no predicates, plain "
                         "path or single meshes) and print its figures: for profiling")
    ap.add_argument("--variant", default=None,
                    help="voxel-pass variant VPLxDEPTH (e.g. 1x8) instead of the autotuner")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="also run one device-to-device copy of 2^30 bytes (a kernel of "
                         "known traffic, to scale FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every box here has one GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    import cybermeshgenerator_amd as cm
    from cybermeshgenerator_amd import workloads as W
    assert cm.hip_available(), "no HIP device"
    log(f"[bench] device {torch.cuda.get_device_name(0)}")

    if args.calibration_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 1073741824 bytes read, 1073741824 bytes written")
        del src, dst

    variant = tuple(int(x) for x in args.variant.split("x")) if args.variant else None
    if args.sdf_only:
        print(json.dumps({"workload": "C5 batched SDF forward+backward, 64 meshes",
                          "sdf": bench_sdf(cm, W, args.steps, args.warmup, variant,
                                           full=False)}), flush=True)
        return

    if args.carve_only:
        print(json.dumps({"workload": "PLC carve, STL input, quality report, point location",
                          "carve": bench_carve(cm, W, args.steps, args.warmup,
                                               cpu=not args.no_cpu),
                          "stl": bench_stl(cm, W, args.steps, args.warmup,
                                           cpu=not args.no_cpu),
                          "quality": bench_quality(cm, W, args.steps, args.warmup,
                                                   cpu=not args.no_cpu),
                          "locate": bench_locate(cm, W, args.steps, args.warmup,
                                                 cpu=not args.no_cpu)}), flush=True)
        return

    pr = bench_predicates(cm, W, args.steps, args.warmup)
    n = pr["n"]
    value = n / (pr["ms"] / 1e3)
    achieved = ORIENT3D_BYTES * n / (pr["filter_ms"] / 1e3) / 1e9
    traffic = measured_traffic(n)
    st = pr["stats"]
    extra = {
        "orient3d": {"tuples": n, "ms_per_batch": pr["ms"], "streams": 2,
                     "single_stream_ms_per_batch": pr["single_ms"],
                     "single_stream_signs_per_s": n / (pr["single_ms"] / 1e3),
                     "filter_kernel_ms": pr["filter_ms"],
                     "host_resident_signs_per_s": n / (pr["pcie_ms"] / 1e3),
                     "host_resident_ms_per_batch": pr["pcie_ms"],
                     "stage_fraction": {"fp32": st["fp32"] / n, "fp64": st["fp64"] / n,
                                        "exact": st["exact"] / n,
                                        "exact_inline": st["exact_inline"] / n,
                                        "exact_pass": (st["exact"] - st["exact_inline"]) / n},
                     "wide_lanes": st["wide"], "failed_lanes": st["failed"]},
    }
    ins = pr["insphere"]
    n5 = ins["n"]
    s5 = ins["stats"]
    extra["insphere"] = {
        "tuples": n5, "signs_per_s": n5 / (ins["ms"] / 1e3), "ms_per_batch": ins["ms"],
        "streams": 2, "single_stream_ms_per_batch": ins["single_ms"],
        "single_stream_signs_per_s": n5 / (ins["single_ms"] / 1e3),
        "filter_kernel_ms": ins["filter_ms"],
        "filter_hbm_gbs": INSPHERE_BYTES * n5 / (ins["filter_ms"] / 1e3) / 1e9,
        "stage_fraction": {"fp32": s5["fp32"] / n5, "fp64": s5["fp64"] / n5,
                           "exact": s5["exact"] / n5,
                           "exact_fixed_point": s5["exact_inline"] / n5,
                           "exact_pass": (s5["exact"] - s5["exact_inline"]) / n5},
        "wide_lanes": s5["wide"], "failed_lanes": s5["failed"]}

    if not args.no_sdf:
        try:
            extra["sdf"] = bench_sdf(cm, W, args.steps, args.warmup, variant)
            if not args.no_cpu:
                log("[bench] CPU reference SDF search sample (1 thread)")
                extra["sdf"]["cpu_baseline"] = cpu_sdf_sample(W)
                log("[bench] CPU reference SDF search over C5, harness threads")
                extra["sdf"]["cpu_harness"] = cpu_harness_c5(W, W.batch_c5(seed=5))
        except Exception as e:  # reported, never hides the predicate line
            log(f"[bench] sdf workloads failed: {e!r}")
            extra["sdf"] = {"error": repr(e)}

    if not args.no_carve:
        try:
            extra["stl"] = bench_stl(cm, W, args.steps, args.warmup, not args.no_cpu)
        except Exception as e:  # reported, never hides the predicate line
            log(f"[bench] stl workload failed: {e!r}")
            extra["stl"] = {"error": repr(e)}
        try:
            extra["quality"] = bench_quality(cm, W, args.steps, args.warmup,
                                             not args.no_cpu)
        except Exception as e:  # reported, never hides the predicate line
            log(f"[bench] quality workload failed: {e!r}")
            extra["quality"] = {"error": repr(e)}
        try:
            extra["locate"] = bench_locate(cm, W, args.steps, args.warmup, not args.no_cpu)
        except Exception as e:  # reported, never hides the predicate line
            log(f"[bench] locate workload failed: {e!r}")
            extra["locate"] = {"error": repr(e)}
        try:
            extra["carve"] = bench_carve(cm, W, args.steps, args.warmup, not args.no_cpu)
        except Exception as e:  # reported, never hides the predicate line
            log(f"[bench] carve workload failed: {e!r}")
            extra["carve"] = {"error": repr(e)}

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU oracle baseline (1 thread)")
        cpu = cpu_baseline_predicates(pr["soa_np"], pr["signs"])
        log("[bench] CPU oracle, harness threads")
        extra["orient3d"]["cpu_harness"] = cpu_harness_predicates(pr["soa_np"], pr["signs"])

    line = {
        "metric": "SDF voxels/s fwd+bwd batched vs plain GPU path; orient3d signs/s, exact signs",
        "value": value,
        "unit": "signs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pr["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, 2^24 orient3d tuples: 90% uniform in [-1,1]^3, "
                "10% exactly coplanar or 1-ulp perturbed)",
        "config": {"workload": "batched exact orient3d, 2^24 tuples, FP64 SoA planes resident "
                               "in HBM, int8 signs", "tuples": n, "mode": "exact",
                   "parallelism": "single GPU",
                   "batches": "consecutive batches alternate between two streams"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic.get("traffic_bytes_per_launch"),
                     "traffic_source": traffic.get("source"),
                     "kernel": "filter_soa_kernel<orient3d, exact>",
                     "bytes_per_launch": ORIENT3D_BYTES * n},
        "cpu_baseline": ({k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}
                         if cpu else None),
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
