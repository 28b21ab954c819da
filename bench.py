#!/usr/bin/env python3
"""Benchmark of the MI355X filtration front end (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the front end over one synthetic cloud already
resident in HBM. The default workload is the single-cloud configuration:
n = 20000 float32 points i.i.d. N(0, I_3), L2, full-layout distance matrix
(1.6 GB) with the enclosing radius fused in. `value` is the whole-job
front-end throughput in GB/s of algorithmic bytes (4 n^2 matrix bytes
written + 4 n d point bytes read per step).

Prints ONE JSON line. Besides the contract keys it carries:
  roofline        the dominant kernel's achieved algorithmic GB/s (HIP events that
                  the library records around the distance-tile launch alone, on
                  the stream it is launched on) against the 8 TB/s HBM peak;
                  step_frac: the same bytes over the whole step's events
  cpu_baseline    the reference's own make_dense + enclosing_radius (compiled
                  from its sources into oracle/_ref, SIMD build, all allotted
                  host cores) on the same cloud
  whole_run       persistence() of the same cloud at maxdim 1, end to end
  batched         the 512-cloud batched front end (device time, GB/s), whole
                  runs of all 512 clouds (diagrams/sec), and the batched form
                  against the plain per-cloud path on the clouds with n <= 2000
  batched_linf    the same set under L-inf: front end and one whole run
  differentiable  forward + backward of a top-8 H1 persistence loss on a
                  noisy circle (n = 2048), the backward kernel alone, and the
                  64-circle batch
  graph_stages    seconds per call of the exact kNN-15 graph and the radius
                  graph (n = 2e5, d = 8) and of Boruvka H0 (n = 1e6, d = 3)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "persistence diagrams/sec + distance-kernel HBM GB/s, batched vs plain GPU path"
HBM_PEAK_GBS = 8000.0
CPU_CORES_ALLOTTED = 16   # host-core share of one GPU on the benchmark box


def cloud(name):
    if name == "c2a":       # N(0, I_3), n = 20000, seed 2
        rng = np.random.default_rng(2)
        return rng.normal(size=(20000, 3)).astype(np.float32), "l2"
    if name.startswith("c2b"):   # U[0,1]^16, n = 20000, seed 3
        rng = np.random.default_rng(3)
        metric = {"c2b-l1": "l1", "c2b-linf": "linf", "c2b-lp": "lp:1.5"}[name]
        return rng.uniform(size=(20000, 16)).astype(np.float32), metric
    raise SystemExit(f"unknown workload {name}")


def traffic_from_profiles(kernel_tag):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy kernel), committed as text under
    profiles/. None if not measured for this kernel."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel_tag:
            return float(parts[1])
    return None


def cpu_baseline(x, metric, algo_bytes):
    """The reference's make_dense + enclosing_radius on the host cores."""
    from oracle import Reference, Restatement, reference_available
    threads = min(CPU_CORES_ALLOTTED, os.cpu_count() or 1)
    if reference_available(simd=True):
        ref = Reference(simd=True)
        t_make, t_rad, _ = ref.time_dense(x, metric, threads)
        secs = t_make + t_rad
        return {"value": round(algo_bytes / secs / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "reference",
                "sample": f"full workload once: make_dense {t_make:.2f} s + enclosing_radius {t_rad:.2f} s "
                          f"(reference -O3 x86-64-v3 SIMD build, {threads} threads)"}
    # Fallback: the scalar C restatement on one core, on a row sample.
    R = Restatement()
    m = 2000
    t0 = time.perf_counter()
    R.make_dense(x[:m], metric, True)
    secs = time.perf_counter() - t0
    per_byte = secs / (4.0 * m * m)
    return {"value": round(1.0 / per_byte / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"first {m} points, full layout, scalar restatement"}


def run_dense(args):
    import torch
    from turbotops_amd import dist_matrix_dev
    x, metric = cloud(args.workload)
    n, d = x.shape
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(x).to(dev)
    mat = torch.empty(n * n, dtype=torch.float32, device=dev)
    rowmax = torch.empty(n, dtype=torch.int32, device=dev)
    radius = torch.empty(1, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    # per step: events around the whole step (row-maxima fill, distance
    # tiles, radius reduction) and, recorded by the library on the same
    # stream, around the distance-tile kernel alone
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    kbeg = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    kend = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]

    def step(k=None):
        ev = None if k is None else (kbeg[k], kend[k])
        dist_matrix_dev(xt, metric, "full", out=mat, rowmax=rowmax, radius=radius, stream=stream, events=ev)

    for _ in range(args.warmup):
        step()
    for k in range(args.steps):   # torch creates an event's handle on its first record
        for e in (starts[k], ends[k], kbeg[k], kend[k]):
            e.record(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        step(k)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps
    kern_ms = sum(s.elapsed_time(e) for s, e in zip(kbeg, kend)) / args.steps

    algo_bytes = 4.0 * n * n + 4.0 * n * d
    ms_per_step = wall * 1e3 / args.steps
    value = algo_bytes / (wall / args.steps) / 1e9
    achieved = algo_bytes / (kern_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(f"dist_tiles:{args.workload}")
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: single cloud n={n} d={d} float32, metric {metric}, "
                               "full-layout distance matrix + fused enclosing radius",
                   "n": n, "dim": d, "metric": metric, "layout": "full", "seed": 2 if d == 3 else 3},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "dist_tiles (HIP events around its launch alone)",
                     "launch_ms": round(kern_ms, 5), "algorithmic_bytes": algo_bytes,
                     "step_ms_events": round(step_ms, 5),
                     "step_frac": round(algo_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, metric, algo_bytes)
    r = torch.empty(1)
    r.copy_(radius.cpu())
    out["config"]["enclosing_radius"] = float(r[0])
    if args.whole_run:
        # the same cloud end to end at maxdim 1 (the configuration's own
        # persistence call): GPU front end + host copies + host reduction
        from turbotops_amd import persistence
        t0 = time.perf_counter()
        dgms, info = persistence(x, maxdim=1, metric=metric)
        out["whole_run"] = {"seconds": round(time.perf_counter() - t0, 3), "engine": info["engine"],
                            "seconds_front": round(info["seconds_front"], 3),
                            "seconds_reduce": round(info["seconds_reduce"], 3), "edges": int(info["edges"]),
                            "h1_bars": int(len(dgms[1]))}
    return out


def run_edges(args):
    """Dimension 0 of the coboundary reducer on the dense complex of the
    single cloud (tt_dense_edges_dev over the matrix in HBM): count, compact,
    radix sort, block minima and the apparent scan; edges/sec."""
    import torch
    from turbotops_amd import dense_edges_dev, dist_matrix_dev
    x, metric = cloud(args.workload.replace("-edges", ""))
    n, d = x.shape
    xt = torch.from_numpy(x).cuda()
    radius = torch.empty(1, dtype=torch.float32, device="cuda")
    rowmax = torch.empty(n, dtype=torch.int32, device="cuda")
    mat = dist_matrix_dev(xt, metric, "full", rowmax=rowmax, radius=radius).view(n, n)
    torch.cuda.synchronize()
    thr = float(radius.item())
    for _ in range(args.warmup):
        m = dense_edges_dev(mat, thr, count_only=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m = dense_edges_dev(mat, thr, count_only=True)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / args.steps
    return {"metric": "dense dimension-0 edges + apparent scan, edges/sec", "value": round(m / per, 1),
            "unit": "edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(per * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{args.workload}: dimension-0 edges of the n={n} d={d} {metric} matrix at its "
                                   f"enclosing radius {thr:.7g}", "edges": m}}


VALU_PEAK_TFLOPS = 157.3   # MI355X dense FP32 vector peak


def run_graph(args):
    """The point-space front-end stages on their own clouds: exact kNN graph
    (k = 15) and radius graph (tau = 0.15) on n = 200000 points U[0,1]^8
    (seed 5), and H0 by Boruvka on n = 1e6 points U[0,1]^3 (seed 4).
    Host-to-host calls (points in, CSR / deaths out); the PCIe share is under
    2 ms of each. Roofline: FP32 VALU, 3d + 1 flops per pair evaluation."""
    from turbotops_amd import knn_graph_exact, mst_deaths, radius_graph
    w = args.workload
    if w == "c2c-mst":
        x = np.random.default_rng(4).uniform(size=(1000000, 3)).astype(np.float32)
    else:
        x = np.random.default_rng(5).uniform(size=(200000, 8)).astype(np.float32)
    n, d = x.shape
    fn = {"c3a-knn": lambda: knn_graph_exact(x, "l2", 15), "c3b-radius": lambda: radius_graph(x, "l2", 0.15),
          "c2c-mst": lambda: mst_deaths(x, "l2")}[w]
    for _ in range(args.warmup):
        res = fn()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = fn()
    per = (time.perf_counter() - t0) / args.steps
    if w == "c3a-knn":
        pairs = float(n) * (n - 1)                 # every query scans every candidate
        size = {"edges": int(res.row_ptr[-1]) // 2}
    elif w == "c3b-radius":
        pairs = float(n) * (n - 1) / 2             # lower-triangle tiles
        size = {"edges": int(res.row_ptr[-1]) // 2}
    else:
        # the Boruvka rounds, vertex searches and point pairs measured by one
        # call (tt_mst_deaths_dev): the searches skip boxes of points that
        # cannot hold a nearer vertex outside the component, so the pairs are
        # far fewer than searches x (n - 1)
        import ctypes as C
        import torch
        from turbotops_amd import capi
        from turbotops_amd.frontend import parse_metric
        kind, p = parse_metric("l2")
        xt = torch.from_numpy(x).cuda()
        dd = torch.empty(n - 1, dtype=torch.float32, device="cuda")
        rounds, scans, measured = C.c_uint32(), C.c_uint64(), C.c_uint64()
        capi.lib().call("tt_mst_deaths_dev", C.c_void_p(xt.data_ptr()), n, d, kind, C.c_float(p),
                        C.c_void_p(dd.data_ptr()), C.byref(rounds), C.byref(scans), C.byref(measured), None)
        torch.cuda.synchronize()
        pairs = float(measured.value)
        size = {"deaths": int(res.size), "rounds": int(rounds.value), "vertex_searches": int(scans.value),
                "pairs_measured": int(measured.value),
                "pairs_brute_force": int(scans.value) * (n - 1)}
    out = {"metric": f"{w} graph stage, seconds per call", "value": round(per, 5), "unit": "s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(per * 1e3, 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": dict({"workload": f"{w}: n={n} d={d} L2"}, **size)}
    if pairs is not None:
        tflops = pairs * (3 * d + 1) / per / 1e12
        out["roofline"] = {"bound": "valu", "achieved": round(tflops, 2), "peak": VALU_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(tflops / VALU_PEAK_TFLOPS, 4), "traffic": None,
                           "pairs": pairs,
                           "note": {"c3a-knn": "every pair measured (brute force over a spatial order)",
                                    "c3b-radius": "brute-force pair count: tiles whose boxes lie farther "
                                                  "apart than tau are skipped, so fewer pairs are measured",
                                    "c2c-mst": "pairs actually measured by the box-pruned searches; the "
                                               "searches are bound by box tests and scalar loads, not "
                                               "FP32 issue"}[w]}
    if args.whole_run and w in ("c3a-knn", "c3b-radius"):
        # the configuration's own persistence call end to end: the kNN graph
        # complex at maxdim 2, the radius-graph complex at maxdim 1
        from turbotops_amd import persistence
        kw = dict(maxdim=2, graph="knn", k=15) if w == "c3a-knn" else dict(maxdim=1, threshold=0.15)
        t0 = time.perf_counter()
        dgms, info = persistence(x, metric="l2", **kw)
        out["whole_run"] = {"seconds": round(time.perf_counter() - t0, 3), "maxdim": kw["maxdim"],
                            "engine": info["engine"], "seconds_graph": round(info["seconds_graph"], 3),
                            "seconds_reduce": round(info["seconds_reduce"], 3),
                            "bars": [int(len(a)) for a in dgms]}
    if not args.no_cpu:
        from oracle import Reference, reference_available
        if reference_available(simd=True):
            threads = min(CPU_CORES_ALLOTTED, os.cpu_count() or 1)
            ref = Reference(simd=True)
            if w == "c2c-mst":
                m = 100000                         # Prim is O(n^2): time a sample, scale by n^2
                secs, _ = ref.time_graph(x[:m], "l2", "mst", threads=threads)
                out["cpu_baseline"] = {"value": round(secs * (n / m) ** 2, 2), "unit": "s", "cores": threads,
                                       "kind": "reference",
                                       "sample": f"mst::h0 on the first {m} points: {secs:.2f} s, scaled by (n/m)^2"}
            else:
                which = "knn" if w == "c3a-knn" else "radius"
                secs, _ = ref.time_graph(x, "l2", which, tau=0.15, k=15, threads=threads)
                out["cpu_baseline"] = {"value": round(secs, 3), "unit": "s", "cores": threads, "kind": "reference",
                                       "sample": f"the full workload once ({which}, k-d tree, {threads} threads)"}
    return out


def c4_clouds():
    """The 512-cloud batched set: n_i ~ U{200..4000}, d_i in {2, 3, 8}, noisy
    unions of 1-4 circles/spheres (sigma 0.02), seed 6."""
    from turbotops_amd.synthetic import uneven_clouds
    return uneven_clouds(count=512, nmin=200, nmax=4000, dims=(2, 3, 8), seed=6)


def _front_timing(bp, clouds):
    bp.front_end(clouds)                 # autotune + graph capture
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        bp.front_end(clouds)
        runs.append((time.perf_counter() - t0, bp.device_ms))
    return min(runs, key=lambda r: r[1])


def run_batched(args, metric="l2"):
    """The batched form on the 512-cloud set.

    front end: all 512 clouds through the recorded launch sequence (distances,
    enclosing radii, the edge sort, copies to host memory), timed with the
    library's HIP events on its own stream, in the boundary engine's edge
    order (the default: every kept edge by (diameter, comb index), with its
    ready-witness flag) and in the
    cohomology engine's (apparent test, then the non-apparent edges sorted);
    algorithmic bytes of the distance stage are sum(4 n_i^2 + 4 n_i d_i).
    diagrams/sec: whole runs (front end + host reductions) of the batched
    form on all 512 clouds, and against the plain per-cloud path
    (persistence() on each cloud in turn) on the clouds with n_i <= `cap`.
    """
    from turbotops_amd import BatchPersistence, persistence
    clouds = c4_clouds()
    algo = float(sum(4.0 * c.shape[0] ** 2 + 4.0 * c.shape[0] * c.shape[1] for c in clouds))
    t0 = time.perf_counter()
    bp = BatchPersistence([c.shape for c in clouds], metric=metric)
    setup_s = time.perf_counter() - t0
    wall, dev_ms = _front_timing(bp, clouds)
    dist_ms = bp.time_distance_stage()
    kept, _ = bp.edges
    out = {"workload": f"512 clouds, n_i ~ U{{200..4000}}, d_i in {{2,3,8}}, circles/spheres sigma 0.02, seed 6, "
                       f"{metric}, maxdim 1",
           "front_end_device_ms": round(dev_ms, 3), "front_end_wall_ms": round(wall * 1e3, 3),
           "front_end_edge_order": "boundary engine: every kept edge by (diameter, comb index), "
                                   "with its ready-witness flag",
           "distance_stage_bytes": algo, "distance_stage_ms": round(dist_ms, 3),
           "distance_stage_gbs": round(algo / (dist_ms * 1e-3) / 1e9, 1),
           "distance_stage_frac": round(algo / (dist_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "edges_kept": int(kept), "rows_per_wave": bp.rows_per_wave, "setup_s": round(setup_s, 3)}
    if args.c4_full:
        bp(clouds)                       # first whole run: host buffers
        t0 = time.perf_counter()
        bp(clouds)
        tf = time.perf_counter() - t0
        out["diagrams_per_sec_all_512"] = {"batched": round(len(clouds) / tf, 3), "whole_run_s": round(tf, 3)}
    del bp
    bn = BatchPersistence([c.shape for c in clouds], metric=metric, witness_flags=False)
    wall_n, dev_n = _front_timing(bn, clouds)
    out["front_end_boundary_no_witness"] = {"device_ms": round(dev_n, 3), "wall_ms": round(wall_n * 1e3, 3)}
    del bn
    bc = BatchPersistence([c.shape for c in clouds], metric=metric, mode="coboundary")
    wall_c, dev_c = _front_timing(bc, clouds)
    out["front_end_coboundary_order"] = {"device_ms": round(dev_c, 3), "wall_ms": round(wall_c * 1e3, 3),
                                         "edges_not_apparent": int(bc.edges[1])}
    del bc
    cap, count = args.c4_cap, args.c4_sample
    if count <= 0:
        return out
    sample = [c for c in clouds if c.shape[0] <= cap][:count]
    bs = BatchPersistence([c.shape for c in sample], metric=metric)
    bs(sample)                           # autotune + graph capture
    t0 = time.perf_counter()
    res = bs(sample)
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain = []
    for i, c in enumerate(sample):
        plain.append(persistence(c, maxdim=1, metric=metric)[0])
        if (i + 1) % 64 == 0:
            print(f"bench: plain path {i + 1}/{len(sample)} clouds", file=sys.stderr, flush=True)
    t_plain = time.perf_counter() - t0
    same = all(all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(r[0], p))
               for r, p in zip(res, plain))
    out["diagrams_per_sec"] = {
        "batched": round(len(sample) / t_batch, 3), "plain": round(len(sample) / t_plain, 3),
        "ratio": round(t_plain / t_batch, 2), "identical_diagrams": bool(same),
        "sample": f"the {len(sample)} clouds with n_i <= {cap} of the 512 (sum n_i = "
                  f"{sum(c.shape[0] for c in sample)}); whole runs: GPU front end + host reductions "
                  f"(batched: all host threads; plain: persistence() per cloud)"}
    return out


def run_batched_knn(args, k=15, plain_count=128):
    """The 512-cloud set through the batch's kNN stage (graph="knn", k = 15,
    maxdim 1): every cloud's exact kNN lists and symmetric graph in one
    recorded device sequence, then the sparse reductions on host threads;
    against persistence(cloud, graph="knn", k=15) per cloud on the first
    `plain_count` clouds. Diagrams compared cloud by cloud on that sample."""
    from turbotops_amd import BatchPersistence, persistence
    clouds = c4_clouds()
    bp = BatchPersistence([c.shape for c in clouds], graph="knn", k=k)
    bp.front_end(clouds)                 # graph capture
    fronts = []
    for _ in range(3):
        t0 = time.perf_counter()
        bp.front_end(clouds)
        fronts.append((time.perf_counter() - t0, bp.device_ms))
    wall, dev = min(fronts, key=lambda r: r[1])
    res = bp(clouds)
    t0 = time.perf_counter()
    res = bp(clouds)
    tb = time.perf_counter() - t0
    sample = clouds[:plain_count]
    t0 = time.perf_counter()
    plain = [persistence(c, maxdim=1, graph="knn", k=k)[0] for c in sample]
    tp = time.perf_counter() - t0
    same = all(all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(r[0], p))
               for r, p in zip(res, plain))
    return {"workload": f"the 512 clouds, symmetric exact kNN graph k={k}, l2, maxdim 1",
            "front_end_device_ms": round(dev, 3), "front_end_wall_ms": round(wall * 1e3, 3),
            "csr_entries": int(bp.nnz),
            "diagrams_per_sec": {"batched_all_512": round(len(clouds) / tb, 3),
                                 "plain": round(len(sample) / tp, 3),
                                 "ratio": round((tp / len(sample)) / (tb / len(clouds)), 2),
                                 "identical_diagrams": bool(same),
                                 "sample": f"plain: persistence() per cloud on the first {len(sample)} clouds; "
                                           f"batched: whole runs of all 512 (all host threads)"}}


def run_batched_linf(args):
    """The 512-cloud set under L-inf (the configuration's second metric):
    front-end device time, the distance stage, and one whole run of every
    cloud (front end + host reductions, all host threads)."""
    from turbotops_amd import BatchPersistence
    clouds = c4_clouds()
    algo = float(sum(4.0 * c.shape[0] ** 2 + 4.0 * c.shape[0] * c.shape[1] for c in clouds))
    bp = BatchPersistence([c.shape for c in clouds], metric="linf")
    wall, dev_ms = _front_timing(bp, clouds)
    dist_ms = bp.time_distance_stage()
    t0 = time.perf_counter()
    bp(clouds)
    tf = time.perf_counter() - t0
    out = {"front_end_device_ms": round(dev_ms, 3), "front_end_wall_ms": round(wall * 1e3, 3),
           "distance_stage_ms": round(dist_ms, 3), "distance_stage_gbs": round(algo / (dist_ms * 1e-3) / 1e9, 1),
           "edges_kept": int(bp.edges[0]), "rows_per_wave": bp.rows_per_wave,
           "diagrams_per_sec_all_512": {"batched": round(len(clouds) / tf, 3), "whole_run_s": round(tf, 3)}}
    del bp
    return out


def run_differentiable(args):
    """Forward + backward of -sum(death - birth) over the top-8 H1 bars on a
    noisy circle (radius 1, sigma 0.05, R^3, seed 7), L2; and the backward
    kernel alone on a large edge set."""
    import torch
    from turbotops_amd import diagram, diagram_backward, top_k_persistence_loss
    from turbotops_amd.synthetic import noisy_circle
    n = args.c5_n
    # one tiny forward + backward off the clock: loads the kernels that torch
    # and the library use the first time
    xw = torch.tensor(noisy_circle(n=64, seed=1), device="cuda", requires_grad=True)
    top_k_persistence_loss(diagram(xw, maxdim=1)[1], 8).backward()
    torch.cuda.synchronize()
    x = torch.tensor(noisy_circle(n=n, seed=7), device="cuda", requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dg = diagram(x, maxdim=1)
    loss = top_k_persistence_loss(dg[1], 8)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # the kernel alone: 2e6 critical edges over 2e5 points in R^3
    g = torch.Generator(device="cuda").manual_seed(0)
    pts = torch.randn(200000, 3, device="cuda", generator=g)
    e = torch.randint(0, 200000, (2000000, 2), device="cuda", dtype=torch.int32, generator=g)
    gr = torch.randn(2000000, device="cuda", generator=g)
    for _ in range(3):
        diagram_backward(pts, e, gr)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(10):
        diagram_backward(pts, e, gr)
    torch.cuda.synchronize()
    kern = (time.perf_counter() - t3) / 10
    out = {"workload": f"noisy circle n={n} in R^3 (radius 1, sigma 0.05, seed 7), L2, loss -sum(death - birth) "
                       f"over the top-8 H1 bars",
           "forward_s": round(t1 - t0, 3), "backward_ms": round((t2 - t1) * 1e3, 3),
           "loss": round(float(loss.detach()), 6), "h1_bars": int(dg[1].shape[0]),
           "backward_call_ms_2e6_edges": round(kern * 1e3, 3)}
    if args.c5_batch > 0:
        from turbotops_amd import DiagramBatch
        from turbotops_amd.synthetic import noisy_circles
        clouds = noisy_circles(count=64, nmin=512, nmax=4096, seed=9)
        clouds = [c for c in clouds if c.shape[0] <= args.c5_batch_nmax][:args.c5_batch]
        db = DiagramBatch([c.shape for c in clouds])

        def batch_step():
            xs = [torch.tensor(c, device="cuda", requires_grad=True) for c in clouds]
            loss = sum(top_k_persistence_loss(o[1], 8) for o in db(xs))
            loss.backward()
            torch.cuda.synchronize()
            return xs

        batch_step()                     # autotune + graph capture
        t0 = time.perf_counter()
        batch_step()
        tb = time.perf_counter() - t0
        out["batch"] = {"clouds": len(clouds), "step_s": round(tb, 3),
                        "diagrams_per_sec": round(len(clouds) / tb, 3),
                        "sample": f"{len(clouds)} of the 64 noisy circles n_i ~ U{{512..4096}} (seed 9), those with "
                                  f"n_i <= {args.c5_batch_nmax}; forward + backward of the summed top-8 H1 losses "
                                  f"per step"}
    return out


def copy_probe(args):
    """A plain device copy of known size, for scaling the rocprofv3 FETCH_SIZE
    / WRITE_SIZE counters (1 GiB read + 1 GiB written per launch)."""
    import torch
    a = torch.ones(2**28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(args.steps, 1)):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": a.numel() * 4, "launches": max(args.steps, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2a",
                    help="c2a | c2b-l1 | c2b-linf | c2b-lp (distance matrix); c2a-edges (dimension-0 edges); "
                         "c3a-knn | c3b-radius | c2c-mst (graph stages)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-whole-run", dest="whole_run", action="store_false",
                    help="skip the end-to-end persistence() run of the single cloud")
    ap.add_argument("--copy-probe", action="store_true", help="run only the copy kernel (PMC scaling)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 512-cloud batched leg")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable leg")
    ap.add_argument("--c4-cap", type=int, default=4000,
                    help="largest n_i in the batched vs plain sample (4000: all 512 clouds; the plain leg takes "
                         "about two minutes)")
    ap.add_argument("--c4-sample", type=int, default=512, help="most clouds in the batched vs plain sample")
    ap.add_argument("--no-c4-full", dest="c4_full", action="store_false",
                    help="skip the whole batched run of all 512 clouds")
    ap.add_argument("--no-c4-linf", dest="c4_linf", action="store_false",
                    help="skip the 512-cloud set under L-inf")
    ap.add_argument("--no-c4-knn", dest="c4_knn", action="store_false",
                    help="skip the 512-cloud set through the batch's kNN-graph stage")
    ap.add_argument("--no-graph-stages", dest="graph_stages", action="store_false",
                    help="skip the compact C3a kNN / C3b radius / C2c Boruvka timings of the default line")
    ap.add_argument("--c5-n", type=int, default=2048, help="points of the differentiable circle")
    ap.add_argument("--c5-batch", type=int, default=64, help="clouds in the differentiable batch")
    ap.add_argument("--c5-batch-nmax", type=int, default=4096, help="largest n_i taken into that batch")
    args = ap.parse_args()
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU: this build runs the front end on a single MI355X "
                         "(--gpus 1); multi-GPU sharding is design-only")
    if args.copy_probe:
        return copy_probe(args)
    if args.workload.endswith("-edges"):
        print(json.dumps(run_edges(args)), flush=True)
        return
    if args.workload in ("c3a-knn", "c3b-radius", "c2c-mst"):
        print(json.dumps(run_graph(args)), flush=True)
        return
    out = run_dense(args)
    if not args.no_batched:
        print("bench: batched leg", file=sys.stderr, flush=True)
        out["batched"] = run_batched(args)
        if args.c4_linf:
            print("bench: batched leg, L-inf", file=sys.stderr, flush=True)
            out["batched_linf"] = run_batched_linf(args)
        if args.c4_knn:
            print("bench: batched leg, kNN graphs", file=sys.stderr, flush=True)
            out["batched_knn"] = run_batched_knn(args)
    if not args.no_diff:
        print("bench: differentiable leg", file=sys.stderr, flush=True)
        out["differentiable"] = run_differentiable(args)
    if args.graph_stages:
        # the point-space stages on their own configurations, host to host
        # (seconds per call over 3 calls after a warm-up; no CPU baseline
        # or whole run here: bench.py --workload NAME gives those)
        print("bench: graph stages", file=sys.stderr, flush=True)
        stages = {}
        for w in ("c3a-knn", "c3b-radius", "c2c-mst"):
            g = run_graph(argparse.Namespace(workload=w, warmup=1, steps=3, whole_run=False, no_cpu=True))
            entry = {"seconds_per_call": g["value"], "config": g["config"]}
            if "roofline" in g:
                entry["pairs"] = g["roofline"]["pairs"]
                entry["pairs_note"] = g["roofline"]["note"]
            stages[w] = entry
        out["graph_stages"] = stages
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
