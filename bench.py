#!/usr/bin/env python3
"""PageRank benchmark on one MI355X: GTEPS and fraction of the HBM roofline.

Workload (one "step" = one full PageRank solve to tolerance on the resident
graph): directed R-MAT scale 22, edge factor 16 (a=0.57, b=0.19, c=0.19),
self-loops and duplicates removed, seed 27491095, uint32 in-CSR; damping
0.85, uniform teleport, dangling mass redistributed uniformly, tolerance 1e-6
(L1), at most 100 sweeps; FP32 ranks with FP32 accumulation.

value = GTEPS = edges x sweeps / wall time of the timed solves (inputs
resident in HBM; each solve includes its init, all sweeps and the output
permutation).  roofline.achieved = algorithmic bytes per sweep
(8*M + 16*N + 4: a uint32 index and an FP32 gathered contribution per edge;
offset, degree, old and new rank per vertex) x sweeps / summed device time of
the sweep kernels (HIP events around every captured chunk).

cpu_baseline: the OpenMP restatement of the reference's default page_rank
(v11, in-place Gauss-Seidel) on the same graph, FP32, solved to the same
tolerance and repeated for about --cpu-seconds on the host cores.

fwd_adjoint_vs_fwd: the metric's third part -- weighted PageRank on R-MAT s20
ef16, device time of forward + adjoint (dL/dw, dL/dt) over forward alone.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PageRank GTEPS & % HBM roofline, R-MAT s22; fwd+adjoint vs fwd time"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=27491095)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iters", type=int, default=100)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--engine", choices=["auto", "binned", "tiled"], default="auto",
                    help="sweep engine (auto: the library's choice)")
    ap.add_argument("--tile-bits", type=int, default=0, help="tiled engine source tile width 2^bits (0 = default)")
    ap.add_argument("--combine", choices=["auto", "gather", "accumulate"], default="auto",
                    help="tiled engine second pass (auto: the library's choice)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="multi and shard workloads: pick the decomposition with nwgr_amd.autotune before "
                         "timing (0: defaults)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the baseline sample (0 disables it)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = OMP_NUM_THREADS or all cores")
    ap.add_argument("--workload", choices=["pagerank", "er", "ppr", "multi", "wpr", "delta", "shard"],
                    default="pagerank",
                    help="pagerank: the headline solve on R-MAT s22; er: the same solve on Erdos-Renyi "
                         "G(n = 2^25, m = 16n); ppr: 256 personalized columns on R-MAT s20; "
                         "multi: 2048 R-MAT graphs (n log-uniform in [1k, 64k], ef 8) in one launch; "
                         "wpr: weighted forward + adjoint on R-MAT s20; delta: residual PageRank "
                         "(page_rank_v13, threshold 1e-4) on R-MAT s22; shard: each GPU's row share of R-MAT s22 "
                         "swept alone on its own layout (--shards shares)")
    ap.add_argument("--shards", type=int, default=8, help="shard workload: number of row slices")
    ap.add_argument("--balance-rounds", type=int, default=2,
                    help="shard workload: re-cuts of the shares by their measured sweep times (0: model cut only)")
    ap.add_argument("--fwd-adjoint", type=int, default=1,
                    help="1: also time the weighted forward and forward + adjoint on R-MAT s20 (the metric's "
                         "fwd+adjoint vs fwd ratio); 0: skip")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (3 times) and exit: the known-size "
                         "reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def calibrate_copy(mib):
    import torch
    n = mib * 2**20 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}), flush=True)


def copy_rate_gbs(mib=1024, reps=3):
    """Bytes read + written per second by a plain device-to-device copy of
    `mib` MiB (torch's copy kernel, the best of `reps`): the box's practical
    streaming rate, against which measured traffic gives a ceiling."""
    import torch
    n = mib * 2**20 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    return 2 * n * 4 / (best / 1e3) / 1e9


def load_traffic(scale, ef, workload="pagerank"):
    """HBM bytes per sweep launch from committed rocprofv3 --pmc summaries."""
    name = "k_sweep_traffic_er_2pow25.json" if workload == "er" else f"k_sweep_traffic_rmat_s{scale}_ef{ef}.json"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_sweep")


def load_profile_bytes(name, key):
    """A measured byte count from a committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def host_info(threads):
    """The host the CPU baseline runs on: CPU model, sockets, physical cores,
    logical CPUs, the CPUs this process may use (affinity, cgroup quota) and
    the threads the baseline used."""
    info = {"threads_used": threads}
    try:
        model, phys, logical = None, set(), 0
        sock = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    sock = v
                elif k == "core id":
                    phys.add((sock, v))
        info.update({"cpu_model": model, "sockets": len({s for s, _ in phys}) or None,
                     "physical_cores": len(phys) or None, "logical_cpus": logical})
    except OSError:
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
        info["cgroup_cpu_quota"] = None if q == "max" else round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    info["omp_num_threads_env"] = os.environ.get("OMP_NUM_THREADS")
    pc = info.get("physical_cores")
    info["all_physical_cores_used"] = bool(pc) and threads >= pc
    return info


def dropin_calls(g, deg, tol, max_iters, reps, pg, out_d):
    """The drop-in as pr.exe times it (bench/pr.cpp:94-95 brackets the whole
    page_rank call): nwgr_amd.page_rank on host arrays, LEAK dangling as the
    reference, ranks written back into the caller's float array.  The first
    call builds the resident device graph (upload, relabelling, tiled layout);
    later calls find it in the cache and pay for the solve and the copy of the
    ranks.  Beside them: the same solve on the resident handle with the ranks
    left in device memory, and with them copied to the host."""
    import torch
    import nwgr_amd
    nwgr_amd.cache_clear()
    ranks = np.zeros(g.n, np.float32)
    t0 = time.perf_counter()
    nwgr_amd.page_rank(g, deg, ranks, 0.85, tol, max_iters)
    first = (time.perf_counter() - t0) * 1e3
    rep = []
    for _ in range(reps):
        t0 = time.perf_counter()
        nwgr_amd.page_rank(g, deg, ranks, 0.85, tol, max_iters)
        rep.append((time.perf_counter() - t0) * 1e3)
    kw = dict(damping=0.85, tol=tol, max_iters=max_iters, dangling="leak")
    dev, host = [], []
    out = np.empty(g.n, np.float32)
    res = pg.run_device(out_d.data_ptr(), **kw)
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = pg.run_device(out_d.data_ptr(), **kw)
        dev.append((time.perf_counter() - t0) * 1e3)
        t0 = time.perf_counter()
        pg.run(out=out, **kw)
        host.append((time.perf_counter() - t0) * 1e3)
    st = nwgr_amd.cache_stats()
    same = bool(np.array_equal(out, ranks))
    rep_ms, dev_ms = float(np.median(rep)), float(np.median(dev))
    return {"semantics": "nwgr_amd.page_rank(graph, degrees, ranks, 0.85, tol, max_iters) wall time per call, "
                         "dangling leaked as the reference, host vectors in and out (bench/pr.cpp:94-95)",
            "first_call_ms": round(first, 3), "repeat_call_ms_median": round(rep_ms, 3),
            "repeat_call_ms_min_max": [round(min(rep), 3), round(max(rep), 3)],
            "resident_solve_device_output_ms_median": round(dev_ms, 3),
            "resident_solve_host_output_ms_median": round(float(np.median(host)), 3),
            "repeat_over_resident": round(rep_ms / dev_ms, 3), "sweeps_per_solve": int(res.iters),
            "calls": reps + 1, "cache_hits": st["hits"], "ranks_equal_resident": same}


def cpu_baseline(g, deg, target_s, threads, tol, max_iters):
    """The reference's default page_rank (v11, in-place Gauss-Seidel) as the
    OpenMP restatement, FP32 as bench/pr.cpp:89, solved to the same tolerance
    as the GPU solve; whole solves are repeated (as pr.exe's trials) until
    about 3/4 of target_s seconds have been spent, after one untimed warm-up
    solve, and reported as min/avg/max per solve (Times::print,
    bench/common.hpp:205-231).  The Jacobi form (page_rank_v10, OpenMP, FP32)
    is timed beside it with the rest of the budget."""
    import oracle  # the CPU restatement: baseline leg only
    t0 = time.perf_counter()
    _, it = oracle.gs_par(g.offsets, g.indices, deg, 0.85, tol, max_iters, nthreads=threads, dtype=np.float32)
    per_solve = time.perf_counter() - t0
    trials = int(max(1, min(1000, round(0.75 * target_s / max(per_solve, 1e-6)))))
    sweeps = 0
    times = []
    for _ in range(trials):
        t1 = time.perf_counter()
        _, it = oracle.gs_par(g.offsets, g.indices, deg, 0.85, tol, max_iters, nthreads=threads, dtype=np.float32)
        times.append(time.perf_counter() - t1)
        sweeps += it
    dt = sum(times)
    # Jacobi (v10) restatement: same tolerance, a few solves
    t1 = time.perf_counter()
    _, jit = oracle.jacobi_par(g.offsets, g.indices, deg, 0.85, tol, max_iters, nthreads=threads)
    j_one = time.perf_counter() - t1
    j_trials = int(max(1, min(100, round(0.25 * target_s / max(j_one, 1e-6)))))
    t1 = time.perf_counter()
    for _ in range(j_trials):
        _, jit = oracle.jacobi_par(g.offsets, g.indices, deg, 0.85, tol, max_iters, nthreads=threads)
    jdt = time.perf_counter() - t1
    return {"value": g.m * sweeps / dt / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"R-MAT s{int(np.log2(g.n))} ef16 (the same graph): {trials} solves to L1 tol {tol:g} "
                      f"({it} in-place Gauss-Seidel sweeps each, {sweeps} in all) of the OpenMP restatement of "
                      f"page_rank v11 (FP32 ranks, dangling mass leaked as the reference, schedule(dynamic,64)), "
                      f"{dt:.2f} s",
            "ms_per_solve": dt * 1e3 / trials, "sweeps_per_solve": it, "seconds_per_sweep": dt / sweeps,
            "ms_per_solve_min_avg_max": [round(min(times) * 1e3, 2), round(dt * 1e3 / trials, 2),
                                         round(max(times) * 1e3, 2)],
            "jacobi_v10": {"ms_per_solve": round(jdt * 1e3 / j_trials, 2), "sweeps_per_solve": jit,
                           "gteps": round(g.m * jit * j_trials / jdt / 1e9, 3), "solves": j_trials}}


def fwd_adjoint_cost(seed, steps=5, warmup=2):
    """Weighted PageRank on R-MAT s20 ef16 (weights U(0,1], teleport
    Dirichlet(1), loss L = sum_v wl_v r_v with wl ~ N(0,1)): device time of
    the forward solve and of forward + adjoint (dL/dw, dL/dt), HIP events on
    the stream both run on; median over `steps` trials."""
    import torch
    import nwgr_amd
    from nwgr_amd import PageRankGraph
    from nwgr_amd.autograd import weighted_pagerank
    g = nwgr_amd.rmat(20, 16, seed=seed)
    rng = np.random.default_rng(seed)
    w = torch.from_numpy((1.0 - rng.random(g.m)).astype(np.float32)).to("cuda:0")
    t = torch.from_numpy(rng.dirichlet(np.ones(g.n)).astype(np.float32)).to("cuda:0")
    wl = torch.from_numpy(rng.standard_normal(g.n).astype(np.float32)).to("cuda:0")
    pg = PageRankGraph(g)
    kw = dict(tol=1e-7, max_iters=300, adj_tol=1e-7)
    fwd, both = [], []
    for i in range(warmup + steps):
        wd = w.clone().requires_grad_(True)
        td = t.clone().requires_grad_(True)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        r = weighted_pagerank(pg, wd, td, **kw)
        e1.record()
        (r * wl).sum().backward()
        e2.record()
        torch.cuda.synchronize()
        if i >= warmup:
            fwd.append(e0.elapsed_time(e1))
            both.append(e0.elapsed_time(e2))
    f, b = float(np.median(fwd)), float(np.median(both))
    return {"workload": f"weighted pagerank rmat s20 ef16 (n {g.n}, m {g.m}), w ~ U(0,1], teleport Dirichlet(1), "
                        f"L = sum wl_v r_v, forward tol 1e-7, adjoint tol 1e-7 relative, dangling leak",
            "fwd_ms": round(f, 3), "fwd_adjoint_ms": round(b, 3), "ratio": round(b / f, 3),
            "fwd_sweeps": int(pg.last_forward.iters) if hasattr(pg, "last_forward") else None,
            "adjoint_sweeps": int(pg.last_adjoint.iters) if getattr(pg, "last_adjoint", None) else None}


def bench_wpr(args):
    """Forward + adjoint vs forward time of the differentiable weighted form."""
    c = fwd_adjoint_cost(args.seed, steps=max(args.steps, 1), warmup=args.warmup)
    line = {
        "metric": "weighted PageRank forward+adjoint time / forward time", "value": c["ratio"], "unit": "x",
        "n_gpus": 1, "steps": max(args.steps, 1), "warmup": args.warmup, "ms_per_step": c["fwd_adjoint_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": c, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_ppr(args):
    """256 personalized columns (one-hot seeds) on R-MAT s20 ef16 in one
    batched solve; algorithmic bytes per sweep 4M + 4kM + 8kN + 8N."""
    import torch
    import nwgr_amd
    from nwgr_amd import PageRankGraph
    scale = 20 if args.scale == 22 else args.scale
    k = 256
    g = nwgr_amd.rmat(scale, args.edge_factor, seed=args.seed)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    seeds = np.random.default_rng(args.seed).integers(0, g.n, k).astype(np.int32)
    sd = torch.from_numpy(seeds).to("cuda:0")
    R = torch.empty((g.n, k), dtype=torch.float32, device="cuda:0")
    kw = dict(damping=0.85, tol=args.tol, max_iters=args.max_iters, dangling="uniform")
    for _ in range(args.warmup):
        pg.ppr_batch_device(sd.data_ptr(), k, R.data_ptr(), **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sweeps, dev_ms = 0, 0.0
    for _ in range(args.steps):
        res = pg.ppr_batch_device(sd.data_ptr(), k, R.data_ptr(), **kw)
        sweeps += res.iters
        dev_ms += res.ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n, m = g.n, g.m
    b_iter = 4 * m + 4 * k * m + 8 * k * n + 8 * n
    achieved = b_iter * sweeps / (dev_ms / 1e3) / 1e9
    traffic = (load_profile_bytes(f"k_ppr_traffic_rmat_s{scale}_k{k}.json", "hbm_bytes_per_sweep")
               if args.edge_factor == 16 else None)
    line = {
        "metric": "batched personalized PageRank, edge-columns per second", "value": round(m * k * sweeps / wall / 1e9, 3),
        "unit": "G edge-columns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"ppr rmat s{scale} ef{args.edge_factor}, {k} one-hot seeds, dangling uniform, "
                               f"tol {args.tol:g} per column", "n": n, "m": m, "k": k,
                   "sweeps_per_solve": res.iters, "device_ms_per_solve": round(dev_ms / args.steps, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_ppr_sweep+k_ppr_hub+k_ppr_reduce", "algorithmic_bytes_per_sweep": b_iter,
                     "traffic_unit": "bytes per sweep (FETCH_SIZE + WRITE_SIZE, rocprofv3 --pmc)",
                     "traffic_rate_gbs": None if traffic is None else round(traffic * sweeps / (dev_ms / 1e3) / 1e9, 1)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_delta(args):
    """Residual ("delta") PageRank as pr.exe version 13 (page_rank_v13) on
    R-MAT s22 ef16: threshold 1e-4 (pr.exe's default tolerance), up to 100
    iterations.  One iteration is one sweep (8M + 16N algorithmic bytes:
    index and gathered delta per edge; degree, rank in and out, delta out
    per vertex).  CPU baseline: the oracle's serial float restatement of
    v13 for a few iterations (kind "port", one thread)."""
    import torch
    import nwgr_amd
    import oracle
    from nwgr_amd import PageRankGraph
    thr = 1e-4
    g = nwgr_amd.rmat(args.scale, args.edge_factor, seed=args.seed)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    kw = dict(damping=0.85, threshold=thr, max_iters=args.max_iters, chunk=args.chunk)
    for _ in range(args.warmup):
        pg.run_delta_device(out.data_ptr(), **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, sweep_ms, launches = 0, 0.0, 0
    for _ in range(args.steps):
        res = pg.run_delta_device(out.data_ptr(), **kw)
        prof = pg.last_profile()
        iters += res.iters
        sweep_ms += prof["sweep_ms"]
        launches += prof["launches"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pushes = pg.last_errors()
    n, m = g.n, g.m
    b_iter = 8 * m + 16 * n
    sweeps = iters - args.steps  # iteration 1 of each solve is the init kernel's push
    achieved = b_iter * sweeps / (sweep_ms / 1e3) / 1e9
    cpu = None
    if args.cpu_seconds > 0:
        k_cpu = 20
        t1 = time.perf_counter()
        oracle.delta(g.offsets, g.indices, deg, 0.85, thr, k_cpu, np.float32)
        t_cpu = time.perf_counter() - t1
        cpu = {"value": round(m * k_cpu / t_cpu / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"oracle float restatement of page_rank_v13, {k_cpu} iterations on the same graph "
                         f"(serial, {t_cpu:.1f} s)"}
    line = {
        "metric": "residual PageRank (page_rank_v13) GTEPS", "value": round(m * iters / wall / 1e9, 3),
        "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"page_rank_v13 on rmat s{args.scale} ef{args.edge_factor}, threshold {thr:g}, "
                               f"max {args.max_iters} iterations, FP32", "n": n, "m": m,
                   "iterations_per_solve": res.iters, "pushes_first_last": [int(pushes[0]), int(pushes[-1])],
                   "engine": pg.last_engine()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": {"tiled": "k_tile+k_combine<delta>", "tiled-accumulate": "k_tile+k_accum<delta>"}.get(
                         pg.last_engine(), "k_sweep<delta>"),
                     "algorithmic_bytes_per_sweep": b_iter, "avg_launch_us": round(sweep_ms / launches * 1e3, 2),
                     "avg_sweep_us": round(sweep_ms / max(sweeps, 1) * 1e3, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_shard(args):
    """One GPU's share of a row-partitioned solve of R-MAT s22 ef16: the
    destination rows (in the handle's relabelled order) cut into `--shards`
    consecutive ranges of about equal sweep cost (the cost model's cut,
    then re-cut by measured share times: nwgr_pr_share_balance), each swept alone on a
    tiled layout that holds only its entries and stores only its rows
    (nwgr_pr_share_run_device_f32), the other rows' contributions resident
    and fixed, as one GPU holds them between the all-gathers of its slice.
    Per share: device time per sweep over fixed-length solves (tol 0, 20
    sweeps, LEAK); algorithmic bytes 8 M_k + 16 N_k.  Beside it the whole
    graph's sweep on the same handle."""
    import torch
    import nwgr_amd
    from nwgr_amd import PageRankGraph
    G = args.shards
    g = nwgr_amd.rmat(args.scale, args.edge_factor, seed=args.seed)
    deg = nwgr_amd.build_degrees(g)
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    kw = dict(damping=0.85, tol=0.0, max_iters=20)
    pg = PageRankGraph(g, deg)

    def per_sweep_us(run):
        for _ in range(max(args.warmup, 1)):
            run()
        ms, sw = 0.0, 0
        for _ in range(args.steps):
            res = run()
            ms += pg.last_profile()["sweep_ms"]
            sw += res.iters
        return ms / sw * 1e3

    full_us = per_sweep_us(lambda: pg.run_device(out.data_ptr(), dangling="leak", **kw))
    log(f"whole graph: {full_us:.1f} us per sweep ({pg.last_engine()})")
    indeg_caller = np.diff(g.offsets.astype(np.int64))

    def measure_shares(tag):
        order, cuts = pg.gs_phases(G)
        indeg = indeg_caller[order]
        shares = []
        for k in range(G):
            lo, hi = int(cuts[k]), int(cuts[k + 1])
            us = per_sweep_us(lambda: pg.run_share_device(out.data_ptr(), G, k, **kw))
            m_k, n_k = int(indeg[lo:hi].sum()), hi - lo
            b_k = 8 * m_k + 16 * n_k
            shares.append({"rows": [lo, hi], "m": m_k, "us_per_sweep": round(us, 2),
                           "gbs": round(b_k / (us / 1e6) / 1e9, 1)})
            log(f"{tag} share {k}: rows [{lo}, {hi}) m={m_k} {us:.1f} us per sweep")
        return shares

    model = measure_shares("model cut")
    model_worst = max(s["us_per_sweep"] for s in model)
    shares = model
    if args.balance_rounds > 0:
        t0 = time.perf_counter()
        est = pg.share_balance(G, rounds=args.balance_rounds)
        log(f"measured balance: {args.balance_rounds} rounds in {time.perf_counter() - t0:.2f} s, "
            f"slowest share {est:.1f} us per sweep while balancing")
        shares = measure_shares("balanced")
    worst = max(shares, key=lambda s: s["us_per_sweep"])
    b_w = 8 * worst["m"] + 16 * (worst["rows"][1] - worst["rows"][0])
    line = {
        "metric": f"per-GPU sweep time of one of {G} row shares, R-MAT s{args.scale} (slowest share)",
        "value": worst["us_per_sweep"], "unit": "us", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"rmat s{args.scale} ef{args.edge_factor}, relabelled rows cut into {G} ranges of "
                               f"equal sweep cost (modelled, then re-cut by measured share times over "
                               f"{args.balance_rounds} rounds), each swept alone on its own layout (only its "
                               f"entries, only its rows stored), the other contributions resident; 20 sweeps per "
                               f"solve, LEAK",
                   "n": g.n, "m": g.m, "whole_graph_us_per_sweep": round(full_us, 2),
                   "ideal_us_per_sweep": round(full_us / G, 2), "shares": shares,
                   "sum_of_shares_us": round(sum(s["us_per_sweep"] for s in shares), 2),
                   "model_cut_slowest_us": model_worst, "balance_rounds": args.balance_rounds},
        "roofline": {"bound": "hbm", "achieved": worst["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(worst["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_tile+k_combine (one share's layout)", "algorithmic_bytes_per_sweep": b_w},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_multi(args):
    """2,048 R-MAT graphs, n log-uniform in [1k, 64k], edge factor 8, every
    graph solved to its own tolerance in one persistent launch."""
    import torch
    import nwgr_amd
    from nwgr_amd import MultiPageRank
    rng = np.random.default_rng(args.seed)
    ns = np.exp(rng.uniform(np.log(1024), np.log(65536), 2048)).astype(np.uint32)
    g, vbase = nwgr_amd.rmat_many(ns, 8, seed=args.seed)
    deg = nwgr_amd.build_degrees(g)
    mp = MultiPageRank(g, vbase, deg)
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    kw = dict(damping=0.85, tol=args.tol, max_iters=args.max_iters, dangling="uniform")
    if args.autotune:
        from nwgr_amd.autotune import autotune_multi
        tr = autotune_multi(mp, sweeps=8, reps=1, rounds=2)
        log(f"autotuned {tr['tuning']} in {tr['seconds']:.1f}s ({tr['trials']} trials)")
    for _ in range(args.warmup):
        mp.run_device(out.data_ptr(), **kw)
    offs = g.offsets.astype(np.int64)
    m_g = offs[vbase[1:].astype(np.int64)] - offs[vbase[:-1].astype(np.int64)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(args.steps):
        res = mp.run_device(out.data_ptr(), **kw)
        dev_ms += res.ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # every run of the same input takes the same sweeps per graph
    iters, _ = mp.graph_results()
    edge_sweeps = args.steps * int(np.sum(m_g * iters.astype(np.int64)))
    n_g = np.diff(vbase.astype(np.int64))
    b_sweeps = int(np.sum((8 * m_g + 16 * n_g) * iters.astype(np.int64)))  # last run's sweeps per graph
    achieved = b_sweeps / (dev_ms / args.steps / 1e3) / 1e9
    traffic = load_profile_bytes("k_multi_traffic_2048_rmat.json", "hbm_bytes_per_launch")
    line = {
        "metric": "many-graph PageRank GTEPS (sum over graphs of edges x sweeps / s)",
        "value": round(edge_sweeps / wall / 1e9, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"2048 rmat graphs, n log-uniform in [1k, 64k] (sum {g.n}), ef 8 (m {g.m}), "
                               f"dangling uniform, tol {args.tol:g} per graph, one launch",
                   "sweeps_min": int(iters.min()), "sweeps_max": int(iters.max()),
                   "device_ms_per_solve": round(dev_ms / args.steps, 3), "tuning": mp.tuning},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_multi",
                     "algorithmic_bytes": "sum over graphs of (8 m_g + 16 n_g) x sweeps_g",
                     "algorithmic_bytes_per_launch": b_sweeps,
                     "traffic_unit": "bytes per launch (FETCH_SIZE + WRITE_SIZE, rocprofv3 --pmc); below the "
                                     "algorithmic bytes: 2-byte local ids, contributions gathered from LDS",
                     "traffic_rate_gbs": None if traffic is None else round(traffic / (dev_ms / args.steps / 1e3) / 1e9, 1)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.workload == "ppr":
        bench_ppr(args)
        return
    if args.workload == "multi":
        bench_multi(args)
        return
    if args.workload == "wpr":
        bench_wpr(args)
        return
    if args.workload == "delta":
        bench_delta(args)
        return
    if args.workload == "shard":
        bench_shard(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the PageRank path here runs on a single MI355X "
              "(multi-GPU sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import nwgr_amd
    from nwgr_amd import PageRankGraph

    t0 = time.perf_counter()
    if args.workload == "er":
        n_er = 1 << 25
        g = nwgr_amd.erdos_renyi(n_er, 16 * n_er, seed=args.seed)
        gname = f"erdos-renyi G(n=2^25, m=16n drawn with replacement, dedup, no self-loops, seed {args.seed})"
    else:
        g = nwgr_amd.rmat(args.scale, args.edge_factor, seed=args.seed)
        gname = (f"rmat s{args.scale} ef{args.edge_factor} (a=0.57,b=0.19,c=0.19, seed {args.seed}, dedup, "
                 f"no self-loops)")
    deg = nwgr_amd.build_degrees(g)
    log(f"generated {gname}: n={g.n} m={g.m} in {time.perf_counter() - t0:.1f}s")
    pg = PageRankGraph(g, deg)
    if args.engine != "auto" or args.tile_bits or args.combine != "auto":
        tune = {}
        if args.engine != "auto":
            tune["engine"] = {"binned": 0, "tiled": 1}[args.engine]
        if args.tile_bits:
            tune["tile_bits"] = args.tile_bits
        if args.combine != "auto":
            tune["tile_combine"] = {"gather": 0, "accumulate": 1}[args.combine]
        pg.tuning = tune
    log(f"device graph: {pg.device_bytes / 2**20:.0f} MiB, setup {pg.setup_ms:.1f} ms")
    import torch
    kw = dict(damping=0.85, tol=args.tol, max_iters=args.max_iters, dangling="uniform", chunk=args.chunk)
    out_d = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    for i in range(args.warmup):
        res = pg.run_device(out_d.data_ptr(), **kw)
        log(f"warmup {i}: sweeps={res.iters} err={res.error:.3e} {res.ms:.3f} ms, engine {pg.last_engine()}")
    sweeps = 0
    sweep_ms = 0.0
    launches = 0
    dev_ms = 0.0
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solve_ms = []
    for i in range(args.steps):
        res = pg.run_device(out_d.data_ptr(), **kw)
        prof = pg.last_profile()
        sweeps += res.iters
        iters.append(res.iters)
        sweep_ms += prof["sweep_ms"]
        launches += prof["launches"]
        dev_ms += res.ms
        solve_ms.append(res.ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    engine = pg.last_engine()
    mass = float(out_d.double().sum())
    # the drop-in's host-vector form (ranks copied back over PCIe): median of
    # five solves after one that allocates the handle's pinned staging
    out = np.empty(g.n, np.float32)
    pg.run(out=out, **kw)
    host_runs = []
    for _ in range(5):
        t1 = time.perf_counter()
        pg.run(out=out, **kw)
        host_runs.append((time.perf_counter() - t1) * 1e3)
    host_ms = float(np.median(host_runs))
    log(f"timed {args.steps} solves: {wall * 1e3:.2f} ms wall, {dev_ms:.2f} ms device, sweeps/solve={iters[0]}, "
        f"sum(r)={mass:.7f}")
    if abs(mass - 1.0) > 1e-4:
        log("WARNING: ranks do not sum to 1")

    n, m = g.n, g.m
    b_iter = 8 * m + 16 * n + 4
    achieved = b_iter * sweeps / (sweep_ms / 1e3) / 1e9
    default_layout = args.engine == "auto" and args.combine == "auto" and not args.tile_bits
    traffic = load_traffic(args.scale, args.edge_factor, args.workload) if default_layout else None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": {"tiled": "k_tile+k_combine", "tiled-accumulate": "k_tile+k_accum"}.get(engine, "k_sweep"),
            "engine": engine,
            "algorithmic_bytes_per_sweep": b_iter,
            "avg_launch_us": round(sweep_ms / launches * 1e3, 2), "launches": launches, "sweeps": sweeps,
            "avg_sweep_us": round(sweep_ms / sweeps * 1e3, 2),
            "launch": "one sweep: k_tile + k_combine or k_accum (tiled engine) or k_sweep (binned); the stopping launch "
                      "of each solve is timed with the sweeps"}
    if traffic:
        # what the design could reach: its measured bytes moved at the copy rate
        cr = copy_rate_gbs()
        t_min = traffic / (cr * 1e9)
        roof["ceiling"] = {"copy_rate_gbs": round(cr, 1), "traffic_bytes_per_sweep": traffic,
                           "sweep_us_at_copy_rate": round(t_min * 1e6, 2),
                           "frac_at_copy_rate": round(b_iter / t_min / 1e9 / HBM_PEAK_GBS, 4),
                           "traffic_rate_gbs": round(traffic * sweeps / (sweep_ms / 1e3) / 1e9, 1),
                           "traffic_rate_over_copy_rate": round(traffic * sweeps / (sweep_ms / 1e3) / 1e9 / cr, 3)}
    value = m * sweeps / wall / 1e9

    cpu = None
    if args.cpu_seconds > 0 and args.workload == "pagerank":
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads (~{args.cpu_seconds:.0f}s)")
        cpu = cpu_baseline(g, deg, args.cpu_seconds, threads, args.tol, args.max_iters)
        cpu["value"] = round(cpu["value"], 4)
        cpu["host"] = host_info(threads)
        log(f"cpu baseline {cpu['value']} GTEPS, {cpu['ms_per_solve']:.1f} ms per solve")
    fa = None
    if args.fwd_adjoint and args.workload == "pagerank":
        fa = fwd_adjoint_cost(args.seed)
        log(f"fwd+adjoint / fwd = {fa['ratio']} ({fa['fwd_adjoint_ms']} / {fa['fwd_ms']} ms)")

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"pagerank {gname}, damping 0.85, dangling uniform, tol {args.tol:g} L1, max "
                               f"{args.max_iters} sweeps, FP32",
                   "n": n, "m": m, "sweeps_per_solve": iters[0], "device_ms_per_solve": round(dev_ms / args.steps, 4),
                   "device_ms_per_solve_median": round(float(np.median(solve_ms)), 4),
                   "setup_ms": round(pg.setup_ms, 1), "engine": engine,
                   "host_output_ms_per_solve": round(host_ms, 3)},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if cpu:
        # time to tolerance on the problem the CPU path solves: the reference
        # leaks the dangling mass (page_rank.hpp:94,133), which takes more
        # sweeps to reach the same L1 tolerance than the UNIFORM headline
        kl = dict(kw, dangling="leak")
        res_l = pg.run_device(out_d.data_ptr(), **kl)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            res_l = pg.run_device(out_d.data_ptr(), **kl)
        torch.cuda.synchronize()
        leak_ms = (time.perf_counter() - t1) * 1e3 / args.steps
        line["gpu_leak_solve"] = {"ms_per_solve": round(leak_ms, 4), "sweeps_per_solve": res_l.iters}
        # the in-place (block Gauss-Seidel) form of the same LEAK solve: fewer
        # sweeps, two launches per phase and sweep
        gs = {}
        for phases in (2, 4, 8):
            res_g = pg.run_gs_device(out_d.data_ptr(), phases, damping=0.85, tol=args.tol, max_iters=args.max_iters)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.steps):
                res_g = pg.run_gs_device(out_d.data_ptr(), phases, damping=0.85, tol=args.tol,
                                         max_iters=args.max_iters)
            torch.cuda.synchronize()
            gs_ms = (time.perf_counter() - t1) * 1e3 / args.steps
            gs[str(phases)] = {"ms_per_solve": round(gs_ms, 4), "sweeps_per_solve": res_g.iters,
                               "us_per_sweep": round(gs_ms * 1e3 / res_g.iters, 2)}
        line["gpu_leak_solve_block_gauss_seidel"] = {
            "phases": gs, "cpu_sweeps_per_solve": cpu["sweeps_per_solve"],
            "note": "rows cut into phases updated in place one after the other (nwgr_pr_gs_run_device_f32); "
                    "time to tolerance beside the Jacobi LEAK solve (gpu_leak_solve)"}
        line["gpu_vs_cpu_per_sweep"] = round(value / cpu["value"], 1) if cpu["value"] > 0 else None
        line["gpu_vs_cpu_time_to_tolerance"] = round(cpu["ms_per_solve"] / leak_ms, 1)
    if args.workload == "pagerank":
        di = dropin_calls(g, deg, args.tol, args.max_iters, max(args.steps, 3), pg, out_d)
        log(f"drop-in: first call {di['first_call_ms']} ms, repeated {di['repeat_call_ms_median']} ms, resident "
            f"{di['resident_solve_device_output_ms_median']} ms")
        line["dropin"] = di
        if cpu:
            # whole calls on both sides: pr.exe's timed region on the host path
            # and on the drop-in (the GPU's includes the copy of the ranks)
            line["dropin"]["cpu_whole_call_ms"] = round(cpu["ms_per_solve"], 2)
            line["dropin"]["repeat_call_vs_cpu_whole_call"] = round(cpu["ms_per_solve"] / di["repeat_call_ms_median"], 1)
    line["fwd_adjoint_vs_fwd"] = fa
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
