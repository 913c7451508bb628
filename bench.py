#!/usr/bin/env python3
"""Benchmark of the PageRank hot path on one MI355X.

Headline (`value`, `roofline`, `cpu_baseline`): plain PageRank on a directed RMAT graph,
scale 25 (33,554,432 vertices), edge factor 16, A/B/C/D = .57/.19/.19/.05, self-loops and
duplicates removed, int32 CSC, float32 scores, damping 0.85, L1 tolerance 1e-4, 20 sweeps.
The graph is generated and built on the device (not timed).  One step = one page_rank
solve (init + up to 20 sweeps, one captured HIP graph) with the CSC resident in HBM.  The
CPU baseline is the oracle's restatement of the reference's v11 loop on this host's cores,
over a bounded sample of sweeps on the same graph.

The default run (--workload all) adds every other configuration under `configs`, each
with its own timing and roofline: RMAT-16 and the 4096^2 grid (with CPU baselines), 64
personalised vectors on RMAT-22 batched against 64 sequential solves, 256 RMAT graphs of
scale 10..18 in one captured loop, and the weighted forward + implicit adjoint on RMAT-20.
--workload picks one of them alone (plain, grid, ppr, multi, wpr), or the bandwidth
calibrations (copy, ceilings).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes of one sweep measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (each scaled by
# what the counter reads for a stream copy of known size), for the configurations profiled in
# profiles/: (workload, RMAT scale or grid width, binned, relabel, bin_bits) -> (bytes per sweep,
# profile file)
MEASURED_TRAFFIC = {
    ("rmat", 25, 1, 1, 14): (3176601920, "profiles/pr_binned_sweep_rmat25.md"),
    ("rmat", 25, 0, 1, 0): (23300197580, "profiles/pr_sweep_rmat25_relabel.md"),
    ("grid", 4096, 1, 0, 14): (708700864, "profiles/pr_binned_sweep_grid4096.md"),
}
METRIC = "GTEPS per PR iteration + % HBM roofline; batched PPR/adjoint vs plain"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bytes_per_iter(n, nnz):
    # col ids 4E + row_ptr 4(N+1) + compulsory contrib gather 4N + x read 4N + outdeg 4N
    # + x write 4N + contrib write 4N
    return 4 * nnz + 4 * (n + 1) + 20 * n


def host_threads():
    """Every CPU this process may run on (the reference's TBB arena takes all of them)."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def host_cpu():
    """CPU model, logical CPUs, the usable ones, and the physical cores / SMT behind them
    (from /proc/cpuinfo's physical id and core id of each usable logical CPU)."""
    model, cores = "unknown", {}
    try:
        usable = set(os.sched_getaffinity(0))
    except AttributeError:
        usable = set(range(os.cpu_count() or 1))
    try:
        with open("/proc/cpuinfo") as f:
            blocks = f.read().split("\n\n")
        for b in blocks:
            kv = dict(line.split(":", 1) for line in b.splitlines() if ":" in line)
            kv = {k.strip(): v.strip() for k, v in kv.items()}
            if "processor" not in kv:
                continue
            model = kv.get("model name", model)
            cpu = int(kv["processor"])
            if cpu in usable:
                cores.setdefault((kv.get("physical id", "0"), kv.get("core id", str(cpu))), []).append(cpu)
    except (OSError, ValueError):
        pass
    phys = len(cores) or len(usable)
    return {"model": model, "logical_cpus": os.cpu_count() or 1, "usable_cpus": len(usable),
            "usable_physical_cores": phys, "smt_per_core": round(len(usable) / max(phys, 1), 2)}


_BASELINE = {}


def cpu_baseline(ptr, idx, deg, sweeps):
    """The oracle's restatement of the reference's v11 loop (page_rank.hpp:124-139, in-place
    contributions, OpenMP static chunks) on every usable host CPU, built with the reference's
    -Ofast -march=native on this host (oracle/Makefile `baseline`; the prebuilt -O3
    x86-64-v3 oracle if that build fails).  `sweeps` timed sweeps (threshold 0), reported
    min / avg / max as the reference's -n trials are, plus one sweep on one thread."""
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    if "lib" not in _BASELINE:
        _BASELINE["dir"] = tempfile.mkdtemp(prefix="nwg_baseline_")
        _BASELINE["lib"], _BASELINE["flags"] = O.baseline_lib(_BASELINE["dir"])
    lib, flags = _BASELINE["lib"], _BASELINE["flags"]
    if lib is None:
        flags = f"-O3 -march=x86-64-v3 -fopenmp (prebuilt; {flags})"
    threads = host_threads()
    n = ptr.size - 1
    _, it, _, sec = O.pr_v11_parallel(n, ptr, idx, deg, 0.85, 0.0, sweeps, threads, lib=lib)
    t = float(np.mean(sec))
    # one sweep on one thread: the sequential Gauss-Seidel form of the same loop
    _, _, _, sec1 = O.pr_v11_parallel(n, ptr, idx, deg, 0.85, 0.0, 1, 1, lib=lib)
    return {
        "value": round(idx.size / t / 1e9, 4),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{it} sweeps of the v11 in-place loop (page_rank.hpp:124-139 restated in C, OpenMP static "
                  f"chunks, {flags}) on the same CSC copied to the host, {threads} threads; mean sweep "
                  f"{t * 1e3:.2f} ms",
        "sweep_ms_min_avg_max": [round(float(np.min(sec)) * 1e3, 3), round(t * 1e3, 3),
                                 round(float(np.max(sec)) * 1e3, 3)],
        "one_thread_sweep_ms": round(float(sec1[0]) * 1e3, 3),
        "host": host_cpu(),
    }


def validate(g, dev_buf, ptr, idx, deg, d=0.85):
    """Full-size checks: one more sweep equals the FP64 update of the previous iterate on
    sampled rows (hubs included), and mass balance over all edges."""
    import torch

    n = g.n
    g.pagerank_device(dev_buf.data_ptr(), None, d, 0.0, 19)
    x19 = dev_buf.cpu().numpy().astype(np.float64)
    g.pagerank_device(dev_buf.data_ptr(), None, d, 0.0, 20)
    torch.cuda.synchronize()
    x20 = dev_buf.cpu().numpy().astype(np.float64)
    c = x19 / np.where(deg > 0, deg, 1)
    base = float(np.float32((1.0 - np.float64(np.float32(d))) / n))
    dd = float(np.float32(d))
    rng = np.random.default_rng(0)
    sample = np.concatenate([rng.integers(0, n, 20000), np.argsort(np.diff(ptr.astype(np.int64)))[-100:]])
    worst = 0.0
    for i in sample:
        want = base + dd * c[idx[ptr[i]:ptr[i + 1]]].sum()
        worst = max(worst, abs(x20[i] - want) / max(want, 1e-30))
    lhs = x20.sum()
    rhs = n * base + dd * x19[deg > 0].sum()
    return {"sampled_rows": int(sample.size), "max_rel_err": worst, "mass_balance_rel": abs(lhs - rhs) / rhs,
            "ok": bool(worst < 1e-5 and abs(lhs - rhs) / rhs < 1e-5)}


def base_line(args, value, unit, ms_per_step, dtype, config):
    return {
        "metric": METRIC,
        "value": value,
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
    }


def apply_plan(g, args):
    if args.plan:
        plan = g.get_plan()
        plan.update({k: int(v) for k, v in (kv.split("=") for kv in args.plan.split(","))})
        g.set_plan(**plan)
    if args.autotune:
        t0 = time.perf_counter()
        chosen = g.autotune()
        log(f"autotune chose {chosen} in {time.perf_counter() - t0:.1f} s")
    return g.get_plan()


# ---------------------------------------------------------------------------- plain / grid
def run_plain(args, ctx, grid=False):
    import torch

    t0 = time.perf_counter()
    if grid:
        g = ctx.grid(args.width, 8500, 50, seed=args.seed)
        iters = args.iters if args.iters_given else 100
        workload = (f"plain PageRank, {args.width}x{args.width} 4-neighbour grid, each undirected edge kept with "
                    f"p=0.85 and stored both ways, +0.5% random long-range edges, cleaned, d=0.85, L1 tol "
                    f"{args.tol}, max {iters} sweeps")
    else:
        g = ctx.rmat(args.scale, args.edge_factor, seed=args.seed)
        iters = args.iters
        workload = (f"plain PageRank, directed RMAT scale {args.scale}, edge factor {args.edge_factor}, "
                    f"A/B/C/D .57/.19/.19/.05, cleaned, int32 CSC, d=0.85, L1 tol {args.tol}, max {iters} sweeps")
    build_s = time.perf_counter() - t0
    n, nnz = g.n, g.nnz
    log(f"n={n} nnz={nnz} generated+built on device in {build_s:.2f} s")
    plan = apply_plan(g, args)

    buf = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        g.pagerank_device(buf.data_ptr(), None, 0.85, args.tol, iters, inplace=args.inplace)
    log(f"warmup {args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = [g.pagerank_device(buf.data_ptr(), None, 0.85, args.tol, iters, inplace=args.inplace)
             for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    sweeps = infos[-1]["iterations"]
    ev_ms = float(np.mean([i["ms_total"] for i in infos]))
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step wall, {ev_ms:.3f} ms/step events, {sweeps} sweeps")

    # live per-sweep kernel timing (HIP events on the library's stream around each sweep's launches)
    kinfo = [g.pagerank_device(buf.data_ptr(), None, 0.85, args.tol, iters, kernel_timing=True, inplace=args.inplace)
             for _ in range(3)]
    sweep_ms = float(np.median([k["sweep_ms_avg"] for k in kinfo]))
    b_iter = bytes_per_iter(n, nnz)
    achieved = b_iter / (sweep_ms * 1e-3) / 1e9
    log(f"sweep {sweep_ms:.4f} ms -> {achieved:.1f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.3f} of peak)")

    value = nnz * sweeps / (ms_per_step * 1e-3) / 1e9
    binned = bool(plan.get("binned")) and not args.inplace  # the in-place schedule runs the pull kernels
    if binned:
        kernel_desc = "one sweep (pr_chunk_kernel + pr_bin_kernel + pr_bin_combine_kernel)"
    else:
        kernel_desc = ("one sweep (pr_tile_kernel + pr_group_kernel<4..64> + pr_sgroup_kernel<4..64> + pr_hubseg_kernel"
                       " + pr_finish_kernel, those the plan uses)")
    traffic, traffic_src = args.traffic, "--traffic" if args.traffic else None
    if traffic is None and (grid or args.edge_factor == 16):
        key = ("grid", args.width) if grid else ("rmat", args.scale)
        hit = MEASURED_TRAFFIC.get(key + (int(binned), int(plan.get("relabel", 0)),
                                          int(plan.get("bin_bits", 0)) if binned else 0))
        if hit:
            traffic, traffic_src = hit[0], f"rocprofv3 FETCH_SIZE + WRITE_SIZE, {hit[1]}"
    out = base_line(args, round(value, 3), "GTEPS", ms_per_step, "f32", {
        "workload": workload,
        "n": n,
        "nnz": nnz,
        "sweeps_per_step": sweeps,
        "ms_per_sweep_events": round(ev_ms / max(sweeps, 1), 5),
        "build_s": round(build_s, 3),
        "plan": plan,
        "schedule": "in-place (v11)" if args.inplace else "Jacobi (v10)",
        "parallelism": "single GPU",
    })
    out["roofline"] = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "kernel": kernel_desc,
        "kernel_ms": round(sweep_ms, 5),
        "algorithmic_bytes": b_iter,
    }
    if not args.no_cpu or not args.no_validate:
        ptr, idx, deg = g.csc()
    if not args.no_validate:
        out["validation"] = validate(g, buf, ptr, idx, deg)
        log(f"validation {out['validation']}")
    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(ptr, idx, deg, args.cpu_sweeps)
        log(f"cpu baseline {out['cpu_baseline']['value']} GTEPS in {time.perf_counter() - t0:.1f} s")
    g.close()
    return out


# ---------------------------------------------------------------------------- batched PPR
def run_ppr(args, ctx):
    import torch

    import nwgraph_amd

    g = ctx.rmat(args.ppr_scale, args.edge_factor, seed=args.seed)
    n, nnz, k = g.n, g.nnz, args.k
    plan = apply_plan(g, args)
    seeds = nwgraph_amd.ppr_seeds(n, k, 8, args.seed)
    tile = g.ppr_autotune(seeds)
    log(f"RMAT-{args.ppr_scale}: n={n} nnz={nnz}; k={k}, autotuned k-tile {tile}")
    X = torch.empty((n, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        g.ppr_device(seeds, X.data_ptr(), 0.85, args.tol, args.iters)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        its, info = g.ppr_device(seeds, X.data_ptr(), 0.85, args.tol, args.iters)
    torch.cuda.synchronize()
    ms_batched = (time.perf_counter() - t0) * 1e3 / args.steps
    # 64 sequential personalised solves with the plain kernel
    x1 = torch.empty(n, dtype=torch.float32, device="cuda")
    teles = []
    for j in range(k):
        v = torch.zeros(n, dtype=torch.float32, device="cuda")
        v[torch.as_tensor(seeds[j].astype(np.int64), device="cuda")] = 1.0 / seeds.shape[1]
        teles.append(v)
    torch.cuda.synchronize()
    for j in range(min(k, 2)):
        g.pagerank_device(x1.data_ptr(), teles[j].data_ptr(), 0.85, args.tol, args.iters)
    t0 = time.perf_counter()
    seq_its, seq_dev = [], 0.0
    for j in range(k):
        si = g.pagerank_device(x1.data_ptr(), teles[j].data_ptr(), 0.85, args.tol, args.iters)
        seq_its.append(si["iterations"])
        seq_dev += si["ms_total"]
    torch.cuda.synchronize()
    ms_seq = (time.perf_counter() - t0) * 1e3
    # 64 sequential plain solves (uniform teleport) on the same graph and plan
    t0 = time.perf_counter()
    plain_its, plain_dev = [], 0.0
    for j in range(k):
        si = g.pagerank_device(x1.data_ptr(), 0, 0.85, args.tol, args.iters)
        plain_its.append(si["iterations"])
        plain_dev += si["ms_total"]
    torch.cuda.synchronize()
    ms_plain = (time.perf_counter() - t0) * 1e3
    vec_sweeps = int(np.sum(its))
    value = nnz * vec_sweeps / (ms_batched * 1e-3) / 1e9
    out = base_line(args, round(value, 3), "GTEPS (vector-sweeps x edges / s)", ms_batched, "f32", {
        "workload": f"batched personalised PageRank, k={k} teleports uniform over 8 seeded sources each, RMAT scale "
                    f"{args.ppr_scale} ef {args.edge_factor}, d=0.85, tol {args.tol}, max {args.iters} sweeps",
        "n": n, "nnz": nnz, "k": k, "k_tile": tile, "plan": plan,
        "vector_sweeps": vec_sweeps, "sequential_vector_sweeps": int(np.sum(seq_its)),
        "ms_batched": round(ms_batched, 3), "ms_64_sequential": round(ms_seq, 3),
        "device_ms_batched": round(info["ms_total"], 3), "device_ms_64_sequential": round(seq_dev, 3),
        "speedup_vs_sequential": round(ms_seq / ms_batched, 3),
        "plain_solve_sweeps": int(np.sum(plain_its)), "ms_64_sequential_plain": round(ms_plain, 3),
        "device_ms_64_sequential_plain": round(plain_dev, 3),
        "speedup_vs_sequential_plain": round(ms_plain / ms_batched, 3),
        "useful_vector_sweep_fraction": round(vec_sweeps / (k * max(1, int(np.max(its)))), 3),
    })
    # algorithmic bytes of the batched solve: per batched sweep the shared streams (ids, row
    # pointers, out-degrees) once, per vector-sweep a contribution read, a score read, a
    # score write and a contribution write (16 N)
    sweeps = int(np.max(its))
    alg = sweeps * (4 * nnz + 4 * (n + 1) + 4 * n) + 16 * n * vec_sweeps
    out["roofline"] = roofline_entry(alg, info["ms_total"],
                                     "batched solve (ppr_init + ppr_sweep_kernel + ppr_finish_kernel per sweep), "
                                     "device time of the whole solve",
                                     "sweeps x (4E + 4(N+1) + 4N) + 16N x vector-sweeps")
    g.close()
    return out


# ---------------------------------------------------------------------------- many graphs
def run_multi(args, ctx):
    rng = np.random.default_rng(args.seed)
    scales = rng.integers(10, 19, args.ngraphs)
    graphs = [ctx.rmat(int(s), args.edge_factor, seed=1000 + q) for q, s in enumerate(scales)]
    mg = ctx.multigraph(graphs)
    log(f"{mg.ngraphs} graphs, scales 10..18: n={mg.n} nnz={mg.nnz}")
    binned = mg.autotune() if args.autotune else 0
    log(f"schedule: {'binned sweep of the union' if binned else 'persistent item kernel'}")
    for _ in range(args.warmup):
        mg.pagerank(0.85, args.tol, args.iters)
    t0 = time.perf_counter()
    ev = []
    for _ in range(args.steps):
        _, its, _, info = mg.pagerank(0.85, args.tol, args.iters)
        ev.append(info["ms_total"])
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    edge_sweeps = int(sum(int(g.nnz) * int(i) for g, i in zip(graphs, its)))
    value = edge_sweeps / (np.mean(ev) * 1e-3) / 1e9
    alg = int(sum(bytes_per_iter(int(g.n), int(g.nnz)) * int(i) for g, i in zip(graphs, its)))
    out = base_line(args, round(value, 3), "GTEPS (sum over graphs, device time)", ms, "f32", {
        "workload": f"{mg.ngraphs} independent RMAT graphs, scales U{{10..18}} (seed {args.seed}), ef "
                    f"{args.edge_factor}, all sweeps in one HIP graph, per-graph tol "
                    f"{args.tol}, max {args.iters}",
        "n": mg.n, "nnz": mg.nnz, "ms_device": round(float(np.mean(ev)), 3),
        "schedule": "binned" if binned else "items",
        "iterations_min_max": [int(its.min()), int(its.max())],
    })
    out["roofline"] = roofline_entry(alg, float(np.mean(ev)), "all sweeps of the union in one captured HIP graph "
                                      "(device time)", "sum over graphs of sweeps x (4E + 4(N+1) + 20N)")
    if binned and args.ngraphs == 256:
        out["traffic_per_sweep"] = {"bytes": 1358957008, "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                    "profiles/pr_binned_multigraph_256.md"}
    mg.close()
    for g in graphs:
        g.close()
    return out


# ---------------------------------------------------------------------------- weighted
def run_wpr(args, ctx):
    import torch

    g = ctx.rmat(args.wpr_scale, args.edge_factor, seed=args.seed)
    n, nnz = g.n, g.nnz
    g.random_weights(seed=args.seed, normalize=True)
    v = torch.full((n,), 1.0 / n, dtype=torch.float32, device="cuda")
    x = torch.empty(n, dtype=torch.float32, device="cuda")
    c = torch.randn(n, generator=torch.Generator().manual_seed(args.seed)).float().cuda()
    gw = torch.empty(nnz, dtype=torch.float32, device="cuda")
    gv = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    # both solves stop on the relative L1 tolerance (NWG_INFO_RELATIVE_THRESHOLD): the forward at
    # tol * ||v||_1 (= tol: the same stop as an absolute tolerance), the adjoint at
    # tol * ||c||_1 / (1 - d), the same relative accuracy for a right-hand side of any scale
    for _ in range(args.warmup):
        g.wpr_forward_device(v.data_ptr(), x.data_ptr(), 0.85, args.tol, args.iters, relative=True)
        g.wpr_backward_device(x.data_ptr(), c.data_ptr(), gw.data_ptr(), gv.data_ptr(), 0.85, args.tol, args.iters,
                              relative=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fi = g.wpr_forward_device(v.data_ptr(), x.data_ptr(), 0.85, args.tol, args.iters, relative=True)
    torch.cuda.synchronize()
    ms_f = (time.perf_counter() - t0) * 1e3 / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bi = g.wpr_backward_device(x.data_ptr(), c.data_ptr(), gw.data_ptr(), gv.data_ptr(), 0.85, args.tol,
                                   args.iters, relative=True)
    torch.cuda.synchronize()
    ms_b = (time.perf_counter() - t0) * 1e3 / args.steps
    # the adjoint with an absolute tolerance: its L1 changes start near d ||A^T c||_1 ~ 1e5-1e6,
    # about ten orders of magnitude above 1e-4 (some 140 sweeps at d = 0.85), where the forward's
    # start near 1 is four
    ba = g.wpr_backward_device(x.data_ptr(), c.data_ptr(), gw.data_ptr(), gv.data_ptr(), 0.85, args.tol, args.iters)
    torch.cuda.synchronize()
    value = nnz * fi["iterations"] / (ms_f * 1e-3) / 1e9
    sweep_b = 8 * nnz + 4 * (n + 1) + 12 * n
    alg_f = sweep_b * fi["iterations"]
    alg_b = sweep_b * bi["iterations"] + 12 * nnz + 12 * n
    out = base_line(args, round(value, 3), "GTEPS (forward)", ms_f + ms_b, "f32", {
        "workload": f"weighted PageRank + implicit adjoint, RMAT scale {args.wpr_scale} ef {args.edge_factor}, "
                    f"w ~ U(0.1,1) row-normalised, L = sum c_i x_i, c ~ N(0,1), d=0.85, relative L1 tol "
                    f"{args.tol} (forward: x {args.tol} * ||v||_1; adjoint: x {args.tol} * ||c||_1 / (1-d)), "
                    f"max {args.iters}",
        "n": n, "nnz": nnz, "forward_ms": round(ms_f, 3), "backward_ms": round(ms_b, 3),
        "backward_over_forward": round(ms_b / ms_f, 3),
        "forward_sweeps": fi["iterations"], "adjoint_sweeps": bi["iterations"],
        "forward_device_ms": round(fi["ms_total"], 3), "backward_device_ms": round(bi["ms_total"], 3),
        "backward_over_forward_device": round(bi["ms_total"] / fi["ms_total"], 3),
        "forward_threshold": fi["threshold_used"], "adjoint_threshold": bi["threshold_used"],
        "adjoint_absolute_tolerance": {"threshold": float(np.float32(args.tol)), "sweeps": ba["iterations"],
                                       "final_error": ba["final_error"], "device_ms": round(ba["ms_total"], 3)},
    })
    out["roofline"] = roofline_entry(alg_f, fi["ms_total"], "weighted forward solve (binned weighted sweep: "
                                     "pr_chunk_kernel + pr_bin_kernel + pr_bin_combine_kernel + pr_wfold_kernel "
                                     "per sweep), device time", "sweeps x (8E + 4(N+1) + 12N)")
    out["traffic_per_sweep"] = {"bytes": 285350032, "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                "profiles/wpr_adjoint_rmat20.md"} if args.wpr_scale == 20 else None
    out["roofline_backward"] = roofline_entry(alg_b, bi["ms_total"], "adjoint solve over the CSR + edge and "
                                              "teleport gradients, device time",
                                              "sweeps x (8E + 4(N+1) + 12N) + 12E + 12N")
    g.close()
    return out


def roofline_entry(alg_bytes, ms, kernel, formula):
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel, "kernel_ms": round(ms, 4),
            "algorithmic_bytes": int(alg_bytes), "bytes_formula": formula}


# ---------------------------------------------------------------------------- every config
def sub_args(args, **kw):
    a = argparse.Namespace(**vars(args))
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def summary(out, keep=()):
    """The parts of a workload's line that the default run carries per configuration."""
    r = {"value": out["value"], "unit": out["unit"], "ms_per_step": out["ms_per_step"]}
    cfg = out["config"]
    r["workload"] = cfg["workload"]
    for k in ("n", "nnz", "sweeps_per_step", "ms_per_sweep_events", "plan") + tuple(keep):
        if k in cfg:
            r[k] = cfg[k]
    for k in ("roofline", "roofline_backward", "traffic_per_sweep", "validation", "cpu_baseline"):
        if k in out:
            r[k] = out[k]
    return r


def run_default(args, ctx):
    """The headline (plain PageRank on RMAT-25: `value`, `roofline`, `cpu_baseline`) and,
    under `configs`, every other benchmark configuration with its own timing and roofline
    (and, for the plain solves, its CPU baseline): RMAT-16, the 4096^2 grid, 64 batched
    personalised vectors on RMAT-22 against 64 sequential solves, 256 graphs in one captured
    loop, and the weighted forward / adjoint on RMAT-20."""
    t_all = time.perf_counter()
    out = run_plain(args, ctx)
    configs = {}

    def add(name, fn):
        t0 = time.perf_counter()
        configs[name] = fn()
        configs[name]["bench_s"] = round(time.perf_counter() - t0, 1)
        log(f"config {name}: {configs[name]['value']} {configs[name]['unit']} in {configs[name]['bench_s']} s")

    add("rmat16", lambda: summary(run_plain(sub_args(args, scale=16, steps=50, warmup=10), ctx)))
    add("grid4096", lambda: summary(run_plain(sub_args(args, iters=100, iters_given=True, steps=10, warmup=3), ctx,
                                              grid=True)))
    add("ppr_rmat22_k64", lambda: summary(run_ppr(sub_args(args, steps=3, warmup=1), ctx), keep=(
        "k", "k_tile", "vector_sweeps", "ms_batched", "device_ms_batched", "ms_64_sequential_plain",
        "device_ms_64_sequential_plain", "speedup_vs_sequential_plain", "ms_64_sequential",
        "speedup_vs_sequential", "useful_vector_sweep_fraction")))
    add("multigraph_256", lambda: summary(run_multi(sub_args(args, steps=3, warmup=1), ctx), keep=(
        "ms_device", "schedule", "iterations_min_max")))
    add("wpr_rmat20", lambda: summary(run_wpr(sub_args(args, iters=100, steps=3, warmup=1), ctx), keep=(
        "forward_ms", "backward_ms", "forward_device_ms", "backward_device_ms", "backward_over_forward",
        "backward_over_forward_device", "forward_sweeps", "adjoint_sweeps", "forward_threshold", "adjoint_threshold",
        "adjoint_absolute_tolerance")))
    out["configs"] = configs
    log(f"all configurations in {time.perf_counter() - t_all:.1f} s")
    return out


# ---------------------------------------------------------------------------- calibration
def run_copy(args, ctx):
    """Streaming device copy of a known size: the measured HBM ceiling, and the reference
    transfer for scaling the FETCH_SIZE / WRITE_SIZE counters of the sweep."""
    import ctypes

    import torch

    nbytes = args.copy_mib << 20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    ms = ctypes.c_double()
    for _ in range(args.warmup):
        ctx.check(ctx.lib.nwg_stream_copy(ctx.handle, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                                          nbytes, ctypes.byref(ms)))
    times = []
    for _ in range(args.steps):
        ctx.check(ctx.lib.nwg_stream_copy(ctx.handle, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                                          nbytes, ctypes.byref(ms)))
        times.append(ms.value)
    t = float(np.median(times))
    gbs = 2 * nbytes / (t * 1e-3) / 1e9
    return base_line(args, round(gbs, 1), "GB/s (read + write)", t, "u8", {
        "workload": f"device stream copy of {args.copy_mib} MiB (16-byte vector loads/stores)",
        "bytes_read": nbytes, "bytes_written": nbytes, "kernel_ms": round(t, 4),
        "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4),
    })


def run_ceilings(args, ctx):
    """The device's streaming ceilings: copy, read-only and write-only passes over
    --copy-mib MiB, with and without nontemporal loads / stores, at 2, 4 and 8 workgroups
    per CU.  Rates count the bytes the pass must move (read + written)."""
    import ctypes

    import torch

    nbytes = args.copy_mib << 20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    ms = ctypes.c_double()
    names = {0: "copy", 1: "read", 2: "write"}
    results = []
    for kind in (0, 1, 2):
        for flags in (0, 1, 2, 3):
            if (kind == 1 and flags & 2) or (kind == 2 and flags & 1):
                continue  # a read pass has no stores, a write pass no loads
            for wg in (2, 4, 8):
                times = []
                for _ in range(args.warmup + args.steps):
                    ctx.check(ctx.lib.nwg_stream_probe(ctx.handle, kind, flags, wg, ctypes.c_void_p(b.data_ptr()),
                                                       ctypes.c_void_p(a.data_ptr()), nbytes, ctypes.byref(ms)))
                    times.append(ms.value)
                t = float(np.median(times[args.warmup:]))
                moved = nbytes * (2 if kind == 0 else 1)
                gbs = moved / (t * 1e-3) / 1e9
                results.append({"pass": names[kind], "nt_load": bool(flags & 1), "nt_store": bool(flags & 2),
                                "wg_per_cu": wg, "ms": round(t, 4), "GBps": round(gbs, 1)})
                log(f"{names[kind]:5s} ntl={flags & 1} nts={(flags >> 1) & 1} wg/cu={wg}: {gbs:8.1f} GB/s")
    best = {k: max(r["GBps"] for r in results if r["pass"] == k) for k in ("copy", "read", "write")}
    out = base_line(args, best["copy"], "GB/s (best copy, read + write)", 0.0, "u8", {
        "workload": f"streaming ceilings over {args.copy_mib} MiB: copy / read-only / write-only, nontemporal "
                    f"variants, 2-8 workgroups per CU, four 16-byte accesses in flight per lane",
        "best_GBps": best, "frac_of_8TBs": {k: round(v / HBM_PEAK_GBS, 4) for k, v in best.items()},
        "probes": results,
    })
    return out


def run_mall(args, ctx):
    """Whether a buffer is served from the last-level cache (MALL) rather than HBM, by size:
    the rate of a pass that reads one buffer 16 times in one launch, and of a read that follows
    a write of the same buffer (as a pass that consumes what the previous pass produced)."""
    import ctypes

    import torch

    ms = ctypes.c_double()
    sizes = [32, 64, 128, 192, 256, 384, 512, 1024, 4096]
    a = torch.empty((max(sizes) << 20) // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()

    def probe(kind, flags, dst, src, nbytes):
        ctx.check(ctx.lib.nwg_stream_probe(ctx.handle, kind, flags, 2, ctypes.c_void_p(dst.data_ptr()),
                                           ctypes.c_void_p(src.data_ptr()), nbytes, ctypes.byref(ms)))
        return ms.value

    rows = []
    for mib in sizes:
        nb = mib << 20
        rr, wr, rd = [], [], []
        for _ in range(args.warmup + args.steps):
            rr.append(probe(3, 0, b, a, nb))
            wr.append(probe(2, 0, b, a, nb))
            rd.append(probe(1, 0, a, b, nb))
        k = args.warmup
        row = {"MiB": mib, "reread16_GBps": round(16 * nb / (np.median(rr[k:]) * 1e-3) / 1e9, 1),
               "write_GBps": round(nb / (np.median(wr[k:]) * 1e-3) / 1e9, 1),
               "read_after_write_GBps": round(nb / (np.median(rd[k:]) * 1e-3) / 1e9, 1),
               "read_after_write_us": round(float(np.median(rd[k:])) * 1e3, 2)}
        rows.append(row)
        log(f"{mib:5d} MiB: reread x16 {row['reread16_GBps']:8.1f} GB/s, write {row['write_GBps']:8.1f}, read after "
            f"write {row['read_after_write_GBps']:8.1f} GB/s ({row['read_after_write_us']} us)")
    return base_line(args, rows[0]["reread16_GBps"], "GB/s (reread x16 at the smallest size)", 0.0, "u8", {
        "workload": "cache residency by buffer size: one buffer read 16 times per launch; a read after a write of "
                    "the same buffer; 2 workgroups of 256 lanes per CU, four 16-byte accesses in flight per lane",
        "rows": rows})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all", "plain", "grid", "ppr", "multi", "wpr", "copy", "ceilings", "mall"],
                    help="all (default): the RMAT-25 headline plus every other configuration under 'configs'")
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--ppr-scale", type=int, default=22)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--ngraphs", type=int, default=256)
    ap.add_argument("--wpr-scale", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=4096)
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sweeps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default sweep plan instead of autotuning it before the timed region")
    ap.add_argument("--plan", default="", help="k=v,... overrides of the sweep plan")
    ap.add_argument("--inplace", action="store_true",
                    help="plain/grid: the reference's in-place (v11) schedule instead of Jacobi (v10)")
    ap.add_argument("--traffic", type=float, default=None, help="measured HBM bytes per sweep (rocprofv3 PMC)")
    args = ap.parse_args()
    args.iters_given = args.iters is not None
    if args.iters is None:
        args.iters = 20

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; multi-GPU sharding is designed "
                                   "(see the sharding section of the design notes) but not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch  # noqa: F401  (device buffers; loaded before the library: one HIP runtime)

    import nwgraph_amd

    ctx = nwgraph_amd.Context(0)
    if args.workload == "all":
        out = run_default(args, ctx)
    elif args.workload == "plain":
        out = run_plain(args, ctx)
    elif args.workload == "grid":
        out = run_plain(args, ctx, grid=True)
    elif args.workload == "ppr":
        out = run_ppr(args, ctx)
    elif args.workload == "multi":
        out = run_multi(args, ctx)
    elif args.workload == "copy":
        out = run_copy(args, ctx)
    elif args.workload == "ceilings":
        out = run_ceilings(args, ctx)
    elif args.workload == "mall":
        out = run_mall(args, ctx)
    else:
        out = run_wpr(args, ctx)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
