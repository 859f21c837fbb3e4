"""Benchmark: exact-cost A* queries/s, batched vs one-at-a-time GA*, on one MI355X.

Workload (default): 4,096 source/target queries on a random geometric graph
of 2,000,000 nodes (integer lattice [0,100000]^2, seed 4, mean degree ~10,
edges both ways, w = ceil(Euclidean distance), h = floor(Euclidean distance
to the target)), query separations log-uniform in [100, 100000] lattice
units.  One step = one batched launch answering all 4,096 queries with their
paths.  `value` = queries/s of the batched launches (graph and queries already
resident in HBM); the same queries are also answered one at a time by the
single-query GA* kernel and their rate and latencies are reported beside it.

Prints one JSON line.  Only --gpus 1 is supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "exact-cost A* queries/sec + p50/p99 latency, batched vs one-at-a-time GA*"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def geometric_workload(nq, seed_graph=4, seed_pairs=4):
    import ddp_amd
    hg = ddp_amd.HostGraph.geometric(2000000, 10, 100000, seed_graph)
    src, dst = hg.geo_pairs(nq, 100, 100000, seed_pairs)
    return hg, src, dst


def algorithmic_bytes(st):
    # per expansion: 8 B row offsets + 4 B g(u) + 8 B coordinates of u;
    # per scanned edge: 4 B column + 4 B weight + 4 B g(v) + 8 B coordinates of v;
    # per improvement: 8 B (g, parent) write
    return 20 * st["expansions"] + 20 * st["edges"] + 8 * st["improvements"]


def euclid_h(xy, t):
    """floor(sqrt(dx^2 + dy^2)) for every node, exact integer arithmetic."""
    p = xy.reshape(-1, 2).astype(np.int64)
    d2 = ((p - p[t]) ** 2).sum(1)
    r = np.floor(np.sqrt(d2.astype(np.float64))).astype(np.int64)
    r -= (r * r > d2)
    r += ((r + 1) * (r + 1) <= d2)
    return r.astype(np.int32)


def cpu_baseline(hg, src, dst, costs, budget_s):
    """The oracle's restatement of GA_star/A_star_cpu.cpp (binary heap,
    linear-scan decrease-key) on one host core, over uniformly chosen queries
    until the time budget is spent.  Every cost is checked against the GPU."""
    from oracle import pyoracle
    rng = np.random.default_rng(0)
    order = rng.permutation(len(src))
    done, el = 0, 0.0
    for i in order:
        h = euclid_h(hg.xy, int(dst[i]))
        t0 = time.perf_counter()
        found, cost, _, _ = pyoracle.astar_ref(hg.off, hg.col, hg.w, h, int(src[i]), int(dst[i]), quirk=False)
        el += time.perf_counter() - t0
        if not found or cost != int(costs[i]):
            raise SystemExit(f"cpu baseline disagrees on query {i}: {cost} vs {costs[i]}")
        done += 1
        if el >= budget_s:
            break
    return {"value": round(done / el, 3), "unit": "queries/s", "cores": 1, "kind": "port",
            "sample": f"{done} of {len(src)} queries drawn uniformly, oracle restatement of A_star_cpu.cpp "
                      f"(linear-scan decrease-key), h = floor(Euclid) array per target, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=4096)
    ap.add_argument("--block", type=int, default=0, help="lanes per query workgroup (0 = autotune)")
    ap.add_argument("--single-sample", type=int, default=256, help="queries answered one at a time")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); every box has a single MI355X "
              f"and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    import ddp_amd
    t0 = time.time()
    hg, src, dst = geometric_workload(args.queries)
    log(f"workload ready in {time.time() - t0:.1f}s: N={hg.N} E={hg.E}")
    ctx = ddp_amd.Context(0)
    dg = ddp_amd.DeviceGraph(ctx, hg)

    block = args.block
    if block == 0:
        from ddp_amd import autotune
        block = autotune.pick_batch_block(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID)
    b = ddp_amd.Batch(ctx, dg, len(src), heur=ddp_amd.H_EUCLID, block=block, path_cap=4096)
    b.upload(src, dst)
    for _ in range(args.warmup):
        b.launch()
        b.wait()
    ctx.synchronize()
    kms, stats = [], None
    t1 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
        stats = b.wait()
        kms.append(stats["kernel_ms"])
    ctx.synchronize()
    elapsed = time.perf_counter() - t1
    out = b.download(paths=True)
    if stats["errors"]:
        raise SystemExit(f"{stats['errors']} batched queries reported errors")
    costs = out["cost"]
    lat = out["latency_ms"]
    info = b.info()
    b.close()

    # the same queries, one at a time through the single-query GA* kernel
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(len(src), size=min(args.single_sample, len(src)), replace=False))
    single_lat = []
    ts = time.perf_counter()
    for i in sample:
        q0 = time.perf_counter()
        r = ddp_amd.astar(ctx, dg, int(src[i]), int(dst[i]), heur=ddp_amd.H_EUCLID, K=4096)
        single_lat.append((time.perf_counter() - q0) * 1e3)
        if r.cost != costs[i]:
            raise SystemExit(f"single-query GA* disagrees with the batch on query {i}: {r.cost} vs {costs[i]}")
    single_qps = len(sample) / (time.perf_counter() - ts)

    value = len(src) * args.steps / elapsed
    kmean = float(np.mean(kms))
    abytes = algorithmic_bytes(stats)
    achieved = abytes / (kmean * 1e-3) / 1e9
    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "C4 batched A*: random geometric graph N=2,000,000 (seed 4, lattice 1e5, "
                               "mean degree ~10), 4,096 queries with log-uniform separation, Euclidean "
                               "heuristic, one persistent launch per step",
                   "N": hg.N, "E": hg.E, "queries": len(src), "block": info["block"], "slots": info["slots"],
                   "path_cap": 4096},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "one_at_a_time": {"queries_per_s": round(single_qps, 1), "sample": int(len(sample)),
                          "p50_ms": round(float(np.percentile(single_lat, 50)), 3),
                          "p99_ms": round(float(np.percentile(single_lat, 99)), 3), "K": 4096},
        "speedup_vs_one_at_a_time": round(value / single_qps, 2),
        "expansions_per_s": round(stats["expansions"] / (kmean * 1e-3), 1),
        "edges_per_s": round(stats["edges"] / (kmean * 1e-3), 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "batch_kernel", "kernel_ms": round(kmean, 3),
                     "algorithmic_bytes_per_launch": int(abytes)},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(hg, src, dst, costs, args.cpu_budget)
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
