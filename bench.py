"""Benchmark: exact-cost A* queries/s, batched vs one-at-a-time GA*, on one MI355X.

Default workload (--workload c4): 4,096 source/target queries on a random
geometric graph of 2,000,000 nodes (integer lattice [0,100000]^2, seed 4, mean
degree ~10, edges both ways, w = ceil(Euclidean distance), h = floor(Euclidean
distance to the target)), query separations log-uniform in [100, 100000]
lattice units.  One step = one batched launch answering all 4,096 queries with
their paths.  `value` = queries/s of the batched launches (graph and queries
already resident in HBM); the same queries are also answered one at a time by
the single-query GA* kernel, and served open-loop with update batches
interleaved, and those rates and latencies are reported beside it.

Other workloads (one JSON line each): c1 single GA* query on a random graph,
c2 grid maze with a K sweep, c3 DA* update batches vs full recompute, c5
served R-MAT queries at offered loads.  Only --gpus 1 is supported.
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
ENGINE_NAMES = {0: "piles", 1: "heaps", 2: "piles", 3: "bidi"}
MODE_NAMES = {0: "auto", 1: "grid", 2: "workgroup", 3: "piles", 4: "bidi"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(x, q):
    return round(float(np.percentile(np.asarray(x), q)), 3)


def base_record(args, value, unit, ms_per_step, workload, higher=True, dtype="int32"):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": higher,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": {"workload": workload}}


def roofline(abytes, kernel_ms, kernel, copy_gbps=None):
    achieved = abytes / (kernel_ms * 1e-3) / 1e9
    r = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": kernel,
         "kernel_ms": round(kernel_ms, 3), "algorithmic_bytes_per_launch": int(abytes)}
    if copy_gbps:
        r["copy_gbps"] = round(copy_gbps, 1)
        r["frac_of_copy"] = round(achieved / copy_gbps, 5)
    return r


def measured_traffic(roof, **key):
    """HBM bytes per launch from the committed counter summary of this exact
    configuration (profiles/<kernel>_<workload>_traffic.json, written from two
    rocprofv3 --pmc runs of bench.py); left null for any other configuration."""
    path = os.path.join(ROOT, "profiles", f"{roof['kernel']}_{key['workload']}_traffic.json")
    if not os.path.exists(path):
        return roof
    with open(path) as f:
        t = json.load(f)
    if all(t.get(k) == v for k, v in key.items()):
        roof["traffic"] = t["hbm_bytes_per_launch"]
        roof["traffic_over_algorithmic"] = round(t["hbm_bytes_per_launch"] / roof["algorithmic_bytes_per_launch"], 3)
        roof["traffic_source"] = os.path.relpath(path, ROOT)
    return roof


def copy_bandwidth(ctx):
    """Measured device copy bandwidth (copy_kernel, 2 GiB; also the known-traffic
    kernel the HBM counters in profiles/ are scaled by)."""
    from ctypes import byref, c_float

    import ddp_amd
    ms, gb = c_float(), c_float()
    if ddp_amd._capi.lib().ddp_copy_bandwidth(ctx.handle, 1 << 31, 3, byref(ms), byref(gb)) != 0:
        return None
    return gb.value


# --------------------------------------------------------------------- C4

def algorithmic_bytes_geo(st):
    # per expansion: 8 B row offsets + 4 B g(u) + 8 B coordinates of u;
    # per scanned edge: 4 B column + 4 B weight + 4 B g(v) + 8 B coordinates of v;
    # per improvement: 8 B (g, parent) write
    return 20 * st["expansions"] + 20 * st["edges"] + 8 * st["improvements"]


def cpu_baseline_geo(hg, src, dst, costs, budget_s, threads):
    """The oracle's restatement of GA_star/A_star_cpu.cpp (binary heap,
    linear-scan decrease-key, N-sized state re-initialised per query) on a
    pool of `threads` host threads, one query per thread at a time, h =
    floor(Euclid) to the query's target; chunks of uniformly drawn queries
    until the time budget is spent.  Every cost is checked against the GPU."""
    from oracle import pyoracle
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(src))
    chunk = max(threads * 4, 16)
    done, el, pos = 0, 0.0, 0
    while el < budget_s and pos < len(perm):
        idx = perm[pos:pos + chunk]
        pos += chunk
        t0 = time.perf_counter()
        got, _, started = pyoracle.astar_batch_euclid(hg.off, hg.col, hg.w, hg.xy, src[idx], dst[idx], threads)
        el += time.perf_counter() - t0
        if not np.array_equal(got, costs[idx]):
            bad = int(np.flatnonzero(got != costs[idx])[0])
            raise SystemExit(f"cpu baseline disagrees on query {idx[bad]}: {got[bad]} vs {costs[idx[bad]]}")
        done += len(idx)
    return {"value": round(done / el, 3), "unit": "queries/s", "cores": int(started), "kind": "port",
            "sample": f"{done} of {len(src)} queries drawn uniformly, oracle restatement of A_star_cpu.cpp "
                      f"(linear-scan decrease-key) on {started} threads, h = floor(Euclid) to each target, "
                      f"{el:.1f} s"}


def geometric_updates(hg, n, per, seed):
    """Update batches that keep the Euclidean heuristic admissible: each insert
    u->v weighs at least ceil(dist(u, v)); deletes take existing edges."""
    rng = np.random.default_rng(seed)
    p = hg.xy.reshape(-1, 2).astype(np.int64)
    o = np.append(hg.off, hg.E)
    out = []
    for _ in range(n):
        lines = []
        for u in rng.integers(0, hg.N, per):
            if o[u + 1] > o[u]:
                lines.append((0, int(u), int(hg.col[rng.integers(o[u], o[u + 1])]), 0))
        for u, v in zip(rng.integers(0, hg.N, per), rng.integers(0, hg.N, per)):
            d2 = int(((p[u] - p[v]) ** 2).sum())
            r = int(np.floor(np.sqrt(float(d2))))
            while r * r > d2:
                r -= 1
            while (r + 1) * (r + 1) <= d2:
                r += 1
            w = r if r * r == d2 else r + 1
            lines.append((1, int(u), int(v), max(1, w) + int(rng.integers(0, 1000))))
        out.append(lines)
    return out


def serve_trace(ctx, dg, heur, src, dst, qps, nreq, upd_batches, seed, heap_cap=0, block=0, engine=0, grid_k=0):
    """Open-loop replay: Poisson arrivals at `qps` (all at once when qps <= 0),
    95% point queries and 5% update batches.  Returns (results, types, query
    index per request, secs, server info with the server's counters)."""
    import ddp_amd
    rng = np.random.default_rng(seed)
    types = (rng.random(nreq) < 0.05).astype(np.int32)
    types[0] = 0
    types[np.cumsum(types) > len(upd_batches)] = 0
    qi = rng.integers(0, len(src), nreq)
    ups, k = [], 0
    for ty in types:
        ups.append(upd_batches[k] if ty else None)
        k += int(ty)
    arrival = np.zeros(nreq) if qps <= 0 else np.cumsum(rng.exponential(1000.0 / qps, nreq))
    sv = ddp_amd.Server(ctx, dg, heur=heur, heap_cap=heap_cap, block=block, engine=engine, grid_k=grid_k)
    t0 = time.perf_counter()
    out = sv.replay(arrival, types, src[qi], dst[qi], ups)
    el = time.perf_counter() - t0
    info = sv.info()
    info.update(sv.stats())
    sv.close()
    return out, types, qi, el, info


def served_summary(out, types, qi, el, info, qps, costs=None):
    q = [r for r, ty in zip(out, types) if ty == 0]
    bad = [r for r in out if r["status"] != 0]
    if bad:
        raise SystemExit(f"{len(bad)} served requests failed, first: {bad[0]}")
    if costs is not None:  # requests before the first update saw the unmodified graph
        first = int(np.argmax(types)) if types.any() else len(types)
        for i in range(first):
            if out[i]["cost"] != costs[qi[i]]:
                raise SystemExit(f"served query {i} disagrees with the batch: {out[i]['cost']} vs {costs[qi[i]]}")
    lat = [r["done_ms"] - r["submit_ms"] for r in q]
    summ = {"offered_qps": qps, "requests": len(out), "updates": int(types.sum()), "slots": info["slots"],
            "queries_per_s": round(len(q) / el, 1), "p50_ms": pct(lat, 50), "p90_ms": pct(lat, 90),
            "p99_ms": pct(lat, 99),
            "device_p50_ms": pct([r["device_ms"] for r in q], 50)}
    if info.get("grid_queries"):
        summ["grid_queries"] = info["grid_queries"]
    return summ


def run_c4(args):
    import ddp_amd
    t0 = time.time()
    hg = ddp_amd.HostGraph.geometric(2000000, 10, 100000, 4)
    src, dst = hg.geo_pairs(args.queries, 100, 100000, 4)
    log(f"workload ready in {time.time() - t0:.1f}s: N={hg.N} E={hg.E}")
    ctx = ddp_amd.Context(0)
    # transposed rows too: the bidirectional engine walks them backward from t
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=not args.no_reorder)
    block, engine, heap_cap, lds_cap = args.block, args.engine, 0, 0
    if block == 0:
        from ddp_amd import autotune
        tuned = autotune.tune_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, sample=len(src))
        block, engine = tuned["block"], engine or tuned["engine"]
        heap_cap, lds_cap = tuned["heap_cap"], tuned["lds_cap"]
        log(f"tuned: {tuned}")
    b = ddp_amd.Batch(ctx, dg, len(src), heur=ddp_amd.H_EUCLID, block=block, path_cap=4096, engine=engine,
                      heap_cap=heap_cap, lds_cap=lds_cap)
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
    costs, lat = out["cost"], out["latency_ms"]
    bad = [i for i in range(len(src)) if costs[i] != ddp_amd.INF_COST and
           (len(out["paths"][i]) == 0 or out["paths"][i][0] != src[i] or out["paths"][i][-1] != dst[i])]
    if bad:
        raise SystemExit(f"{len(bad)} batched paths do not run from s to t (first: query {bad[0]})")
    info = b.info()
    b.close()
    log(f"batched: {args.steps} steps in {elapsed:.3f}s")
    # every engine on the same queries: identical costs, their kernel times side by side
    engines = {}
    for name, eng in () if args.no_engines else (("piles", ddp_amd.ENGINE_PILES), ("bidi", ddp_amd.ENGINE_BIDI),
                                                  ("heaps", ddp_amd.ENGINE_HEAPS)):
        blk = 512 if eng == ddp_amd.ENGINE_HEAPS else 1024
        c_e, _, s_e = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=blk, engine=eng)
        if not np.array_equal(c_e, costs):
            raise SystemExit(f"engine {name} disagrees with the timed batch")
        engines[name] = {"block": blk, "kernel_ms": round(s_e["kernel_ms"], 3),
                         "queries_per_s": round(len(src) / (s_e["kernel_ms"] * 1e-3), 1),
                         "expansions": s_e["expansions"], "edges": s_e["edges"]}
    log(f"engines: {engines}")

    # the same queries, one at a time through the single-query GA* kernels,
    # in the mode and shape the single-query tuner picks on 16 of them
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(len(src), size=min(args.single_sample, len(src)), replace=False))
    single_lat = []
    one = {}
    if len(sample):
        from ddp_amd import autotune
        one = autotune.tune_astar(ctx, dg, src[sample], dst[sample], heur=ddp_amd.H_EUCLID, sample=16)
        log(f"single-query tuner: {one}")
    ts = time.perf_counter()
    for i in sample:
        q0 = time.perf_counter()
        r = ddp_amd.astar(ctx, dg, int(src[i]), int(dst[i]), heur=ddp_amd.H_EUCLID, K=one["K"], block=one["block"],
                          heap_cap=one["heap_cap"], lds_cap=one["lds_cap"], mode=one["mode"])
        single_lat.append((time.perf_counter() - q0) * 1e3)
        if r.cost != costs[i]:
            raise SystemExit(f"single-query GA* disagrees with the batch on query {i}: {r.cost} vs {costs[i]}")
    single_qps = len(sample) / max(time.perf_counter() - ts, 1e-9)
    log(f"one at a time: {len(sample)} queries, {single_qps:.1f} q/s")

    value = len(src) * args.steps / elapsed
    kmean = float(np.mean(kms))
    res = base_record(args, round(value, 1), "queries/s", round(elapsed * 1e3 / args.steps, 3),
                      "C4 batched A*: random geometric graph N=2,000,000 (seed 4, lattice 1e5, mean degree ~10), "
                      "4,096 queries with log-uniform separation, Euclidean heuristic, one persistent launch "
                      "per step")
    res["config"].update({"N": hg.N, "E": hg.E, "queries": len(src), "block": info["block"],
                          "slots": info["slots"], "path_cap": 4096, "relabel": not args.no_reorder,
                          "engine": ENGINE_NAMES.get(engine, "piles"), "heap_cap": info["heap_cap"],
                          "lds_cap": info["lds_cap"]})
    res["engines"] = engines
    if len(sample):
        res["one_at_a_time"] = {"queries_per_s": round(single_qps, 1), "sample": int(len(sample)),
                                "p50_ms": pct(single_lat, 50), "p99_ms": pct(single_lat, 99), "K": one["K"],
                                "block": one["block"], "mode": MODE_NAMES.get(one["mode"], one["mode"]),
                                "chosen_by": "ddp_tune_astar on 16 of the sampled queries"}
        res["speedup_vs_one_at_a_time"] = round(value / single_qps, 2)
    res.update({
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
        "expansions_per_s": round(stats["expansions"] / (kmean * 1e-3), 1),
        "edges_per_s": round(stats["edges"] / (kmean * 1e-3), 1),
        "roofline": measured_traffic(roofline(algorithmic_bytes_geo(stats), kmean, "batch_kernel", copy_bandwidth(ctx)),
                                     workload="c4", block=info["block"], relabel=not args.no_reorder,
                                     queries=len(src), engine=res["config"]["engine"]),
        "p50_p99_note": "p50_ms / p99_ms: completion times inside one batched launch of all queries (submitted at "
                        "t = 0); request latencies under load are the served leg's",
    })
    if not args.no_serve:
        ups = geometric_updates(hg, args.serve_requests // 10, 100, 8)
        o, ty, qi, el, sinfo = serve_trace(ctx, dg, ddp_amd.H_EUCLID, src, dst, args.serve_qps, args.serve_requests,
                                           ups, 7, block=args.serve_block)
        res["served"] = served_summary(o, ty, qi, el, sinfo, args.serve_qps, costs)
        log(f"served: {res['served']}")
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_geo(hg, src, dst, costs, args.cpu_budget, args.cpu_threads)
    ctx.close()
    return res


# ----------------------------------------------------------------- C1 / C2

def time_single(ctx, dg, s, t, heur, K, steps, warmup, block=0, mode=0):
    import ddp_amd
    for _ in range(warmup):
        ddp_amd.astar(ctx, dg, s, t, heur=heur, K=K, block=block, mode=mode)
    walls, kms, r = [], [], None
    for _ in range(steps):
        w0 = time.perf_counter()
        r = ddp_amd.astar(ctx, dg, s, t, heur=heur, K=K, block=block, mode=mode)
        walls.append((time.perf_counter() - w0) * 1e3)
        kms.append(r.stats["kernel_ms"])
    return r, float(np.mean(walls)), float(np.mean(kms))


def run_c1(args):
    import ddp_amd
    from oracle import pyoracle
    hg = ddp_amd.HostGraph.random(100000, 6, 1, 100, 1)
    src, dst = hg.reachable_pairs(1, 1)
    s, t = int(src[0]), int(dst[0])
    ctx = ddp_amd.Context(0)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=not args.no_reorder)
    minw = int(hg.w.min())
    h, _ = dg.bfs_hops(t, reverse=True, scale=minw, store_as_h=True)
    r, wall, kms = time_single(ctx, dg, s, t, ddp_amd.H_ARRAY, 1024, args.steps, args.warmup)
    other = {}
    for name, mode, heur, blk in (("grid_block256", ddp_amd.MODE_GRID, ddp_amd.H_ARRAY, 256),
                                  ("grid_block1024", ddp_amd.MODE_GRID, ddp_amd.H_ARRAY, 1024),
                                  ("workgroup_mode", ddp_amd.MODE_WORKGROUP, ddp_amd.H_ARRAY, 0),
                                  ("pile_mode", ddp_amd.MODE_PILES, ddp_amd.H_ARRAY, 0),
                                  ("bidi_mode_h0", ddp_amd.MODE_BIDI, ddp_amd.H_ZERO, 0)):
        ro, wall_o, kms_o = time_single(ctx, dg, s, t, heur, 1024, args.steps, args.warmup, block=blk, mode=mode)
        if ro.cost != r.cost:
            raise SystemExit(f"C1: {name} {ro.cost} != grid mode {r.cost}")
        so = ro.stats
        other[name] = {"wall_ms": round(wall_o, 3), "kernel_ms": round(kms_o, 3), "iterations": so["iterations"],
                       "expansions": so["expansions"], "blocks": so["blocks"], "block": so["block"]}
    # CPU: the oracle's restatement of A_star_cpu.cpp on the same inputs
    reps, el = 0, 0.0
    while el < min(args.cpu_budget, 10.0) or reps == 0:
        c0 = time.perf_counter()
        found, cost, _, _ = pyoracle.astar_ref(hg.off, hg.col, hg.w, h, s, t, quirk=False)
        el += time.perf_counter() - c0
        reps += 1
    if not found or cost != r.cost:
        raise SystemExit(f"C1 cost mismatch: GPU {r.cost} CPU {cost}")
    st = r.stats
    abytes = 16 * st["expansions"] + 16 * st["edges"] + 8 * st["improvements"]
    res = base_record(args, round(1000.0 / wall, 2), "queries/s", round(wall, 3),
                      "C1 single GA* query: random N=100,000, 6 out-edges, w in [1,100], seed 1, h = reverse-BFS "
                      "hops x min weight, K=1024 (grid mode)")
    res["config"].update({"s": s, "t": t, "K": st["K"], "block": st["block"], "cost": r.cost})
    res.update(other)
    res.update({"kernel_ms": round(kms, 3), "iterations": st["iterations"], "expansions": st["expansions"],
                "expansions_per_s": round(st["expansions"] / (kms * 1e-3), 1),
                "edges_per_s": round(st["edges"] / (kms * 1e-3), 1),
                "roofline": roofline(abytes, kms, "ga_grid_kernel"),
                "cpu_baseline": {"value": round(reps / el, 3), "unit": "queries/s", "cores": 1, "kind": "port",
                                 "sample": f"the same query {reps} times, oracle restatement of A_star_cpu.cpp"}})
    ctx.close()
    return res


def run_c2(args):
    import ddp_amd
    from oracle import pyoracle
    W = 4096
    hg = ddp_amd.HostGraph.grid_maze(W, W, 300, 2)
    s, t = 0, W * W - 1
    ctx = ddp_amd.Context(0)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)  # transposed rows for the bidirectional mode
    dg.set_grid(W)
    sweep = {}
    cost = None
    for name, K, mode in (("512-workgroup", 512, ddp_amd.MODE_WORKGROUP), ("1024-workgroup", 1024,
                          ddp_amd.MODE_WORKGROUP), ("1024-piles", 1024, ddp_amd.MODE_PILES),
                          ("256-bidi", 256, ddp_amd.MODE_BIDI), ("512-bidi", 512, ddp_amd.MODE_BIDI),
                          ("1024-bidi", 1024, ddp_amd.MODE_BIDI),
                          ("512-grid", 512, ddp_amd.MODE_GRID),
                          ("2048-grid", 2048, ddp_amd.MODE_GRID), ("8192-grid", 8192, ddp_amd.MODE_GRID)):
        r, wall, kms = time_single(ctx, dg, s, t, ddp_amd.H_MANHATTAN, K, args.steps, args.warmup, mode=mode)
        st = r.stats
        if cost is not None and r.cost != cost:
            raise SystemExit(f"C2 cost differs across K: {r.cost} vs {cost}")
        cost = r.cost
        abytes = 8 * st["expansions"] + 12 * st["edges"] + 8 * st["improvements"]  # closed-form h reads nothing
        kern = "ga_grid_kernel" if mode == ddp_amd.MODE_GRID else "ga_wg_kernel"  # one-workgroup modes
        sweep[name] = {"K": K, "wall_ms": round(wall, 3), "kernel_ms": round(kms, 3), "iterations": st["iterations"],
                       "expansions": st["expansions"], "edges": st["edges"], "block": st["block"],
                       "blocks": st["blocks"], "roofline": roofline(abytes, kms, kern)}
        log(f"C2 {name}: {sweep[name]}")
    best = min(sweep, key=lambda k: sweep[k]["wall_ms"])
    xs = np.arange(W * W, dtype=np.int64)
    h = ((W - 1 - xs // W) + (W - 1 - xs % W)).astype(np.int32)
    c0 = time.perf_counter()
    found, ccost, cexp = pyoracle.astar_indexed(hg.off, hg.col, hg.w, h, s, t)
    cel = time.perf_counter() - c0
    if not found or ccost != cost:
        raise SystemExit(f"C2 cost mismatch: GPU {cost} CPU {ccost}")
    res = base_record(args, round(1000.0 / sweep[best]["wall_ms"], 3), "queries/s", sweep[best]["wall_ms"],
                      "C2 single GA* query: 4096x4096 4-connected grid maze, p_block=0.30, seed 2, carved corridor, "
                      "unit weights, Manhattan heuristic, corner to corner, K and execution mode swept")
    res["config"].update({"N": hg.N, "E": hg.E, "cost": cost, "best": best})
    res.update({"sweep": sweep, "roofline": sweep[best]["roofline"],
                "cpu_baseline": {"value": round(1.0 / cel, 4), "unit": "queries/s", "cores": 1,
                                 "kind": "port", "sample": f"the full query once ({cexp} expansions), indexed-heap "
                                 f"variant of the A_star_cpu.cpp restatement (the linear-scan decrease-key does not "
                                 f"finish on 16.7M nodes)"}})
    ctx.close()
    return res


# --------------------------------------------------------------------- C3

def run_c3(args):
    import ddp_amd
    from oracle import pyoracle
    n_nodes, nb = args.c3_nodes, args.c3_batches
    hg = ddp_amd.HostGraph.random(n_nodes, 8, 1, 1000, 3)
    upd = hg.gen_updates(nb, args.c3_batch // 2, args.c3_batch // 2, 1, 1000, 33).batches
    src, dst = hg.reachable_pairs(1, 3)
    s, t = int(src[0]), int(dst[0])
    # heuristic on the union of every version: admissible for all of them
    o = np.append(hg.off, hg.E)
    us = [np.repeat(np.arange(hg.N), np.diff(o))]
    vs, ws = [hg.col], [hg.w]
    for b in upd:
        a = np.array([l for l in b if l[0] == 1], np.int64)
        us.append(a[:, 1])
        vs.append(a[:, 2])
        ws.append(a[:, 3])
    us, vs, ws = np.concatenate(us), np.concatenate(vs).astype(np.int32), np.concatenate(ws).astype(np.uint32)
    order = np.argsort(us, kind="stable")
    uoff = np.searchsorted(us[order], np.arange(hg.N)).astype(np.int32)
    ctx = ddp_amd.Context(0)
    ug = ddp_amd.DeviceGraph(ctx, ddp_amd.HostGraph(uoff, vs[order], ws[order]), reverse=True)
    h, _ = ug.bfs_hops(t, reverse=True, scale=int(ws.min()))
    ug.close()
    # repair and full recompute, each with the GA* grid kernel (K queues) and
    # with one workgroup's pile engine
    sessions = {
        "repair_grid": (ddp_amd.DAStar.REPAIR, ddp_amd.MODE_AUTO),
        "repair_piles": (ddp_amd.DAStar.REPAIR, ddp_amd.MODE_PILES),
        "recompute_grid": (ddp_amd.DAStar.RECOMPUTE, ddp_amd.MODE_AUTO),
        "recompute_piles": (ddp_amd.DAStar.RECOMPUTE, ddp_amd.MODE_PILES),
    }
    # one session at a time (20 batches each), so that one's device state does
    # not evict another's from the caches
    ms = {k: [] for k in sessions}
    dev = {k: [] for k in sessions}
    costs = {k: [] for k in sessions}
    init = set()
    for k, (mode, m) in sessions.items():
        da = ddp_amd.DAStar(ctx, hg, s, t, h=h, K=args.c3_K, mode=m)
        init.add(da.cost)
        for bi, lines in enumerate(upd):
            c, st = da.apply(lines, mode=mode)
            costs[k].append(c)
            ms[k].append(st["total_ms"])
            dev[k].append(st["repair_ms"] + st["search_ms"])
        da.close()
        log(f"C3 {k}: {float(np.mean(ms[k])):.3f} ms per batch")
    if len(init) != 1 or any(costs[k] != costs["repair_grid"] for k in sessions):
        raise SystemExit(f"C3: sessions disagree ({init}, {costs})")
    per = [{"batch": bi, "cost": costs["repair_grid"][bi], **{k: round(ms[k][bi], 3) for k in sessions}}
           for bi in range(nb)]
    cur = (hg.off, hg.col, hg.w)
    for bi, lines in enumerate(upd):  # the oracle on every 4th version and the last
        cur = pyoracle.apply_batch(*cur, lines)[:3]
        if bi % max(1, nb // 5) == 0 or bi == nb - 1:
            d, _ = pyoracle.dijkstra(*cur, s, t)
            want = int(d[t]) if d[t] >= 0 else ddp_amd.INF_COST
            if costs["repair_grid"][bi] != want:
                raise SystemExit(f"C3 batch {bi}: GPU {costs['repair_grid'][bi]} oracle {want}")
    mean = {k: float(np.mean(v)) for k, v in ms.items()}
    best_rep = min(("repair_grid", "repair_piles"), key=lambda k: mean[k])
    best_rec = min(("recompute_grid", "recompute_piles"), key=lambda k: mean[k])
    res = base_record(args, round(1000.0 / mean[best_rep], 3), "batches/s", round(mean[best_rep], 3),
                      f"C3 DA*: random N={n_nodes:,}, out-degree 8, w in [1,1000], seed 3, {nb} batches of "
                      f"{args.c3_batch // 2:,} inserts + {args.c3_batch // 2:,} deletes, h on the union graph")
    res["config"].update({"N": hg.N, "E": hg.E, "K": args.c3_K, "s": s, "t": t, "repair": best_rep})
    res.update({"ms_per_batch_mean": {k: round(v, 3) for k, v in mean.items()},
                "device_ms_per_batch_mean": {k: round(float(np.mean(v)), 3) for k, v in dev.items()},
                "repair_ms_mean": round(mean[best_rep], 3), "recompute_ms_mean": round(mean[best_rec], 3),
                "best_recompute": best_rec,
                "speedup_vs_recompute": round(mean[best_rec] / mean[best_rep], 2), "batches": per})
    ctx.close()
    return res


# --------------------------------------------------------------------- C5

def rmat_updates(hg, nb, per, seed):
    """nb batches of `per` deletes of existing edges and `per` random inserts
    (w uniform in [1, 1000])."""
    rng = np.random.default_rng(seed)
    o = np.append(hg.off, hg.E).astype(np.int64)
    deg = np.diff(o)
    out = []
    for _ in range(nb):
        u = rng.integers(0, hg.N, per)
        u = u[deg[u] > 0]
        j = o[u] + (rng.random(len(u)) * deg[u]).astype(np.int64)
        dels = [(0, int(a), int(b), 0) for a, b in zip(u, hg.col[j])]
        iu, iv, iw = rng.integers(0, hg.N, per), rng.integers(0, hg.N, per), rng.integers(1, 1001, per)
        out.append(dels + [(1, int(a), int(b), int(c)) for a, b, c in zip(iu, iv, iw)])
    return out


def run_c5(args):
    import ddp_amd
    scale = args.rmat_scale
    t0 = time.time()
    hg = ddp_amd.HostGraph.rmat(scale, 16 * (1 << scale), 0.57, 0.19, 0.19, 1, 1000, 5)
    log(f"R-MAT scale {scale}: N={hg.N} E={hg.E} in {time.time() - t0:.1f}s")
    src, dst = hg.hub_pairs(args.c5_queries, 5)
    ups = rmat_updates(hg, max(1, args.c5_requests // 15), 1000, 9)
    ctx = ddp_amd.Context(0)
    t0 = time.time()
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=not args.no_reorder)
    log(f"uploaded in {time.time() - t0:.1f}s")
    # batched kernel on all queries (unmodified graph), one-way piles and bidirectional
    batched = {}
    cost_b = None
    engines = (("piles", ddp_amd.ENGINE_PILES), ("bidi", ddp_amd.ENGINE_BIDI))
    if args.c5_engine:
        engines = tuple(e for e in engines if e[0] == args.c5_engine)
    for name, eng in engines:
        c, _, bst = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, heap_cap=args.c5_heap_cap, engine=eng)
        if bst["errors"]:
            raise SystemExit(f"C5 batch ({name}): {bst['errors']} queries failed")
        if cost_b is not None and not np.array_equal(c, cost_b):
            raise SystemExit(f"C5 batch: {name} disagrees with piles")
        cost_b = c
        # per expansion: 8 B row offsets + the 8-byte (f, parent) word; per scanned
        # edge: 4 B column + 4 B weight + the target's word; per improvement: the word
        abytes = 16 * bst["expansions"] + 16 * bst["edges"] + 8 * bst["improvements"]
        batched[name] = {"queries_per_s": round(len(src) / (bst["kernel_ms"] * 1e-3), 2),
                         "kernel_ms": round(bst["kernel_ms"], 2), "edges": bst["edges"],
                         "expansions": bst["expansions"], "block": bst["block"], "slots": bst["slots"],
                         "hbm_algorithmic_bytes": abytes,
                         "roofline": measured_traffic(roofline(abytes, bst["kernel_ms"], "batch_kernel"),
                                                      workload="c5", block=bst["block"],
                                                      relabel=not args.no_reorder, queries=len(src), engine=name)}
        log(f"C5 batch {name}: {batched[name]}")
    if args.c5_batch_only:
        ctx.close()
        return {"batched": batched}
    # one at a time: grid mode (a single query is a full search of the giant
    # component; one workgroup takes 0.7-1 s), at the K that is fastest on the
    # sample (default blocks: 256 lanes below K = 65,536, 1,024 from there);
    # every query's cost checked against the batch
    lat1, one_k, by_k = [], 0, {}
    nsingle = min(args.single_sample, 8)
    for K in (8192, 16384, 32768, 65536):
        cap = max(2 * hg.N // K + 256, (8 * hg.N + K - 1) // K)  # 8N queued entries, as the server plans
        lk = []
        for i in range(nsingle):
            q0 = time.perf_counter()
            r = ddp_amd.astar(ctx, dg, int(src[i]), int(dst[i]), heur=ddp_amd.H_ZERO, K=K, heap_cap=cap,
                              mode=ddp_amd.MODE_GRID)
            lk.append((time.perf_counter() - q0) * 1e3)
            if r.cost != cost_b[i]:
                raise SystemExit(f"C5 query {i}: single (K={K}) {r.cost} batch {cost_b[i]}")
        if lk:
            by_k[K] = round(float(np.mean(lk)), 2)
            if not lat1 or np.mean(lk) < np.mean(lat1):
                lat1, one_k = lk, K
    log(f"C5 one at a time, mean ms by K: {by_k}")
    single_qps = 1000.0 / float(np.mean(lat1)) if lat1 else None
    grid_k = args.c5_grid_k or one_k or 32768
    # served: closed loop (every request at once) for the capacity, then open
    # loop at fractions of it for the latency curve
    o_, ty, qi, el, info = serve_trace(ctx, dg, ddp_amd.H_ZERO, src, dst, 0.0, args.c5_requests, ups, 11,
                                       heap_cap=args.c5_heap_cap, engine=args.c5_serve_engine,
                                       block=args.c5_serve_block)
    closed = served_summary(o_, ty, qi, el, info, 0.0, cost_b)
    log(f"C5 served closed loop {closed}")
    cap = closed["queries_per_s"]
    offered = args.c5_qps or [round(f * cap, 1) for f in (0.25, 0.5, 0.75, 0.9)]
    sweep = [closed]
    for qps in offered:
        o_, ty, qi, el, info = serve_trace(ctx, dg, ddp_amd.H_ZERO, src, dst, qps, args.c5_sweep_requests, ups, 12,
                                           heap_cap=args.c5_heap_cap, engine=args.c5_serve_engine,
                                           block=args.c5_serve_block)
        sweep.append(served_summary(o_, ty, qi, el, info, qps, cost_b))
        log(f"C5 served {sweep[-1]}")
    under = [x for x in sweep if 0 < x["offered_qps"] < cap] or sweep
    lat_at = max(under, key=lambda x: x["offered_qps"])
    # low load: the same open-loop trace with and without grid mode for
    # queries that arrive while nothing is in flight
    low = []
    for f in args.c5_low_load:
        qps = round(f * cap, 1)
        pair = {"offered_qps": qps}
        for name, gk in (("slots", 0), ("grid", grid_k)):
            o_, ty, qi, el, info = serve_trace(ctx, dg, ddp_amd.H_ZERO, src, dst, qps, args.c5_low_requests, ups,
                                               13, heap_cap=args.c5_heap_cap, engine=args.c5_serve_engine,
                                               block=args.c5_serve_block, grid_k=gk)
            pair[name] = served_summary(o_, ty, qi, el, info, qps, cost_b)
        pair["grid_k"] = grid_k
        low.append(pair)
        log(f"C5 low load {pair}")
    best_batched = max(v["queries_per_s"] for v in batched.values())
    res = base_record(args, cap, "queries/s", None,
                      f"C5 online serving: R-MAT scale {scale} ({hg.N:,} nodes, {hg.E:,} edges, a/b/c/d = "
                      f"0.57/0.19/0.19/0.05, seed 5), w in [1,1000], h = 0, {args.c5_queries} connected pairs, "
                      f"95% queries, 5% update batches of 1,000 deletes + 1,000 inserts; value = closed-loop "
                      f"served capacity")
    res["config"].update({"N": hg.N, "E": hg.E, "relabel": not args.no_reorder, "slots": info["slots"],
                          "serve_block": info["block"],
                          "serve_engine": ENGINE_NAMES.get(args.c5_serve_engine, "piles")})
    res.update({"p50_ms": lat_at["p50_ms"], "p99_ms": lat_at["p99_ms"], "latency_at_qps": lat_at["offered_qps"],
                "sweep": sweep, "low_load": low, "batched": batched, "batched_queries_per_s": best_batched,
                "one_at_a_time": ({"queries_per_s": round(single_qps, 3), "p50_ms": pct(lat1, 50),
                                   "p99_ms": pct(lat1, 99), "sample": len(lat1), "K": one_k, "mode": "grid",
                                   "mean_ms_by_K": by_k} if lat1 else None),
                "batched_over_one_at_a_time": round(best_batched / single_qps, 2) if single_qps else None,
                "served_over_one_at_a_time": round(cap / single_qps, 2) if single_qps else None})
    ctx.close()
    return res


def summary_c3(r):
    """The C3 line reduced to what the headline line carries."""
    keep = ("value", "unit", "repair_ms_mean", "recompute_ms_mean", "best_recompute", "ms_per_batch_mean",
            "device_ms_per_batch_mean", "speedup_vs_recompute")
    out = {k: r[k] for k in keep}
    out["config"] = r["config"]
    out["batches"] = len(r["batches"])
    out["checked"] = "repair == recompute on every batch; oracle Dijkstra on the merged graph every 4th and the last"
    return out


def summary_c5(r):
    keep = ("value", "unit", "p50_ms", "p99_ms", "latency_at_qps", "sweep", "low_load", "batched",
            "batched_queries_per_s", "one_at_a_time", "batched_over_one_at_a_time", "served_over_one_at_a_time")
    out = {k: r[k] for k in keep}
    out["config"] = r["config"]
    return out


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical CPUs visible"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--queries", type=int, default=4096)
    ap.add_argument("--block", type=int, default=0, help="lanes per query workgroup (0 = autotune)")
    ap.add_argument("--engine", type=int, default=0, help="batched queue engine: 0 auto, 1 heaps, 2 piles")
    ap.add_argument("--single-sample", type=int, default=256, help="queries answered one at a time")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the batched CPU baseline (the GPU box's share is 16)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-engines", action="store_true", help="c4: skip the side-by-side run of every engine")
    ap.add_argument("--no-reorder", action="store_true",
                    help="keep the input node numbering on the device (default: Hilbert/BFS relabelling)")
    ap.add_argument("--serve-qps", type=float, default=5000.0,
                    help="offered load of the served leg (0 = closed loop: every request at once)")
    ap.add_argument("--serve-requests", type=int, default=8192)
    ap.add_argument("--serve-block", type=int, default=0, help="lanes per serving slot (0 = library default)")
    ap.add_argument("--c3-nodes", type=int, default=1000000)
    ap.add_argument("--c3-batches", type=int, default=20)
    ap.add_argument("--c3-batch", type=int, default=20000, help="lines per batch (half inserts, half deletes)")
    ap.add_argument("--c3-K", type=int, default=16384, help="queues of both DA* sessions (repair and recompute)")
    ap.add_argument("--rmat-scale", type=int, default=23)
    ap.add_argument("--c5-queries", type=int, default=256)
    ap.add_argument("--c5-requests", type=int, default=1024)
    ap.add_argument("--c5-heap-cap", type=int, default=0, help="queue entries per lane (0 = library default)")
    ap.add_argument("--c5-qps", type=float, nargs="+", default=None,
                    help="offered loads of the open-loop served points (default: 25/50/75/90%% of the "
                         "closed-loop capacity)")
    ap.add_argument("--c5-sweep-requests", type=int, default=256, help="requests per open-loop point")
    ap.add_argument("--c5-batch-only", action="store_true", help="c5: the batched launches only (for profiling)")
    ap.add_argument("--c5-engine", default="", help="c5: batch engine to run (piles or bidi; default both)")
    ap.add_argument("--c5-serve-engine", type=int, default=0, help="c5: serving engine (0 = library default)")
    ap.add_argument("--c5-serve-block", type=int, default=0, help="c5: lanes per serving slot (0 = library default)")
    ap.add_argument("--c5-low-load", type=float, nargs="*", default=[0.025, 0.05],
                    help="c5: fractions of the closed-loop capacity for the low-load points, each served with "
                         "and without grid mode")
    ap.add_argument("--c5-low-requests", type=int, default=128,
                    help="c5: requests per low-load point (the router settles after its first 20-30)")
    ap.add_argument("--c5-grid-k", type=int, default=0,
                    help="c5: grid-mode lanes of the low-load points (0 = the one-at-a-time leg's fastest K)")
    ap.add_argument("--no-extra", action="store_true",
                    help="c4 only: skip the C3 and C5 legs that the default line carries as extra keys")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); every box has a single MI355X "
              f"and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    res = {"c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4, "c5": run_c5}[args.workload](args)
    if args.workload == "c4" and not args.no_extra:
        # the other served/dynamic configurations ride along as extra keys
        res["c3"] = summary_c3(run_c3(args))
        res["c5"] = summary_c5(run_c5(args))
    if isinstance(res.get("cpu_baseline"), dict):
        res["cpu_baseline"]["cpu"] = cpu_model()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
