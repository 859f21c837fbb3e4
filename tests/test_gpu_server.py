"""GPU parity of the continuous-batching server: every query's cost equals a
from-scratch shortest path on the graph version it was bound to (its
submission position among the update batches), whatever else was running."""
import sys
import time

import numpy as np
import pytest

import ddp_amd
from oracle import pyoracle

from helpers import geometric_updates

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def versions(hg, batches):
    out = [(hg.off, hg.col, hg.w)]
    for b in batches:
        out.append(pyoracle.apply_batch(*out[-1], b)[:3])
    return out


def make_trace(hg, n, upd_every, seed):
    rng = np.random.default_rng(seed)
    src, dst = hg.reachable_pairs(n, seed)
    types, ups = [], []
    ub = hg.gen_updates(n // upd_every + 1, 40, 40, 1, 100, seed + 1).batches
    k = 0
    for i in range(n):
        if i % upd_every == upd_every - 1:
            types.append(1)
            ups.append(ub[k])
            k += 1
        else:
            types.append(0)
            ups.append(None)
    arrival = np.cumsum(rng.exponential(0.05, n))  # ms
    return arrival, np.array(types, np.int32), src, dst, ups


def check(results, types, src, dst, ups, hg):
    batches = [u for u in ups if u is not None]
    vers = versions(hg, batches)
    cache = {}
    v = 0
    for i, r in enumerate(results):
        assert r["tag"] == i and r["status"] == 0, r
        if types[i] == 1:
            v += 1
            assert r["version"] == v
            continue
        assert r["version"] == v
        key = (v, int(src[i]))
        if key not in cache:
            cache[key], _ = pyoracle.dijkstra(*vers[v], int(src[i]))
        d = cache[key][dst[i]]
        assert r["cost"] == (int(d) if d >= 0 else ddp_amd.INF_COST), (i, r, d)


def test_server_versions_and_costs(ctx):
    hg = ddp_amd.HostGraph.random(6000, 5, 1, 100, 41)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    arrival, types, src, dst, ups = make_trace(hg, 400, 20, 42)
    for block, engine in ((64, ddp_amd.ENGINE_HEAPS), (256, ddp_amd.ENGINE_HEAPS), (64, ddp_amd.ENGINE_PILES),
                          (256, ddp_amd.ENGINE_PILES), (1024, ddp_amd.ENGINE_PILES)):
        say("server block", block, "engine", engine)
        sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, block=block, slots=48, engine=engine)
        res = sv.replay(arrival, types, src, dst, ups)
        say("  info", sv.info(), "p50 device ms", np.median([r["device_ms"] for r in res]))
        check(res, types, src, dst, ups, hg)
        sv.close()


def test_server_matches_batch_on_each_version(ctx):
    """Same answer alone (batched kernel on a graph holding only that version)
    and in the mix (server with updates interleaved)."""
    hg = ddp_amd.HostGraph.geometric(20000, 10, 100000, 43)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    ub = geometric_updates(hg, 3, 200, 44)  # inserts keep the Euclidean heuristic admissible
    src, dst = hg.geo_pairs(60, 500, 100000, 45)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_EUCLID, slots=16)
    tags = 0
    plan = []  # (version, query index)
    for v in range(4):
        if v > 0:
            assert sv.submit_update(ub[v - 1], 10000 + v) == 0
        for i in range(v * 15, v * 15 + 15):
            assert sv.submit_query(int(src[i]), int(dst[i]), i) == 0
            plan.append((v, i))
    sv.drain()
    got = {r["tag"]: r for r in sv.poll(1000)}
    sv.close()
    vers = versions(hg, ub)
    for v in range(4):
        g = ddp_amd.HostGraph(*vers[v], xy=hg.xy)
        dgv = ddp_amd.DeviceGraph(ctx, g)
        idx = [i for (vv, i) in plan if vv == v]
        cost, _, _ = ddp_amd.astar_batch(ctx, dgv, src[idx], dst[idx], heur=ddp_amd.H_EUCLID)
        for k, i in enumerate(idx):
            assert got[i]["version"] == v and got[i]["cost"] == cost[k], (v, i)
        dgv.close()


def test_server_idle_exit_and_relaunch(ctx):
    hg = ddp_amd.HostGraph.random(3000, 4, 1, 50, 46)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.reachable_pairs(4, 47)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, slots=8, idle_ms=30)
    for k in range(2):
        assert sv.submit_query(int(src[2 * k]), int(dst[2 * k]), 2 * k) == 0
        assert sv.submit_query(int(src[2 * k + 1]), int(dst[2 * k + 1]), 2 * k + 1) == 0
        sv.drain(60000)
        res = sv.poll()
        assert len(res) == 2
        for r in res:
            d, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[r["tag"]]))
            assert r["cost"] == d[dst[r["tag"]]]
        time.sleep(0.2)  # every workgroup exits after 30 ms idle; the next submit relaunches
    sv.close()


def test_server_ring_backpressure(ctx):
    hg = ddp_amd.HostGraph.random(3000, 4, 1, 50, 48)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.reachable_pairs(300, 49)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, slots=4, ring=64)
    done = {}
    i = 0
    refused = 0
    while i < 300:
        rc = sv.submit_query(int(src[i]), int(dst[i]), i)
        if rc == 0:
            i += 1
        else:
            assert rc == 3  # DDP_ERR_NOMEM: ring full
            refused += 1
            for r in sv.poll():
                done[r["tag"]] = r
    sv.drain()
    for r in sv.poll(1000):
        done[r["tag"]] = r
    assert len(done) == 300 and refused > 0
    for k in range(0, 300, 30):
        d, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[k]))
        assert done[k]["cost"] == d[dst[k]]
    sv.close()


def test_server_config4_full_size(ctx):
    """The C4 workload through the continuous-batching server at full size:
    4,096 queries on the 2,000,000-node geometric graph submitted at once,
    with one update batch (100 deletes + 100 inserts that keep the Euclidean
    heuristic admissible) in the middle.  Queries before the update return
    the batch kernel's costs on the original graph, queries after it the
    batch kernel's costs on the updated graph: the same answer alone and in
    the mix.  Three queries on each side of the update also equal the
    oracle's Dijkstra on their own version of the graph."""
    hg = ddp_amd.HostGraph.geometric(2000000, 10, 100000, 4)
    src, dst = hg.geo_pairs(4096, 100, 100000, 4)
    ub = geometric_updates(hg, 1, 100, 8)
    mid = len(src) // 2
    n = len(src) + 1
    types = np.zeros(n, np.int32)
    types[mid] = 1
    qs, qt = np.insert(src, mid, 0), np.insert(dst, mid, 0)
    ups = [None] * n
    ups[mid] = ub[0]
    dg = ddp_amd.DeviceGraph(ctx, hg, reorder=True)
    say("c4 served: batch before")
    before, _, st = ddp_amd.astar_batch(ctx, dg, src[:mid], dst[:mid], heur=ddp_amd.H_EUCLID, block=1024,
                                        engine=ddp_amd.ENGINE_PILES)
    assert st["errors"] == 0
    say("c4 served: server")
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_EUCLID)
    res = sv.replay(np.zeros(n), types, qs, qt, ups)
    say("  info", sv.info())
    sv.close()
    dg.close()
    say("c4 served: batch after")
    g1 = ddp_amd.HostGraph(*pyoracle.apply_batch(hg.off, hg.col, hg.w, ub[0])[:3], xy=hg.xy)
    dg1 = ddp_amd.DeviceGraph(ctx, g1, reorder=True)
    after, _, st = ddp_amd.astar_batch(ctx, dg1, src[mid:], dst[mid:], heur=ddp_amd.H_EUCLID, block=1024,
                                       engine=ddp_amd.ENGINE_PILES)
    assert st["errors"] == 0
    dg1.close()
    assert [r["tag"] for r in res] == list(range(n)) and all(r["status"] == 0 for r in res)
    got = np.array([r["cost"] for r in res])
    ver = np.array([r["version"] for r in res])
    assert (ver[:mid] == 0).all() and (ver[mid:] == 1).all()
    assert np.array_equal(got[:mid], before)
    assert np.array_equal(got[mid + 1:], after)
    g1_rows = (g1.off, g1.col, g1.w)
    for i in (0, mid // 2, mid - 1, mid + 1, mid + len(src) // 4, n - 1):
        rows = (hg.off, hg.col, hg.w) if i < mid else g1_rows
        dist, _ = pyoracle.dijkstra(*rows, int(qs[i]), int(qt[i]))
        want = int(dist[qt[i]]) if dist[qt[i]] >= 0 else ddp_amd.INF_COST
        say("c4 served: request", i, "version", ver[i], got[i], want)
        assert got[i] == want


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("grid_k,qps", [(0, 150.0), (32768, 10.0)], ids=["slots", "grid_low_load"])
def test_server_config5_served(ctx, grid_k, qps):
    """C5's served workload on its own graph: R-MAT scale 23 (8,388,608
    nodes, 134,217,728 edges, a/b/c/d = 0.57/0.19/0.19/0.05, seed 5),
    w in [1,1000], h = 0.  Poisson arrivals of point queries between connected
    pairs, interleaved with 3 update batches of 1,000 deletes + 1,000 inserts.
    Every request succeeds and is bound to the version of its submission
    position; on every version, sampled queries equal the oracle's Dijkstra on
    the oracle-applied graph, and every served cost equals the batch kernel's
    on that version's graph, whose paths are checked edge by edge.  Served
    at 150 requests/s in slots, and at 10/s with grid mode for low-load
    queries (most of them then run with the whole GPU, updates by their own
    kernel)."""
    from helpers import path_cost, rmat_updates
    hg = ddp_amd.HostGraph.rmat(23, 16 * (1 << 23), 0.57, 0.19, 0.19, 1, 1000, 5)
    say("c5 served: graph", hg.N, hg.E)
    src, dst = hg.hub_pairs(96, 5)
    ub = rmat_updates(hg, 3, 1000, 9)
    n = len(src) + len(ub)
    pos = [24, 48, 72]  # request index of each update batch
    types = np.zeros(n, np.int32)
    types[pos] = 1
    qidx = np.full(n, -1)
    qidx[types == 0] = np.arange(len(src))
    qs = np.where(qidx >= 0, src[np.maximum(qidx, 0)], 0).astype(np.int32)
    qt = np.where(qidx >= 0, dst[np.maximum(qidx, 0)], 0).astype(np.int32)
    ups = [None] * n
    for k, p in enumerate(pos):
        ups[p] = ub[k]
    rng = np.random.default_rng(55)
    arrival = np.cumsum(rng.exponential(1000.0 / qps, n))
    dg = ddp_amd.DeviceGraph(ctx, hg, reorder=True)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, grid_k=grid_k)
    say("c5 served: replay", sv.info(), "grid_k", grid_k, "offered/s", qps)
    res = sv.replay(arrival, types, qs, qt, ups)
    st = sv.stats()
    say("c5 served: stats", st)
    sv.close()
    if grid_k:
        assert st["grid_queries"] > len(src) // 2, st
    dg.close()
    assert [r["tag"] for r in res] == list(range(n)) and all(r["status"] == 0 for r in res), \
        [r for r in res if r["status"]][:3]
    ver = np.cumsum(types)
    assert [r["version"] for r in res] == ver.tolist()
    got = np.array([r["cost"] for r in res])
    lat = np.array([r["done_ms"] - r["submit_ms"] for r in res if r["type"] == 0])
    say("c5 served: latency p50 / p99 ms", np.percentile(lat, 50), np.percentile(lat, 99))
    cur = (hg.off, hg.col, hg.w)
    for v in range(len(ub) + 1):
        if v:
            cur = pyoracle.apply_batch(*cur, ub[v - 1])[:3]
        mine = np.flatnonzero((ver == v) & (types == 0))
        # the batch kernel on this version's graph, with paths
        gv = ddp_amd.HostGraph(*cur)
        dgv = ddp_amd.DeviceGraph(ctx, gv, reorder=True)
        costs, paths, st = ddp_amd.astar_batch(ctx, dgv, qs[mine], qt[mine], heur=ddp_amd.H_ZERO, path_cap=4096)
        dgv.close()
        assert st["errors"] == 0
        assert np.array_equal(got[mine], costs), v
        for k in range(len(mine)):
            p = paths[k]
            assert len(p) and p[0] == qs[mine[k]] and p[-1] == qt[mine[k]] and path_cost(*cur, p) == costs[k]
        for i in mine[:2]:  # sampled against the oracle's Dijkstra
            d, _ = pyoracle.dijkstra(*cur, int(qs[i]), int(qt[i]))
            say("c5 served: version", v, "request", i, "cost", got[i], "oracle", d[qt[i]])
            assert got[i] == d[qt[i]]


def test_server_oversized_update_returns_capacity(ctx):
    """A batch with more inserts than the whole log cannot be taken even after
    a compaction: replay returns DDP_ERR_CAPACITY instead of polling forever;
    the requests before it finished, and later queries run on the version the
    accepted batches made."""
    hg = ddp_amd.HostGraph.random(2000, 4, 1, 50, 61)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.reachable_pairs(8, 62)
    ub = hg.gen_updates(3, 30, 10, 1, 50, 63).batches
    big = hg.gen_updates(1, 100, 0, 1, 50, 64).batches[0]  # 100 inserts; the log holds 64
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, max_inserts=64, slots=4)
    n = 8
    types = np.array([0, 1, 0, 1, 0, 1, 0, 0], np.int32)
    ups = [None, ub[0], None, ub[1], None, big, None, None]
    with pytest.raises(ddp_amd.DDPError) as e:
        sv.replay(np.zeros(n), types, src[:n], dst[:n], ups)
    assert "DDP_ERR_CAPACITY" in str(e.value)
    sv.drain()
    sv.poll(64)
    assert sv.info()["version"] == 2  # the oversized third update was refused, the first two applied
    g2 = pyoracle.apply_batch(*pyoracle.apply_batch(hg.off, hg.col, hg.w, ub[0])[:3], ub[1])[:3]
    for k in range(4):
        assert sv.submit_query(int(src[k]), int(dst[k]), 100 + k) == 0
    sv.drain()
    done = {r["tag"]: r for r in sv.poll(64)}
    for k in range(4):
        d, _ = pyoracle.dijkstra(*g2, int(src[k]))
        r = done[100 + k]
        assert r["version"] == 2 and r["status"] == 0
        assert r["cost"] == (d[dst[k]] if d[dst[k]] >= 0 else ddp_amd.INF_COST)
    sv.close()


@pytest.mark.timeout(900)
def test_server_log_compaction_many_batches(ctx):
    """5,000 update batches of the C5 mix (deletes of live edges, random
    inserts) on a small graph whose log holds 40 batches: the server folds
    the log into its rows over a hundred times, and every query -- submitted
    between the batches, before and after compactions -- returns the oracle's
    Dijkstra cost on the graph version it was bound to."""
    hg = ddp_amd.HostGraph.random(3000, 6, 1, 1000, 91)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    nb, per = 5000, 25
    rng = np.random.default_rng(92)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, max_inserts=40 * per, slots=16)
    cur = (hg.off, hg.col, hg.w)
    pending = {}  # tag -> (version, s, t)
    checks = {}
    for b in range(nb):
        o = np.append(cur[0], len(cur[1])).astype(np.int64)
        deg = np.diff(o)
        us = rng.integers(0, hg.N, per)
        us = us[deg[us] > 0]
        lines = [(0, int(u), int(cur[1][o[u] + rng.integers(0, deg[u])]), 0) for u in us]
        lines += [(1, int(a), int(c), int(x)) for a, c, x in
                  zip(rng.integers(0, hg.N, per), rng.integers(0, hg.N, per), rng.integers(1, 1001, per))]
        assert sv.submit_update(lines, 10**6 + b) == 0
        cur = pyoracle.apply_batch(*cur, lines)[:3]
        if b % 50 == 49:  # a few queries on this version, checked against the oracle below
            v = b + 1
            for k in range(2):
                s, t = int(rng.integers(0, hg.N)), int(rng.integers(0, hg.N))
                tag = 10 * v + k
                assert sv.submit_query(s, t, tag) == 0
                d, _ = pyoracle.dijkstra(*cur, s)
                checks[tag] = (v, int(d[t]) if d[t] >= 0 else ddp_amd.INF_COST)
        for r in sv.poll(4096):
            pending[r["tag"]] = r
        if b % 1000 == 999:
            say("compaction test: batch", b + 1)
    sv.drain()
    for r in sv.poll(65536):
        pending[r["tag"]] = r
    assert all(r["status"] == 0 for r in pending.values())
    for tag, (v, want) in checks.items():
        r = pending[tag]
        assert r["version"] == v and r["cost"] == want, (tag, r, want)
    assert sv.info()["version"] == nb
    sv.close()


def test_server_bidi_versions_and_costs(ctx):
    """The bidirectional serving engine on a versioned graph: deletes hide
    transposed entries too, inserts appear on per-target lists, and every
    query equals Dijkstra on its version (relabelled and input numbering)."""
    hg = ddp_amd.HostGraph.random(6000, 5, 1, 100, 41)
    arrival, types, src, dst, ups = make_trace(hg, 400, 20, 42)
    for reorder, block in ((False, 256), (True, 1024)):
        dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=reorder)
        say("bidi server block", block, "relabelled", reorder)
        sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, block=block, slots=32, engine=ddp_amd.ENGINE_BIDI)
        res = sv.replay(arrival, types, src, dst, ups)
        check(res, types, src, dst, ups, hg)
        sv.close()
        dg.close()


def test_server_bidi_euclid_matches_batch(ctx):
    hg = ddp_amd.HostGraph.geometric(20000, 10, 100000, 43)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=True)
    ub = geometric_updates(hg, 3, 200, 44)
    src, dst = hg.geo_pairs(60, 500, 100000, 45)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_EUCLID, slots=16, engine=ddp_amd.ENGINE_BIDI)
    plan = []
    for v in range(4):
        if v > 0:
            assert sv.submit_update(ub[v - 1], 10000 + v) == 0
        for i in range(v * 15, v * 15 + 15):
            assert sv.submit_query(int(src[i]), int(dst[i]), i) == 0
            plan.append((v, i))
    sv.drain()
    got = {r["tag"]: r for r in sv.poll(1000)}
    sv.close()
    vers = versions(hg, ub)
    for v in range(4):
        g = ddp_amd.HostGraph(*vers[v], xy=hg.xy)
        dgv = ddp_amd.DeviceGraph(ctx, g)
        idx = [i for (vv, i) in plan if vv == v]
        cost, _, _ = ddp_amd.astar_batch(ctx, dgv, src[idx], dst[idx], heur=ddp_amd.H_EUCLID)
        for k, i in enumerate(idx):
            assert got[i]["version"] == v and got[i]["status"] == 0 and got[i]["cost"] == cost[k], (v, i)
        dgv.close()


@pytest.mark.timeout(600)
def test_server_bidi_compaction(ctx):
    """Compaction with the bidirectional engine rebuilds the transposed rows
    and their delete versions: 600 update batches through a log of 20 batches,
    queries checked against Dijkstra on their versions."""
    hg = ddp_amd.HostGraph.random(2000, 6, 1, 1000, 93)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    nb, per = 600, 20
    rng = np.random.default_rng(94)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, max_inserts=20 * per, slots=16, block=256,
                        engine=ddp_amd.ENGINE_BIDI)
    cur = (hg.off, hg.col, hg.w)
    got, checks = {}, {}
    for b in range(nb):
        o = np.append(cur[0], len(cur[1])).astype(np.int64)
        deg = np.diff(o)
        us = rng.integers(0, hg.N, per)
        us = us[deg[us] > 0]
        lines = [(0, int(u), int(cur[1][o[u] + rng.integers(0, deg[u])]), 0) for u in us]
        lines += [(1, int(a), int(c), int(x)) for a, c, x in
                  zip(rng.integers(0, hg.N, per), rng.integers(0, hg.N, per), rng.integers(1, 1001, per))]
        assert sv.submit_update(lines, 10**6 + b) == 0
        cur = pyoracle.apply_batch(*cur, lines)[:3]
        if b % 15 == 14:
            for k in range(2):
                s, t = int(rng.integers(0, hg.N)), int(rng.integers(0, hg.N))
                tag = 10 * (b + 1) + k
                assert sv.submit_query(s, t, tag) == 0
                d, _ = pyoracle.dijkstra(*cur, s)
                checks[tag] = (b + 1, int(d[t]) if d[t] >= 0 else ddp_amd.INF_COST)
        for r in sv.poll(4096):
            got[r["tag"]] = r
    sv.drain()
    for r in sv.poll(65536):
        got[r["tag"]] = r
    assert all(r["status"] == 0 for r in got.values())
    for tag, (v, want) in checks.items():
        assert got[tag]["version"] == v and got[tag]["cost"] == want, (tag, got[tag], want)
    sv.close()


def grid_mode_session(ctx, hg, dg, heur, engine, block, grid_k, src, dst, ub, max_inserts):
    """Lone queries (each submitted with nothing in flight: grid mode), bursts
    (the first alone, the rest while it runs), update batches -- every other
    one submitted while a burst is still in flight, the rest with nothing in
    flight (applied by a kernel of their own) -- and a log small enough to be
    compacted.  Returns {tag: result}, {tag: (version, s, t)} and the
    server's counters."""
    sv = ddp_amd.Server(ctx, dg, heur=heur, block=block, slots=8, engine=engine, grid_k=grid_k,
                        max_inserts=max_inserts)
    got, plan = {}, {}
    qi, v = 0, 0

    def query(tag):
        nonlocal qi
        s, t = int(src[qi % len(src)]), int(dst[qi % len(dst)])
        qi += 1
        assert sv.submit_query(s, t, tag) == 0
        plan[tag] = (v, s, t)

    def settle():
        sv.drain(120000)
        for r in sv.poll(4096):
            got[r["tag"]] = r

    for b in range(len(ub) + 1):
        for k in range(3):  # one at a time
            query(1000 * b + k)
            settle()
        for k in range(3, 8):  # a burst
            query(1000 * b + k)
        if b < len(ub) and b % 2 == 0:  # while the burst runs
            assert sv.submit_update(ub[b], 10**6 + b) == 0
            v += 1
            query(1000 * b + 8)
        settle()
        if b < len(ub) and b % 2 == 1:  # with nothing in flight: applied by its own kernel
            assert sv.submit_update(ub[b], 10**6 + b) == 0
            v += 1
            # at once a query on the new version, while that kernel may still be
            # queued behind a stopping serve kernel: it must not be claimed
            # there before the version exists
            query(1000 * b + 9)
            settle()
    st = sv.stats()
    say("  grid mode session: stats", st)
    sv.close()
    return got, plan, st


def test_server_grid_mode_at_low_load(ctx):
    """grid_k: a query submitted while no other request is in flight runs in
    grid mode and one submitted meanwhile in a slot; either way its cost is
    Dijkstra's on the graph version it was bound to -- across update batches,
    compactions and an idle serve kernel told to stop."""
    hg = ddp_amd.HostGraph.random(6000, 5, 1, 100, 51)
    src, dst = hg.reachable_pairs(80, 52)
    ub = hg.gen_updates(6, 40, 40, 1, 100, 53).batches
    vers = versions(hg, ub)
    for engine, block, grid_k, reverse in ((ddp_amd.ENGINE_PILES, 1024, 1024, False),
                                           (ddp_amd.ENGINE_BIDI, 256, 4096, True)):
        say("grid mode: engine", engine, "block", block, "K", grid_k)
        dg = ddp_amd.DeviceGraph(ctx, hg, reverse=reverse)
        got, plan, st = grid_mode_session(ctx, hg, dg, ddp_amd.H_ZERO, engine, block, grid_k, src, dst, ub, 100)
        dg.close()
        assert len(got) == len(plan) + len(ub)
        cache = {}
        for tag, (v, s, t) in plan.items():
            r = got[tag]
            if (v, s) not in cache:
                cache[(v, s)], _ = pyoracle.dijkstra(*vers[v], s)
            d = cache[(v, s)][t]
            assert r["status"] == 0 and r["version"] == v, (tag, r)
            assert r["cost"] == (int(d) if d >= 0 else ddp_amd.INF_COST), (tag, r, d)
            assert r["device_ms"] > 0
        # every lone query, and the first of each burst, ran in grid mode (the
        # rest of a burst goes to slots or queues behind it, as the measured
        # device times say); updates submitted with nothing in flight were
        # applied by their own kernel
        assert st["grid_queries"] >= 4 * (len(ub) + 1), st
        assert st["direct_updates"] >= len(ub) // 2, st
        assert st["requests"] == len(plan) + len(ub), st
        assert st["compactions"] >= 1 and st["launches"] >= 1, st


def test_server_grid_mode_euclid(ctx):
    """Grid mode aims the closed-form heuristic at each query's own target:
    Euclidean A* on a relabelled geometric graph with inserts that keep it
    admissible, checked against Dijkstra on every version."""
    hg = ddp_amd.HostGraph.geometric(20000, 10, 100000, 54)
    dg = ddp_amd.DeviceGraph(ctx, hg, reorder=True)
    ub = geometric_updates(hg, 3, 200, 55)
    src, dst = hg.geo_pairs(40, 500, 100000, 56)
    got, plan, st = grid_mode_session(ctx, hg, dg, ddp_amd.H_EUCLID, ddp_amd.ENGINE_PILES, 1024, 2048, src, dst,
                                      ub, 0)
    dg.close()
    vers = versions(hg, ub)
    cache = {}
    for tag, (v, s, t) in plan.items():
        r = got[tag]
        if (v, s) not in cache:
            cache[(v, s)], _ = pyoracle.dijkstra(*vers[v], s)
        d = cache[(v, s)][t]
        assert r["status"] == 0 and r["version"] == v and r["cost"] == int(d), (tag, r, d)
    assert st["grid_queries"] >= 3 * (len(ub) + 1), st


def maze_updates(hg, width, nb, per, seed):
    """Update batches that keep the Manhattan heuristic admissible: deletes of
    existing edges, inserts between any two cells weighted at least their
    Manhattan distance."""
    rng = np.random.default_rng(seed)
    o = np.append(hg.off, len(hg.col)).astype(np.int64)
    deg = np.diff(o)
    out = []
    for _ in range(nb):
        us = rng.integers(0, hg.N, 4 * per)
        us = us[deg[us] > 0][:per]
        lines = [(0, int(u), int(hg.col[o[u] + rng.integers(0, deg[u])]), 0) for u in us]
        for a, b in zip(rng.integers(0, hg.N, per), rng.integers(0, hg.N, per)):
            md = abs(int(a) // width - int(b) // width) + abs(int(a) % width - int(b) % width)
            lines.append((1, int(a), int(b), md + int(rng.integers(0, 3))))
        out.append(lines)
    return out


def test_server_grid_mode_manhattan(ctx):
    """Grid mode aims the Manhattan heuristic at each query's own target cell
    (row and column from the grid width) on a maze with updates that keep it
    admissible; every cost equals Dijkstra's on its version."""
    width = 256
    hg = ddp_amd.HostGraph.grid_maze(width, 200, 300, 57)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    dg.set_grid(width)
    src, dst = hg.reachable_pairs(40, 58)
    ub = maze_updates(hg, width, 3, 30, 59)
    got, plan, st = grid_mode_session(ctx, hg, dg, ddp_amd.H_MANHATTAN, ddp_amd.ENGINE_PILES, 1024, 2048, src, dst,
                                      ub, 0)
    dg.close()
    vers = versions(hg, ub)
    cache = {}
    for tag, (v, s, t) in plan.items():
        r = got[tag]
        if (v, s) not in cache:
            cache[(v, s)], _ = pyoracle.dijkstra(*vers[v], s)
        d = cache[(v, s)][t]
        assert r["status"] == 0 and r["version"] == v, (tag, r)
        assert r["cost"] == (int(d) if d >= 0 else ddp_amd.INF_COST), (tag, r, d)
    assert st["grid_queries"] >= 4 * (len(ub) + 1), st
