"""Zero-weight edges, zero-weight cycles included.  Graphs whose weights
are all >= 1 relax with one atomicMin per edge; a graph with a zero weight
(or a served version after a zero-weight insert) must keep the
strict-decrease CAS so that parent chains stay acyclic.  Costs equal
Dijkstra, paths are real paths of that cost, parent chains end at s."""
import sys

import numpy as np
import pytest

import ddp_amd
from helpers import path_cost, random_graph
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def zero_weight_graph(seed, N=3000):
    rng = np.random.default_rng(seed)
    off, col, w = random_graph(rng, N, max_deg=6, wmax=4)
    w[rng.random(len(w)) < 0.35] = 0
    # explicit zero-weight 2-cycles u <-> v
    for u in rng.integers(0, N, 40):
        u = int(u)
        j = int(off[u])
        end = int(off[u + 1]) if u + 1 < N else len(col)
        if j < end:
            v = int(col[j])
            w[j] = 0
            jv = int(off[v])
            endv = int(off[v + 1]) if v + 1 < N else len(col)
            if jv < endv:
                col[jv], w[jv] = u, 0
    return off, col, w


def check_path(off, col, w, path, s, t, cost):
    assert len(path) >= 1 and path[0] == s and path[-1] == t
    assert path_cost(off, col, w, path) == cost


def test_zero_weight_batches_and_single_queries(ctx):
    off, col, w = zero_weight_graph(51)
    hg = ddp_amd.HostGraph(off, col, w)
    src, dst = hg.reachable_pairs(40, 52)
    want = []
    for s, t in zip(src, dst):
        dist, _ = pyoracle.dijkstra(off, col, w, int(s), int(t))
        want.append(int(dist[t]))
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    assert dg.info()["min_weight"] == 0
    for engine in (ddp_amd.ENGINE_PILES, ddp_amd.ENGINE_HEAPS):
        for block in (64, 256):
            say("zero-weight batch", engine, block)
            cost, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, block=block,
                                                  path_cap=4096, engine=engine)
            assert st["errors"] == 0
            assert [int(c) for c in cost] == want
            for i in range(0, 40, 3):
                check_path(off, col, w, paths[i], int(src[i]), int(dst[i]), want[i])
    for mode in (ddp_amd.MODE_GRID, ddp_amd.MODE_WORKGROUP, ddp_amd.MODE_PILES):
        for i in range(0, 40, 8):
            s, t = int(src[i]), int(dst[i])
            r = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_ZERO, K=256, mode=mode, want_parent=True)
            say("zero-weight single", mode, s, t, r.cost)
            assert r.cost == want[i]
            check_path(off, col, w, r.path, s, t, want[i])
            chain = pyoracle.chain_from_parent(r.parent, t)
            assert chain[-1] == s and chain == [int(x) for x in r.path[::-1]]
    dg.close()


def test_server_zero_weight_insert(ctx):
    """Base weights >= 1 (atomicMin relaxation) until an update inserts
    zero-weight edges; queries on later versions must stay exact."""
    rng = np.random.default_rng(53)
    off, col, w = random_graph(rng, 2500, max_deg=5, wmax=30)
    hg = ddp_amd.HostGraph(off, col, w)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.reachable_pairs(24, 54)
    ins = []
    for k in range(300):
        u, v = int(rng.integers(0, 2500)), int(rng.integers(0, 2500))
        ins.append((1, u, v, 0))
        ins.append((1, v, u, 0))
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_ZERO, slots=8)
    for i in range(12):
        assert sv.submit_query(int(src[i]), int(dst[i]), i) == 0
    assert sv.submit_update(ins, 1000) == 0
    for i in range(12, 24):
        assert sv.submit_query(int(src[i]), int(dst[i]), i) == 0
    sv.drain()
    got = {r["tag"]: r for r in sv.poll(100)}
    sv.close()
    v1 = pyoracle.apply_batch(off, col, w, ins)[:3]
    for i in range(24):
        g = (off, col, w) if i < 12 else v1
        dist, _ = pyoracle.dijkstra(*g, int(src[i]), int(dst[i]))
        want = int(dist[dst[i]]) if dist[dst[i]] >= 0 else ddp_amd.INF_COST
        say("server zero-weight", i, got[i]["cost"], want)
        assert got[i]["version"] == (0 if i < 12 else 1) and got[i]["cost"] == want
