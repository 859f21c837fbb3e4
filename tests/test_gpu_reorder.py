"""GPU parity of relabelled device graphs (DDP_GRAPH_REORDER): the device
numbers nodes along a Hilbert curve (or breadth-first without coordinates),
while every entry point keeps taking and returning the caller's ids.  Costs
must equal the oracle's and the unrelabelled run's; paths, parent arrays and
heuristic arrays must come back in the caller's numbering."""
import sys

import numpy as np
import pytest

import ddp_amd
from helpers import geometric_updates, path_cost, random_graph
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_path(off, col, w, path, s, t, g_t):
    assert len(path) >= 1 and path[0] == s and path[-1] == t
    assert path_cost(off, col, w, path) == g_t


def test_reorder_batch_and_single_geometric(ctx):
    hg = ddp_amd.HostGraph.geometric(40000, 10, 100000, 21)
    src, dst = hg.geo_pairs(200, 100, 100000, 22)
    plain = ddp_amd.DeviceGraph(ctx, hg)
    relab = ddp_amd.DeviceGraph(ctx, hg, reorder=True)
    c0, p0, _ = ddp_amd.astar_batch(ctx, plain, src, dst, heur=ddp_amd.H_EUCLID, path_cap=4096)
    c1, p1, _ = ddp_amd.astar_batch(ctx, relab, src, dst, heur=ddp_amd.H_EUCLID, path_cap=4096)
    assert np.array_equal(c0, c1)
    for i in range(0, 200, 7):
        s, t = int(src[i]), int(dst[i])
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, s, t)
        say("reorder geo", i, s, t)
        assert c1[i] == dist[t]
        check_path(hg.off, hg.col, hg.w, p1[i], s, t, int(dist[t]))
        for mode in (ddp_amd.MODE_GRID, ddp_amd.MODE_WORKGROUP, ddp_amd.MODE_PILES):
            r = ddp_amd.astar(ctx, relab, s, t, heur=ddp_amd.H_EUCLID, K=1024, mode=mode, want_parent=True)
            assert r.cost == dist[t]
            check_path(hg.off, hg.col, hg.w, r.path, s, t, int(dist[t]))
            assert pyoracle.chain_from_parent(r.parent, t) == [int(x) for x in r.path[::-1]]


def test_reorder_bfs_order_and_h_arrays(ctx):
    """Random graph without coordinates: breadth-first relabelling; h arrays
    given and returned in the caller's order."""
    rng = np.random.default_rng(23)
    off, col, w = random_graph(rng, 3000, max_deg=5, wmax=40, tombstones=0.02)
    hg = ddp_amd.HostGraph(off, col, w)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=True)
    roff, rcol, _ = pyoracle.transpose(off, col, w)
    for t in (0, 1234, 2999):
        hops, lv = pyoracle.bfs_hops(roff, rcol, t)
        want = np.where(hops < 0, -1, hops * 3).astype(np.int32)
        got, glv = dg.bfs_hops(t, reverse=True, scale=3)
        assert np.array_equal(got, want) and glv == lv
    srcs, dsts = ddp_amd.HostGraph(off, col, w).reachable_pairs(12, 24)
    minw = int(w[col >= 0].min())
    for s, t in zip(srcs, dsts):
        s, t = int(s), int(t)
        hops, _ = pyoracle.bfs_hops(roff, rcol, t)
        h = np.where(hops < 0, -1, hops * minw).astype(np.int32)
        dist, _ = pyoracle.dijkstra(off, col, w, s)
        for how in ("set_h", "bfs"):
            if how == "set_h":
                dg.set_h(h)
            else:
                got, _ = dg.bfs_hops(t, reverse=True, scale=minw, store_as_h=True)
                assert np.array_equal(got, h)
            r = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_ARRAY, K=256, want_parent=True)
            say("reorder bfs", s, t, how, r.cost)
            assert r.cost == dist[t]
            check_path(off, col, w, r.path, s, t, int(dist[t]))
            assert pyoracle.chain_from_parent(r.parent, t) == [int(x) for x in r.path[::-1]]
    with pytest.raises(ddp_amd.DDPError):
        dg.set_grid(1000)


def test_reorder_server_updates(ctx):
    """Updates name edges by caller ids; queries on each version agree with a
    from-scratch Dijkstra on that version."""
    hg = ddp_amd.HostGraph.geometric(15000, 10, 100000, 25)
    dg = ddp_amd.DeviceGraph(ctx, hg, reorder=True)
    ub = geometric_updates(hg, 3, 150, 26)
    src, dst = hg.geo_pairs(40, 500, 100000, 27)
    sv = ddp_amd.Server(ctx, dg, heur=ddp_amd.H_EUCLID, slots=16)
    plan = []
    for v in range(4):
        if v > 0:
            assert sv.submit_update(ub[v - 1], 10000 + v) == 0
        for i in range(v * 10, v * 10 + 10):
            assert sv.submit_query(int(src[i]), int(dst[i]), i) == 0
            plan.append((v, i))
    sv.drain()
    got = {r["tag"]: r for r in sv.poll(1000)}
    sv.close()
    vers = [(hg.off, hg.col, hg.w)]
    for b in ub:
        vers.append(pyoracle.apply_batch(*vers[-1], b)[:3])
    for v, i in plan:
        dist, _ = pyoracle.dijkstra(*vers[v], int(src[i]), int(dst[i]))
        want = int(dist[dst[i]]) if dist[dst[i]] >= 0 else ddp_amd.INF_COST
        assert got[i]["version"] == v and got[i]["cost"] == want, (v, i)
