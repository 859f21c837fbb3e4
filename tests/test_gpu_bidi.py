"""GPU parity of the bidirectional batch engine (DDP_BATCH_BIDI): a forward
search from s and a backward search from t in one workgroup, meeting in the
middle.  Every cost equals the oracle's shortest distance (and the one-way
pile engine's), every returned path is a real s..t path of that cost, on
graphs with parallel edges, deleted edges, zero weights, unreachable targets
and s == t, under the zero, Manhattan and Euclidean heuristics."""
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


def dijkstra_costs(off, col, w, src, dst):
    out = np.zeros(len(src), np.int64)
    for i, (s, t) in enumerate(zip(src, dst)):
        d, _ = pyoracle.dijkstra(off, col, w, int(s), int(t))
        out[i] = d[t] if d[t] >= 0 else ddp_amd.INF_COST
    return out


def check_paths(off, col, w, src, dst, costs, paths):
    for i in range(len(src)):
        if costs[i] == ddp_amd.INF_COST:
            assert len(paths[i]) == 0
            continue
        p = paths[i]
        assert len(p) and p[0] == src[i] and p[-1] == dst[i], i
        assert path_cost(off, col, w, p) == costs[i], i


@pytest.mark.parametrize("block", [256, 1024])
def test_bidi_random_graphs_zero_h(ctx, block):
    rng = np.random.default_rng(101 + block)
    for k in range(6):
        N = int(rng.integers(50, 3000))
        off, col, w = random_graph(rng, N, max_deg=int(rng.integers(1, 6)), wmax=int(rng.choice([1, 5, 100])),
                                   tombstones=0.05 if k % 2 else 0.0)
        hg = ddp_amd.HostGraph(off, col, w)
        dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=bool(k % 3 == 2))
        nq = 64
        src = rng.integers(0, N, nq).astype(np.int32)
        dst = rng.integers(0, N, nq).astype(np.int32)
        dst[:4] = src[:4]  # s == t
        say("bidi random", k, "N", N, "block", block)
        costs, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, block=block, path_cap=N + 1,
                                               engine=ddp_amd.ENGINE_BIDI, slots=8)
        assert st["errors"] == 0
        want = dijkstra_costs(off, col, w, src, dst)
        assert np.array_equal(costs, want), np.flatnonzero(costs != want)[:5]
        check_paths(off, col, w, src, dst, costs, paths)
        dg.close()


def test_bidi_zero_weight_edges(ctx):
    """w = 0 edges: relaxation by strict-decrease CAS keeps both parent trees acyclic."""
    rng = np.random.default_rng(7)
    N = 2000
    off, col, w = random_graph(rng, N, max_deg=4, wmax=3)
    w = (w - 1).astype(np.uint32)  # about a third of the weights are 0
    hg = ddp_amd.HostGraph(off, col, w)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    src, dst = hg.reachable_pairs(48, 8)
    costs, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, block=512, path_cap=N + 1,
                                           engine=ddp_amd.ENGINE_BIDI)
    assert st["errors"] == 0
    assert np.array_equal(costs, dijkstra_costs(off, col, w, src, dst))
    check_paths(off, col, w, src, dst, costs, paths)


def test_bidi_geometric_euclid_matches_piles(ctx):
    """Euclidean heuristic: the average potential (h_t - h_s) / 2 keeps costs
    exact; the same queries through the one-way pile engine agree, and a
    sample matches the oracle's Dijkstra."""
    hg = ddp_amd.HostGraph.geometric(60000, 10, 100000, 11)
    src, dst = hg.geo_pairs(512, 100, 100000, 12)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=True)
    c1, p1, s1 = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=1024, path_cap=8192,
                                     engine=ddp_amd.ENGINE_BIDI)
    c0, _, s0 = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=1024,
                                    engine=ddp_amd.ENGINE_PILES)
    say("bidi euclid: expansions", s1["expansions"], "one-way", s0["expansions"])
    assert s1["errors"] == 0 and s0["errors"] == 0
    assert np.array_equal(c1, c0)
    check_paths(hg.off, hg.col, hg.w, src, dst, c1, p1)
    idx = np.arange(0, len(src), 32)
    assert np.array_equal(c1[idx], dijkstra_costs(hg.off, hg.col, hg.w, src[idx], dst[idx]))


def test_bidi_grid_manhattan(ctx):
    W, H = 200, 150
    hg = ddp_amd.HostGraph.grid_maze(W, H, 300, 13)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    dg.set_grid(W)
    rng = np.random.default_rng(14)
    src = rng.integers(0, W * H, 40).astype(np.int32)
    dst = rng.integers(0, W * H, 40).astype(np.int32)
    src[0], dst[0] = 0, W * H - 1
    costs, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_MANHATTAN, block=512,
                                           path_cap=W * H, engine=ddp_amd.ENGINE_BIDI)
    assert st["errors"] == 0
    assert np.array_equal(costs, dijkstra_costs(hg.off, hg.col, hg.w, src, dst))
    check_paths(hg.off, hg.col, hg.w, src, dst, costs, paths)


def test_bidi_rmat_hubs(ctx):
    """Power-law rows (hub rows wider than a round) in both directions."""
    hg = ddp_amd.HostGraph.rmat(14, 16 * (1 << 14), 0.57, 0.19, 0.19, 1, 1000, 15)
    src, dst = hg.hub_pairs(64, 16)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=True)
    costs, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, path_cap=4096,
                                           engine=ddp_amd.ENGINE_BIDI)
    c0, _, s0 = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, engine=ddp_amd.ENGINE_PILES)
    say("bidi rmat: edges", st["edges"], "one-way", s0["edges"])
    assert st["errors"] == 0 and np.array_equal(costs, c0)
    check_paths(hg.off, hg.col, hg.w, src, dst, costs, paths)
    idx = np.arange(0, len(src), 8)
    assert np.array_equal(costs[idx], dijkstra_costs(hg.off, hg.col, hg.w, src[idx], dst[idx]))


def test_bidi_refuses_what_it_cannot_do(ctx):
    hg = ddp_amd.HostGraph.random(500, 3, 1, 9, 17)
    dg = ddp_amd.DeviceGraph(ctx, hg)  # no transposed rows
    with pytest.raises(ddp_amd.DDPError):
        ddp_amd.astar_batch(ctx, dg, [0], [1], heur=ddp_amd.H_ZERO, engine=ddp_amd.ENGINE_BIDI)
    dg2 = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    dg2.bfs_hops(1, reverse=True, store_as_h=True)
    with pytest.raises(ddp_amd.DDPError):  # one-target h arrays cannot give h toward s
        ddp_amd.astar_batch(ctx, dg2, [0], [1], heur=ddp_amd.H_ARRAY, engine=ddp_amd.ENGINE_BIDI)
    with pytest.raises(ddp_amd.DDPError):
        ddp_amd.astar_batch(ctx, dg2, [0], [1], heur=ddp_amd.H_ZERO, engine=ddp_amd.ENGINE_BIDI, block=128)


@pytest.mark.parametrize("K", [256, 1024])
def test_bidi_single_query_mode(ctx, K):
    """ddp_astar with DDP_ASTAR_BIDI: the cost of the oracle's Dijkstra, the
    path walked on the device through the meeting node, and a parent array
    whose chain from t leads back to s along that path."""
    rng = np.random.default_rng(K)
    hg = ddp_amd.HostGraph.random(20000, 5, 1, 100, 31)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=True)
    src, dst = hg.reachable_pairs(6, 32)
    src = np.append(src, [int(rng.integers(0, hg.N))] * 2)
    dst = np.append(dst, [src[-1], int(rng.integers(0, hg.N))])  # s == t, and a random pair
    for s, t in zip(src, dst):
        s, t = int(s), int(t)
        r = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_ZERO, K=K, mode=ddp_amd.MODE_BIDI, want_parent=True)
        d, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, s, t)
        want = int(d[t]) if d[t] >= 0 else ddp_amd.INF_COST
        assert r.cost == want, (s, t)
        if want == ddp_amd.INF_COST:
            assert len(r.path) == 0
            continue
        assert r.path[0] == s and r.path[-1] == t and path_cost(hg.off, hg.col, hg.w, r.path) == want
        chain = pyoracle.chain_from_parent(r.parent, t)
        assert chain[::-1] == r.path.tolist()


def test_bidi_single_query_maze(ctx):
    """A 512 x 512 maze corner to corner under the Manhattan heuristic: the
    oracle's A* cost, about half the rounds of the one-way pile mode."""
    W = 512
    hg = ddp_amd.HostGraph.grid_maze(W, W, 300, 2)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    dg.set_grid(W)
    s, t = 0, W * W - 1
    xs = np.arange(W * W, dtype=np.int64)
    h = ((W - 1 - xs // W) + (W - 1 - xs % W)).astype(np.int32)
    found, cost, _ = pyoracle.astar_indexed(hg.off, hg.col, hg.w, h, s, t)
    rb = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_MANHATTAN, K=1024, mode=ddp_amd.MODE_BIDI)
    rp = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_MANHATTAN, K=1024, mode=ddp_amd.MODE_PILES)
    say("maze bidi rounds", rb.stats["iterations"], "one-way", rp.stats["iterations"])
    assert found and rb.cost == cost == rp.cost
    assert path_cost(hg.off, hg.col, hg.w, rb.path) == cost and rb.path[0] == s and rb.path[-1] == t
    assert rb.stats["iterations"] < rp.stats["iterations"]
