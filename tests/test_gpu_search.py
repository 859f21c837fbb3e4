"""GPU parity: reverse BFS, single-query GA* and batched A* through the C ABI
against the CPU oracle on the same seeded inputs.  Integer results must be
bit-exact; every returned path must be a real path of the returned cost."""
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


def bfs_h_oracle(off, col, w, t, scale):
    roff, rcol, rw = pyoracle.transpose(off, col, w)
    hops, lv = pyoracle.bfs_hops(roff, rcol, t)
    return np.where(hops < 0, -1, hops * scale).astype(np.int32), lv


def check_path(off, col, w, path, s, t, g_t):
    assert len(path) >= 1 and path[0] == s and path[-1] == t
    assert path_cost(off, col, w, path) == g_t


def test_bfs_hops_parity(ctx):
    rng = np.random.default_rng(101)
    for k in range(40):
        N = int(rng.integers(1, 400))
        off, col, w = random_graph(rng, N, max_deg=int(rng.integers(1, 6)), tombstones=0.05 if k % 4 == 0 else 0)
        hg = ddp_amd.HostGraph(off, col, w)
        dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
        t = int(rng.integers(0, N))
        scale = int(rng.integers(1, 4))
        say("bfs case", k, "N", N, "t", t)
        want, lv = bfs_h_oracle(off, col, w, t, scale)
        got, glv = dg.bfs_hops(t, reverse=True, scale=scale)
        assert np.array_equal(got, want) and glv == lv
        # reverse=0 on the transposed graph is the reference's own use (bfs_hx on graph_op)
        roff, rcol, rw = pyoracle.transpose(off, col, w)
        dgr = ddp_amd.DeviceGraph(ctx, ddp_amd.HostGraph(roff, rcol, rw))
        got0, _ = dgr.bfs_hops(t, reverse=False, scale=scale)
        assert np.array_equal(got0, want)


def test_bfs_hops_config1(ctx):
    hg = ddp_amd.HostGraph.random(100000, 6, 1, 100, 1)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    for t in (0, 777, 99999):
        say("bfs C1 t", t)
        want, lv = bfs_h_oracle(hg.off, hg.col, hg.w, t, 1)
        got, glv = dg.bfs_hops(t, reverse=True, scale=1)
        assert np.array_equal(got, want) and glv == lv


@pytest.mark.parametrize("K,block,mode", [(1, 64, ddp_amd.MODE_AUTO), (64, 64, ddp_amd.MODE_AUTO),
                                          (200, 0, ddp_amd.MODE_WORKGROUP), (1024, 0, ddp_amd.MODE_WORKGROUP),
                                          (1024, 64, ddp_amd.MODE_GRID), (1024, 256, ddp_amd.MODE_GRID),
                                          (4096, 1024, ddp_amd.MODE_AUTO), (100, 0, ddp_amd.MODE_PILES),
                                          (1024, 0, ddp_amd.MODE_PILES)])
def test_astar_small_graphs(ctx, K, block, mode):
    """Grid mode (K queues over K/block workgroups) and workgroup mode (all K
    queues in one workgroup) against the oracle's CPU A* and Dijkstra."""
    rng = np.random.default_rng(1000 + K + block + 7 * mode)
    for k in range(60):
        N = int(rng.integers(2, 300))
        off, col, w = random_graph(rng, N, max_deg=int(rng.integers(1, 7)), wmax=int(rng.choice([1, 3, 100])),
                                   tombstones=0.05 if k % 5 == 0 else 0)
        s = int(rng.integers(0, N))
        dist, _ = pyoracle.dijkstra(off, col, w, s)
        reach = np.flatnonzero(dist >= 0)
        t = int(rng.choice(reach)) if rng.random() < 0.85 else int(rng.integers(0, N))
        hg = ddp_amd.HostGraph(off, col, w)
        dg = ddp_amd.DeviceGraph(ctx, hg)
        hmode = k % 3
        if hmode == 0:
            h = np.zeros(N, np.int32)
        else:
            h, _ = bfs_h_oracle(off, col, w, t, int(w[col >= 0].min(initial=1)) if hmode == 1 else 1)
        dg.set_h(h)
        say("astar case", K, block, mode, k, "N", N, "s", s, "t", t, "h", hmode)
        r = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_ARRAY, K=K, block=block, want_parent=True, mode=mode)
        found, cost, parent, _ = pyoracle.astar_ref(off, col, w, h, s, t, quirk=False)
        if not found or (h[s] < 0):
            assert r.cost == ddp_amd.INF_COST and len(r.path) == 0
            continue
        assert r.cost == cost == dist[t] + h[t]
        check_path(off, col, w, r.path, s, t, int(dist[t]))
        # the parent array holds the same chain (reversed)
        assert pyoracle.chain_from_parent(r.parent, t) == [int(x) for x in r.path[::-1]]


def test_astar_config1(ctx):
    """Config 1: random N=100,000, 6 out-edges, w in [1,100], seed 1,
    h = reverse-BFS hops x min weight, K = 1024."""
    hg = ddp_amd.HostGraph.random(100000, 6, 1, 100, 1)
    src, dst = hg.reachable_pairs(3, 1)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)
    minw = int(hg.w.min())
    for s, t in zip(src, dst):
        s, t = int(s), int(t)
        h, _ = dg.bfs_hops(t, reverse=True, scale=minw, store_as_h=True)
        say("C1 s", s, "t", t)
        r = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_ARRAY, K=1024)
        for mode in (ddp_amd.MODE_WORKGROUP, ddp_amd.MODE_PILES):
            rg = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_ARRAY, K=1024, mode=mode)
            assert rg.cost == r.cost
        found_q, cost_q, _, _ = pyoracle.astar_ref(hg.off, hg.col, hg.w, h, s, t, quirk=True)
        found, cost, _, _ = pyoracle.astar_ref(hg.off, hg.col, hg.w, h, s, t, quirk=False)
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, s, t)
        assert found and r.cost == cost == dist[t]
        if found_q:  # the reference CPU binary's answer (edge E-1 dropped) agrees unless that edge matters
            assert cost_q >= cost
        check_path(hg.off, hg.col, hg.w, r.path, s, t, int(dist[t]))
        say("  stats", r.stats)


def test_astar_grid_manhattan(ctx):
    hg = ddp_amd.HostGraph.grid_maze(128, 96, 300, 2)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    dg.set_grid(128)
    s, t = 0, 128 * 96 - 1
    dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, s)
    assert dist[t] > 0
    for K, mode in ((512, ddp_amd.MODE_AUTO), (512, ddp_amd.MODE_WORKGROUP), (1024, ddp_amd.MODE_PILES),
                    (2048, ddp_amd.MODE_AUTO),
                    (8192, ddp_amd.MODE_AUTO)):
        say("grid K", K, "mode", mode)
        r = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_MANHATTAN, K=K, mode=mode)
        assert r.cost == dist[t]
        check_path(hg.off, hg.col, hg.w, r.path, s, t, int(dist[t]))


def test_astar_euclid_and_zero(ctx):
    hg = ddp_amd.HostGraph.geometric(20000, 10, 100000, 4)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.geo_pairs(6, 200, 100000, 5)
    for s, t in zip(src, dst):
        s, t = int(s), int(t)
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, s, t)
        for kind in (ddp_amd.H_EUCLID, ddp_amd.H_ZERO):
            say("geo", s, t, kind)
            r = ddp_amd.astar(ctx, dg, s, t, heur=kind, K=1024)
            hx = hg.xy[2 * s] - hg.xy[2 * t]
            assert r.cost == dist[t]
            check_path(hg.off, hg.col, hg.w, r.path, s, t, int(dist[t]))


def test_astar_edge_cases(ctx):
    off = np.array([0, 1, 2], np.int32)
    col = np.array([1, 2, 0], np.int32)
    w = np.array([4, 5, 6], np.uint32)
    dg = ddp_amd.DeviceGraph(ctx, ddp_amd.HostGraph(off, col, w))
    r = ddp_amd.astar(ctx, dg, 1, 1, heur=ddp_amd.H_ZERO)  # s == t
    assert r.cost == 0 and r.path.tolist() == [1]
    r = ddp_amd.astar(ctx, dg, 2, 0, heur=ddp_amd.H_ZERO)  # uses edge E-1 (no row-end quirk on the GPU)
    assert r.cost == 6 and r.path.tolist() == [2, 0]
    dg.set_h(np.array([-1, 0, 0], np.int32))
    r = ddp_amd.astar(ctx, dg, 0, 2, heur=ddp_amd.H_ARRAY)  # h(s) = -1: cannot reach
    assert r.cost == ddp_amd.INF_COST
    with pytest.raises(ddp_amd.DDPError):
        ddp_amd.astar(ctx, dg, 0, 5, heur=ddp_amd.H_ZERO)
    iso = ddp_amd.DeviceGraph(ctx, ddp_amd.HostGraph(np.zeros(4, np.int32), np.zeros(0, np.int32),
                                                      np.zeros(0, np.uint32)))
    r = ddp_amd.astar(ctx, iso, 0, 3, heur=ddp_amd.H_ZERO)
    assert r.cost == ddp_amd.INF_COST and len(r.path) == 0


def test_batch_parity_geometric(ctx):
    hg = ddp_amd.HostGraph.geometric(50000, 10, 100000, 4)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.geo_pairs(300, 100, 100000, 6)
    for block in (64, 256):
        say("batch block", block)
        cost, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=block,
                                              path_cap=4096)
        say("  stats", st)
        for i, (s, t) in enumerate(zip(src, dst)):
            dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(s), int(t))
            assert cost[i] == dist[t], (i, s, t)
            check_path(hg.off, hg.col, hg.w, paths[i], int(s), int(t), int(dist[t]))


def test_heap_spill_to_hbm(ctx):
    """Tiny LDS parts force most heap entries into the HBM spill area."""
    hg = ddp_amd.HostGraph.geometric(20000, 10, 100000, 14)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.geo_pairs(64, 2000, 100000, 15)
    want = []
    for s, t in zip(src, dst):
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(s), int(t))
        want.append(int(dist[t]))
    for i in range(0, 64, 16):
        say("spill single", i)
        for mode in (ddp_amd.MODE_WORKGROUP, ddp_amd.MODE_GRID):
            r = ddp_amd.astar(ctx, dg, int(src[i]), int(dst[i]), heur=ddp_amd.H_EUCLID, K=256, block=256, lds_cap=1,
                              mode=mode)
            assert r.cost == want[i]
    for block, lcap in ((64, 4), (256, 4), (512, 4), (128, 1)):
        say("spill batch", block, lcap)
        cost, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=block, lds_cap=lcap,
                                              slots=7, path_cap=2048)
        assert cost.tolist() == want
        for i in range(64):
            check_path(hg.off, hg.col, hg.w, paths[i], int(src[i]), int(dst[i]), want[i])


def test_batch_same_alone_and_in_the_mix(ctx):
    hg = ddp_amd.HostGraph.geometric(30000, 10, 100000, 8)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.geo_pairs(200, 100, 100000, 9)
    mixed, _, _ = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, slots=37)
    for i in range(0, 200, 20):
        alone, _, _ = ddp_amd.astar_batch(ctx, dg, src[i:i + 1], dst[i:i + 1], heur=ddp_amd.H_EUCLID)
        single = ddp_amd.astar(ctx, dg, int(src[i]), int(dst[i]), heur=ddp_amd.H_EUCLID, K=256)
        assert alone[0] == mixed[i] == single.cost


def test_batch_plan_replay(ctx):
    """The captured hipGraph is replayed several times with host work in
    between; every replay must give the eager result."""
    hg = ddp_amd.HostGraph.random(20000, 6, 1, 100, 21)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.reachable_pairs(64, 22)
    b = ddp_amd.Batch(ctx, dg, 64, heur=ddp_amd.H_ZERO, block=64, path_cap=0)
    b.upload(src, dst)
    first = None
    for rep in range(4):
        b.launch()
        st = b.wait()
        out = b.download()
        assert st["errors"] == 0
        if first is None:
            first = out["cost"].copy()
            for i in range(0, 64, 8):
                dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[i]), int(dst[i]))
                assert first[i] == dist[dst[i]]
        assert np.array_equal(out["cost"], first)
    b.close()


@pytest.mark.parametrize("block", [64, 256, 512])
def test_batch_piles_parity(ctx, block):
    """The pile engine (near ring + far pile, no heaps) returns the oracle's
    costs and real paths on geometric (Euclid), random (zero and BFS
    heuristics) and grid-maze (Manhattan) inputs."""
    hg = ddp_amd.HostGraph.geometric(30000, 10, 100000, 31)
    src, dst = hg.geo_pairs(96, 100, 100000, 32)
    for reorder in (False, True):
        dg = ddp_amd.DeviceGraph(ctx, hg, reorder=reorder)
        say("piles geo", block, reorder)
        cost, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=block,
                                              path_cap=4096, engine=ddp_amd.ENGINE_PILES)
        assert st["errors"] == 0
        for i in range(0, 96, 5):
            dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[i]), int(dst[i]))
            assert cost[i] == dist[dst[i]], (i, cost[i], dist[dst[i]])
            check_path(hg.off, hg.col, hg.w, paths[i], int(src[i]), int(dst[i]), int(dist[dst[i]]))
        heaps, _, _ = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=block,
                                          engine=ddp_amd.ENGINE_HEAPS)
        assert np.array_equal(cost, heaps)
        dg.close()
    rng = np.random.default_rng(33 + block)
    off, col, w = random_graph(rng, 4000, max_deg=6, wmax=100, tombstones=0.03)
    rg = ddp_amd.HostGraph(off, col, w)
    dg = ddp_amd.DeviceGraph(ctx, rg, reverse=True)
    rs, rt = rg.reachable_pairs(40, 34)
    rs = np.append(rs, [0, 5]).astype(np.int32)  # plus two arbitrary pairs (maybe unreachable)
    rt = np.append(rt, [3999, 5]).astype(np.int32)
    say("piles random zero", block)
    cost, paths, st = ddp_amd.astar_batch(ctx, dg, rs, rt, heur=ddp_amd.H_ZERO, block=block, path_cap=4000,
                                          engine=ddp_amd.ENGINE_PILES)
    assert st["errors"] == 0
    for i in range(len(rs)):
        dist, _ = pyoracle.dijkstra(off, col, w, int(rs[i]))
        want = int(dist[rt[i]]) if dist[rt[i]] >= 0 else ddp_amd.INF_COST
        assert cost[i] == want, (i, cost[i], want)
        if want != ddp_amd.INF_COST:
            check_path(off, col, w, paths[i], int(rs[i]), int(rt[i]), want)
    # one target, its reverse-BFS heuristic (h = -1 prunes)
    t = int(rt[0])
    h, _ = dg.bfs_hops(t, reverse=True, scale=int(w[col >= 0].min()), store_as_h=True)
    ss = np.array(rs[:20], np.int32)
    say("piles random bfs-h", block)
    cost, _, st = ddp_amd.astar_batch(ctx, dg, ss, np.full(20, t, np.int32), heur=ddp_amd.H_ARRAY, block=block,
                                      engine=ddp_amd.ENGINE_PILES)
    for i in range(20):
        dist, _ = pyoracle.dijkstra(off, col, w, int(ss[i]))
        want = int(dist[t]) if dist[t] >= 0 else ddp_amd.INF_COST
        assert cost[i] == want
    dg.close()
    mz = ddp_amd.HostGraph.grid_maze(256, 200, 300, 35)
    dg = ddp_amd.DeviceGraph(ctx, mz)
    dg.set_grid(256)
    ms, mt = mz.reachable_pairs(24, 36)
    say("piles maze", block)
    cost, _, st = ddp_amd.astar_batch(ctx, dg, ms, mt, heur=ddp_amd.H_MANHATTAN, block=block,
                                      engine=ddp_amd.ENGINE_PILES)
    for i in range(24):
        dist, _ = pyoracle.dijkstra(mz.off, mz.col, mz.w, int(ms[i]), int(mt[i]))
        assert cost[i] == dist[mt[i]]
    # a maze search runs through many windows: refills are counted, each reads
    # the far pile twice (minimum pass, then the split)
    assert st["pile_refills"] >= 1 and st["pile_scanned"] % 2 == 0
    dg.close()


@pytest.mark.parametrize("block", [64, 256, 1024])
def test_batch_power_law_hub_rows(ctx, block):
    """R-MAT graph (hub rows of thousands of edges, Dijkstra order): rounds
    whose edges span many windows and rows longer than the owner-map runs
    take the binary-search path, smaller rounds the owner hints; both
    engines, both numberings, costs equal to Dijkstra and paths real."""
    hg = ddp_amd.HostGraph.rmat(13, 16 << 13, 0.57, 0.19, 0.19, 1, 100, 41)
    src, dst = hg.hub_pairs(48, 42)
    maxdeg = int(np.diff(np.append(hg.off, hg.E)).max())
    assert maxdeg > 1000  # rows far longer than the 32-edge owner runs
    want = []
    for i in range(48):
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[i]), int(dst[i]))
        want.append(int(dist[dst[i]]) if dist[dst[i]] >= 0 else ddp_amd.INF_COST)
    for reorder in (False, True):
        dg = ddp_amd.DeviceGraph(ctx, hg, reorder=reorder)
        say("power law", block, reorder, "maxdeg", maxdeg)
        cost, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, block=block,
                                              path_cap=8192, engine=ddp_amd.ENGINE_PILES)
        assert st["errors"] == 0
        assert [int(c) for c in cost] == want
        for i in range(0, 48, 7):
            if want[i] != ddp_amd.INF_COST:
                check_path(hg.off, hg.col, hg.w, paths[i], int(src[i]), int(dst[i]), want[i])
        if block <= 256:
            heaps, _, _ = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, block=block,
                                              engine=ddp_amd.ENGINE_HEAPS)
            assert [int(c) for c in heaps] == want
        dg.close()


def test_batch_piles_overflow_is_reported(ctx):
    """Queue memory too small for a query: the query is flagged, never
    answered wrongly; the others stay exact."""
    hg = ddp_amd.HostGraph.geometric(30000, 10, 100000, 37)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src, dst = hg.geo_pairs(32, 100, 100000, 38)
    b = ddp_amd.Batch(ctx, dg, 32, heur=ddp_amd.H_EUCLID, block=64, heap_cap=16, engine=ddp_amd.ENGINE_PILES)
    b.upload(src, dst)
    b.launch()
    st = b.wait()
    out = b.download()
    b.close()
    say("piles overflow", st)
    assert st["errors"] > 0
    for i in range(32):
        if out["qerr"][i] == 0:
            dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[i]), int(dst[i]))
            assert out["cost"][i] == dist[dst[i]]


def test_batch_piles_rebuild_after_overflow(ctx):
    """One expansion overflows both piles: hub 1 has 5,000 leaf edges of
    weight 10^6 after its first edge 1 -> 2 -> 3 (the target).  The dropped
    pushes trigger a rebuild of the far pile from the touched log, which keeps
    only nodes with f below F(t) = 3; the answer stays exact."""
    nleaf = 5000
    N = 4 + nleaf
    off = np.zeros(N, np.int32)
    col = [1, 2] + list(range(4, N)) + [3]
    w = [1, 1] + [1_000_000] * nleaf + [1]
    off[1], off[2] = 1, 2 + nleaf
    off[3:] = 3 + nleaf
    hg = ddp_amd.HostGraph(off, np.array(col, np.int32), np.array(w, np.uint32))
    dg = ddp_amd.DeviceGraph(ctx, hg)
    # block 64, heap_cap 64: 4,096 entries = near ring 1,024 + two far piles of 1,536
    b = ddp_amd.Batch(ctx, dg, 2, heur=ddp_amd.H_ZERO, block=64, heap_cap=64, engine=ddp_amd.ENGINE_PILES,
                      path_cap=8)
    b.upload(np.array([0, 0], np.int32), np.array([3, 2], np.int32))
    b.launch()
    st = b.wait()
    out = b.download(paths=True)
    b.close()
    say("piles rebuild", st)
    assert st["errors"] == 0 and st["pile_rebuilds"] >= 1
    assert out["cost"].tolist() == [3, 2]
    assert out["paths"][0].tolist() == [0, 1, 2, 3]


@pytest.mark.parametrize("engine,block", [(ddp_amd.ENGINE_PILES, 64), (ddp_amd.ENGINE_PILES, 1024),
                                          (ddp_amd.ENGINE_HEAPS, 64)])
def test_batch_edge_cases(ctx, engine, block):
    """Batches with s == t, unreachable targets, out-of-range ids (flagged,
    not answered), a one-node graph without edges, and more queries than
    slots, answered twice from the same upload."""
    one = ddp_amd.HostGraph(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint32))
    d1 = ddp_amd.DeviceGraph(ctx, one)
    b = ddp_amd.Batch(ctx, d1, 2, heur=ddp_amd.H_ZERO, block=block, engine=engine, path_cap=8)
    b.upload(np.array([0, 0], np.int32), np.array([0, 0], np.int32))
    b.launch()
    b.wait()
    out = b.download(paths=True)
    b.close()
    d1.close()
    assert list(out["cost"]) == [0, 0] and list(out["qerr"]) == [0, 0]
    assert [list(p) for p in out["paths"]] == [[0], [0]]
    # chain 0 -> 1 -> 2 plus isolated node 3
    off = np.array([0, 1, 2, 2], np.int32)
    col = np.array([1, 2], np.int32)
    w = np.array([5, 7], np.uint32)
    hg = ddp_amd.HostGraph(off, col, w)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    src = np.array([0, 2, 0, 3, 1, -1, 0, 0] * 40, np.int32)
    dst = np.array([2, 0, 3, 3, 1, 2, 7, 1] * 40, np.int32)
    want = [12, ddp_amd.INF_COST, ddp_amd.INF_COST, 0, 0, None, None, 5] * 40
    b = ddp_amd.Batch(ctx, dg, len(src), heur=ddp_amd.H_ZERO, block=block, engine=engine, path_cap=8, slots=4)
    b.upload(src, dst)
    for rep in range(2):
        b.launch()
        st = b.wait()
        out = b.download(paths=True)
        say("edge cases", engine, block, rep, st["errors"])
        assert st["errors"] == 80  # the two out-of-range queries of each group of eight
        for i, wv in enumerate(want):
            if wv is None:
                assert out["qerr"][i] & 4, i
            else:
                assert out["qerr"][i] == 0 and out["cost"][i] == wv, (i, out["cost"][i], wv)
        assert list(out["paths"][0]) == [0, 1, 2] and len(out["paths"][1]) == 0
    b.close()
    dg.close()


def test_batch_config4_full_size(ctx):
    """The C4 batch at full size (2,000,000-node geometric graph, 4,096 queries
    of log-uniform separation, Euclidean heuristic, relabelled device graph):
    - every path is a chain of real edges from s to t whose weights sum to
      the returned cost (vectorised over all 4,096 paths);
    - the pile engine (1,024-lane slots) and the heap engine (512-lane slots)
      return the same 4,096 costs;
    - 64 uniformly drawn queries match the oracle's restatement of
      A_star_cpu.cpp and 8 match a plain Dijkstra."""
    hg = ddp_amd.HostGraph.geometric(2000000, 10, 100000, 4)
    src, dst = hg.geo_pairs(4096, 100, 100000, 4)
    dg = ddp_amd.DeviceGraph(ctx, hg, reorder=True)
    say("c4 full: piles")
    cost, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=1024, path_cap=4096,
                                          engine=ddp_amd.ENGINE_PILES)
    assert st["errors"] == 0
    say("c4 full: heaps")
    heaps, _, st_h = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=512,
                                         engine=ddp_amd.ENGINE_HEAPS)
    assert st_h["errors"] == 0
    assert np.array_equal(cost, heaps)
    dg.close()
    say("c4 full: paths")
    N = hg.N
    o = np.append(hg.off, hg.E).astype(np.int64)
    u = np.repeat(np.arange(N, dtype=np.int64), np.diff(o))
    key = u * N + hg.col.astype(np.int64)
    order = np.lexsort((hg.w, key))  # by (u, v), the lightest parallel edge first
    ks, ws = key[order], hg.w[order].astype(np.int64)
    found = cost != ddp_amd.INF_COST
    say("c4 full: reachable", int(found.sum()), "of", len(src))
    assert found.mean() > 0.95  # the graph's giant component holds nearly every node
    assert all(len(paths[i]) == 0 for i in np.flatnonzero(~found))
    live = np.flatnonzero(found)
    assert all(paths[i][0] == src[i] and paths[i][-1] == dst[i] for i in live)
    pid = np.concatenate([np.full(len(paths[i]) - 1, i, np.int64) for i in live])
    hop = np.concatenate([paths[i][:-1].astype(np.int64) * N + paths[i][1:] for i in live])
    at = np.minimum(np.searchsorted(ks, hop), len(ks) - 1)
    assert (ks[at] == hop).all(), "a path step is not an edge"
    sums = np.bincount(pid, weights=ws[at], minlength=len(paths)).astype(np.int64)
    assert np.array_equal(sums[live], cost[live].astype(np.int64))
    say("c4 full: oracle sample")
    rng = np.random.default_rng(12)
    pick = np.sort(rng.choice(len(src), size=64, replace=False))
    got, _, _ = pyoracle.astar_batch_euclid(hg.off, hg.col, hg.w, hg.xy, src[pick], dst[pick], 8)
    assert np.array_equal(got, cost[pick])
    for i in pick[:8]:
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[i]), int(dst[i]))
        want = int(dist[dst[i]]) if dist[dst[i]] >= 0 else ddp_amd.INF_COST
        assert cost[i] == want, (i, cost[i], want)


def test_astar_config2_full_size(ctx):
    """C2 at full size: the 4096x4096 maze (p_block 0.30, seed 2), corner to
    corner, Manhattan heuristic, in every mode and K the benchmark sweeps
    (workgroup K = 512 / 1,024, piles K = 1,024, bidirectional K = 256 / 512 /
    1,024, grid K = 512 / 2,048 / 8,192): each returns the cost of the oracle's
    indexed-heap A* and a real path of that cost."""
    W = 4096
    hg = ddp_amd.HostGraph.grid_maze(W, W, 300, 2)
    s, t = 0, W * W - 1
    xs = np.arange(W * W, dtype=np.int64)
    h = ((W - 1 - xs // W) + (W - 1 - xs % W)).astype(np.int32)
    found, want, _ = pyoracle.astar_indexed(hg.off, hg.col, hg.w, h, s, t)
    assert found and want >= 2 * (W - 1)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True)  # transposed rows for the bidirectional mode
    dg.set_grid(W)
    for K, mode in ((512, ddp_amd.MODE_WORKGROUP), (1024, ddp_amd.MODE_WORKGROUP), (1024, ddp_amd.MODE_PILES),
                    (256, ddp_amd.MODE_BIDI), (512, ddp_amd.MODE_BIDI), (1024, ddp_amd.MODE_BIDI),
                    (512, ddp_amd.MODE_GRID), (2048, ddp_amd.MODE_GRID), (8192, ddp_amd.MODE_GRID)):
        say("c2 full", K, mode)
        r = ddp_amd.astar(ctx, dg, s, t, heur=ddp_amd.H_MANHATTAN, K=K, mode=mode)
        assert r.cost == want
        check_path(hg.off, hg.col, hg.w, r.path, s, t, want)
    dg.close()


def test_batch_config5_full_size(ctx):
    """C5's graph at full size (R-MAT scale 23: 8,388,608 nodes, 134,217,728
    edges, a/b/c/d = 0.57/0.19/0.19/0.05, w in [1,1000], seed 5), h = 0, 8
    hub pairs in one batch on the pile engine: every cost matches a plain
    Dijkstra and every path is a real path of that cost."""
    hg = ddp_amd.HostGraph.rmat(23, 16 << 23, 0.57, 0.19, 0.19, 1, 1000, 5)
    src, dst = hg.hub_pairs(8, 5)
    dg = ddp_amd.DeviceGraph(ctx, hg, reorder=True)
    say("c5 full: batch")
    cost, paths, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, block=1024, path_cap=4096,
                                          engine=ddp_amd.ENGINE_PILES)
    assert st["errors"] == 0
    dg.close()
    for i in range(len(src)):
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(src[i]), int(dst[i]))
        want = int(dist[dst[i]]) if dist[dst[i]] >= 0 else ddp_amd.INF_COST
        say("c5 full: query", i, cost[i], want)
        assert cost[i] == want
        if want != ddp_amd.INF_COST:
            check_path(hg.off, hg.col, hg.w, paths[i], int(src[i]), int(dst[i]), want)
