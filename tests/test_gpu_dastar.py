"""GPU parity of DA*: after every update batch the repaired cost must equal a
from-scratch search on the updated graph (oracle Dijkstra on the oracle's own
application of the batch), the returned path must be a path of that cost in
the updated graph, and the library's graph version must equal the oracle's
edge for edge."""
import sys

import numpy as np
import pytest

import ddp_amd
from helpers import full_offsets, path_cost
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def union_h(hg, batches, t):
    """Reverse-BFS hops x min weight on base + every insert: admissible for
    every graph version of the run."""
    o = full_offsets(hg.off, hg.E)
    src = list(np.repeat(np.arange(hg.N), np.diff(o)))
    dst = list(hg.col)
    ws = list(hg.w)
    for b in batches:
        for f, u, v, w in b:
            if f == 1:
                src.append(u)
                dst.append(v)
                ws.append(w)
    src, dst, ws = np.array(src), np.array(dst), np.array(ws, np.uint32)
    order = np.argsort(src, kind="stable")
    off = np.searchsorted(src[order], np.arange(hg.N)).astype(np.int32)
    roff, rcol, rw = pyoracle.transpose(off, dst[order], ws[order])
    hops, _ = pyoracle.bfs_hops(roff, rcol, t)
    minw = int(ws.min())
    return np.where(hops < 0, -1, hops * minw).astype(np.int32)


def run_batches(ctx, hg, batches, s, t, h, K=1024):
    da = ddp_amd.DAStar(ctx, hg, s, t, h=h, K=K)
    rc = ddp_amd.DAStar(ctx, hg, s, t, h=h, K=K)
    cur = (hg.off, hg.col, hg.w)
    d0, _ = pyoracle.dijkstra(*cur, s, t)
    assert da.cost == (d0[t] if d0[t] >= 0 else ddp_amd.INF_COST)
    for bi, lines in enumerate(batches):
        say("batch", bi, "lines", len(lines))
        cost, st = da.apply(lines, mode=ddp_amd.DAStar.REPAIR)
        cost_rc, _ = rc.apply(lines, mode=ddp_amd.DAStar.RECOMPUTE)
        o2, c2, w2, nd = pyoracle.apply_batch(*cur, lines)
        cur = (o2, c2, w2)
        assert st["deletes_effective"] == nd
        g = da.graph()
        assert g.off.tolist() == o2.tolist() and g.col.tolist() == c2.tolist() and g.w.tolist() == w2.tolist()
        dist, _ = pyoracle.dijkstra(*cur, s, t)
        want = int(dist[t]) if dist[t] >= 0 else ddp_amd.INF_COST
        say("  cost", cost, "want", want, "stats", st)
        assert cost == cost_rc == want
        if want != ddp_amd.INF_COST:
            p = da.path()
            assert p[0] == s and p[-1] == t
            assert path_cost(*cur, p) == want
    da.close()
    rc.close()


def test_dastar_random_batches(ctx):
    hg = ddp_amd.HostGraph.random(20000, 8, 1, 1000, 3)
    upd = hg.gen_updates(8, 300, 300, 1, 1000, 4)
    src, dst = hg.reachable_pairs(1, 5)
    s, t = int(src[0]), int(dst[0])
    h = union_h(hg, upd.batches, t)
    run_batches(ctx, hg, upd.batches, s, t, h)


def test_dastar_zero_heuristic_and_small_K(ctx):
    hg = ddp_amd.HostGraph.random(5000, 4, 1, 50, 7)
    upd = hg.gen_updates(6, 150, 200, 1, 50, 8)
    src, dst = hg.reachable_pairs(1, 9)
    run_batches(ctx, hg, upd.batches, int(src[0]), int(dst[0]), None, K=64)


def test_dastar_deletes_on_the_path(ctx):
    """Each batch deletes edges of the current shortest path (forcing subtree
    invalidation and re-pulls) and inserts a few random shortcuts."""
    rng = np.random.default_rng(11)
    hg = ddp_amd.HostGraph.random(8000, 5, 1, 100, 12)
    src, dst = hg.reachable_pairs(1, 13)
    s, t = int(src[0]), int(dst[0])
    batches = []
    cur = (hg.off, hg.col, hg.w)
    for b in range(10):
        dist, _ = pyoracle.dijkstra(*cur, s)
        lines = []
        if dist[t] >= 0:
            # rebuild one shortest path from the distances
            o = full_offsets(cur[0], len(cur[1]))
            path = [t]
            while path[-1] != s:
                v = path[-1]
                for u in range(len(cur[0])):
                    hit = False
                    for j in range(o[u], o[u + 1]):
                        if cur[1][j] == v and dist[u] >= 0 and dist[u] + int(cur[2][j]) == dist[v]:
                            path.append(u)
                            hit = True
                            break
                    if hit:
                        break
            path = path[::-1]
            for k in rng.choice(len(path) - 1, size=min(3, len(path) - 1), replace=False):
                lines.append((0, int(path[k]), int(path[k + 1]), 0))
        lines.append((0, 1, 2, 0))  # most likely absent: a no-op delete
        for _ in range(20):
            lines.append((1, int(rng.integers(0, hg.N)), int(rng.integers(0, hg.N)), int(rng.integers(1, 100))))
        batches.append(lines)
        cur = pyoracle.apply_batch(*cur, lines)[:3]
    run_batches(ctx, hg, batches, s, t, None)


def test_dastar_target_cut_off_and_reconnected(ctx):
    # chain 0 -> 1 -> 2 -> 3 plus a detour 0 -> 4 -> 3
    off = np.array([0, 2, 3, 4, 4], np.int32)
    col = np.array([1, 4, 2, 3, 3], np.int32)
    w = np.array([1, 5, 1, 1, 5], np.uint32)
    hg = ddp_amd.HostGraph(off, col, w)
    batches = [
        [(0, 1, 2, 0)],                    # forces the detour: cost 10
        [(0, 0, 4, 0)],                    # t unreachable
        [(1, 2, 3, 7), (1, 1, 2, 1)],      # reconnected through a new 1->2: cost 3
        [(1, 0, 3, 4)],                    # a direct 0->3 of weight 4 does not beat 3
        [(0, 2, 3, 0), (0, 2, 3, 0)],      # both parallel 2->3 edges gone: direct edge, cost 4
    ]
    run_batches(ctx, hg, batches, 0, 3, None, K=64)


def test_alternative_paths_loop(ctx):
    """ddp_amd.alternative_paths: each cost equals Dijkstra after deleting the
    edge into t of every earlier path (the k-short/GA_merge.cu loop)."""
    hg = ddp_amd.HostGraph.random(5000, 5, 1, 100, 51)
    src, dst = hg.reachable_pairs(3, 52)
    for s, t in zip(src, dst):
        s, t = int(s), int(t)
        res = ddp_amd.alternative_paths(ctx, hg, s, t, 8, K=1024)
        cur = (hg.off, hg.col, hg.w)
        for k, (cost, path) in enumerate(res):
            if k:
                prev = res[k - 1][1]
                cur = pyoracle.apply_batch(*cur, [(0, int(prev[-2]), t, 0)])[:3]
            dist, _ = pyoracle.dijkstra(*cur, s, t)
            assert cost == (int(dist[t]) if dist[t] >= 0 else ddp_amd.INF_COST), (s, t, k)
            if cost != ddp_amd.INF_COST:
                assert path[0] == s and path[-1] == t and path_cost(*cur, path) == cost


def test_dastar_wide_and_narrow_subtree_levels(ctx):
    """Deleting s->a invalidates a's whole subtree: a, a level of 20,000
    leaves (spread over every workgroup of the walk) and a 50-node chain below
    one leaf (50 one-node levels, each ended by the walk's device-wide
    barrier).  The leaf is then re-pulled through a costlier detour; a second
    batch puts s->a back."""
    nl, nc = 20000, 50
    a, leaf0, chain0 = 1, 2, 2 + nl
    b = chain0 + nc
    n = b + 1
    t = chain0 + nc - 1
    rows = [[] for _ in range(n)]
    rows[0] = [(a, 1), (b, 10)]
    rows[a] = [(leaf0 + i, 1) for i in range(nl)]
    rows[leaf0] = [(chain0, 1)]
    for k in range(nc - 1):
        rows[chain0 + k] = [(chain0 + k + 1, 1)]
    rows[b] = [(leaf0, 1)]
    off = np.cumsum([0] + [len(r) for r in rows[:-1]]).astype(np.int32)
    col = np.array([v for r in rows for v, _ in r], np.int32)
    w = np.array([x for r in rows for _, x in r], np.uint32)
    hg = ddp_amd.HostGraph(off, col, w)
    da = ddp_amd.DAStar(ctx, hg, 0, t, h=None, K=1024)
    assert da.cost == 2 + nc
    cost, st = da.apply([(0, 0, a, 0)], mode=ddp_amd.DAStar.REPAIR)
    say("cut", cost, st)
    assert st["invalidated"] == 1 + nl + nc
    assert cost == 10 + 1 + nc
    p = da.path().tolist()
    assert p[:3] == [0, b, leaf0] and p[-1] == t and len(p) == nc + 3
    cost, st = da.apply([(1, 0, a, 1)], mode=ddp_amd.DAStar.REPAIR)
    say("restored", cost, st)
    assert cost == 2 + nc
    assert da.path().tolist()[:3] == [0, a, leaf0]
    da.close()


def test_dastar_many_roots_at_once(ctx):
    """One batch deletes 9,000 tree edges s->m_i: the first two levels
    (9,000 roots, then their 9,000 leaves) are wider than one workgroup's walk
    takes, the chain below leaf 0 is narrow.  The leaves are re-pulled
    through x (s->x weighs 5), the m_i become unreachable."""
    m = 9000
    nc = 40
    x = 1
    m0, leaf0, chain0 = 2, 2 + m, 2 + 2 * m
    n = chain0 + nc
    t = n - 1
    rows = [[] for _ in range(n)]
    rows[0] = [(x, 5)] + [(m0 + i, 1) for i in range(m)]
    rows[x] = [(leaf0 + i, 1) for i in range(m)]
    for i in range(m):
        rows[m0 + i] = [(leaf0 + i, 1)]
    rows[leaf0] = [(chain0, 1)]
    for k in range(nc - 1):
        rows[chain0 + k] = [(chain0 + k + 1, 1)]
    off = np.cumsum([0] + [len(r) for r in rows[:-1]]).astype(np.int32)
    col = np.array([v for r in rows for v, _ in r], np.int32)
    w = np.array([c for r in rows for _, c in r], np.uint32)
    hg = ddp_amd.HostGraph(off, col, w)
    da = ddp_amd.DAStar(ctx, hg, 0, t, h=None, K=1024)
    assert da.cost == 2 + nc
    cost, st = da.apply([(0, 0, m0 + i, 0) for i in range(m)], mode=ddp_amd.DAStar.REPAIR)
    say("cut", cost, st)
    assert st["roots"] == m
    assert st["invalidated"] == 2 * m + nc
    assert cost == 5 + 1 + nc
    p = da.path().tolist()
    assert p[:3] == [0, x, leaf0] and p[-1] == t and len(p) == nc + 3
    da.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", ["grid", "piles"])
def test_dastar_config3_full_size(ctx, mode):
    """C3 at full size (N = 1,000,000, out-degree 8, w in [1,1000], seed 3),
    all 20 of its batches of 10,000 inserts + 10,000 deletes at K = 16,384
    (the K the benchmark runs), h = reverse-BFS hops x min weight on the union
    of every version, the repairing session on the grid kernel or the
    one-workgroup pile engine (both benchmark sessions).  Each batch: repair
    == recompute.  Every 4th batch and the last: the cost equals the oracle's
    Dijkstra on the oracle's own merged graph and the path is a real path of
    that cost.  After the last: the library's graph equals the oracle's edge
    for edge."""
    nb = 20
    hg = ddp_amd.HostGraph.random(1000000, 8, 1, 1000, 3)
    upd = hg.gen_updates(nb, 10000, 10000, 1, 1000, 33).batches
    src, dst = hg.reachable_pairs(1, 3)
    s, t = int(src[0]), int(dst[0])
    o = full_offsets(hg.off, hg.E)
    us = [np.repeat(np.arange(hg.N, dtype=np.int64), np.diff(o))]
    vs, ws = [hg.col.astype(np.int64)], [hg.w.astype(np.int64)]
    for b in upd:
        a = np.array([l for l in b if l[0] == 1], np.int64)
        us.append(a[:, 1])
        vs.append(a[:, 2])
        ws.append(a[:, 3])
    us, vs, ws = np.concatenate(us), np.concatenate(vs), np.concatenate(ws)
    order = np.argsort(us, kind="stable")
    uoff = np.searchsorted(us[order], np.arange(hg.N)).astype(np.int32)
    roff, rcol, _ = pyoracle.transpose(uoff, vs[order].astype(np.int32), ws[order].astype(np.uint32))
    hops, _ = pyoracle.bfs_hops(roff, rcol, t)
    h = np.where(hops < 0, -1, hops * int(ws.min())).astype(np.int32)
    da = ddp_amd.DAStar(ctx, hg, s, t, h=h, K=16384,
                        mode=ddp_amd.MODE_PILES if mode == "piles" else ddp_amd.MODE_AUTO)
    rc = ddp_amd.DAStar(ctx, hg, s, t, h=h, K=16384)
    assert da.cost == rc.cost
    cur = (hg.off, hg.col, hg.w)
    for bi, lines in enumerate(upd):
        cost, st = da.apply(lines, mode=ddp_amd.DAStar.REPAIR)
        cost_rc, _ = rc.apply(lines, mode=ddp_amd.DAStar.RECOMPUTE)
        say("c3 full batch", bi, cost, cost_rc, st["invalidated"], st["iterations"])
        assert cost == cost_rc
        cur = pyoracle.apply_batch(*cur, lines)[:3]
        if bi % 4 == 0 or bi == nb - 1:
            dist, _ = pyoracle.dijkstra(*cur, s, t)
            want = int(dist[t]) if dist[t] >= 0 else ddp_amd.INF_COST
            assert cost == want, bi
            if want != ddp_amd.INF_COST:
                p = da.path()
                assert p[0] == s and p[-1] == t
                assert path_cost(*cur, p) == want
    g = da.graph()
    assert np.array_equal(g.off, cur[0]) and np.array_equal(g.col, cur[1]) and np.array_equal(g.w, cur[2])
    da.close()
    rc.close()


def test_dastar_bad_mode_changes_nothing(ctx):
    """An unknown mode is refused before the batch reaches the device: the
    next batch (same lines, valid mode) then matches Dijkstra on the graph
    with that batch applied once."""
    hg = ddp_amd.HostGraph.random(3000, 5, 1, 60, 71)
    src, dst = hg.reachable_pairs(1, 72)
    s, t = int(src[0]), int(dst[0])
    upd = hg.gen_updates(1, 200, 200, 1, 60, 73).batches[0]
    da = ddp_amd.DAStar(ctx, hg, s, t, K=1024)
    with pytest.raises(ddp_amd.DDPError):
        da.apply(upd, mode=7)
    cost, _ = da.apply(upd, mode=ddp_amd.DAStar.REPAIR)
    cur = pyoracle.apply_batch(hg.off, hg.col, hg.w, upd)[:3]
    d, _ = pyoracle.dijkstra(*cur, s, t)
    assert cost == (int(d[t]) if d[t] >= 0 else ddp_amd.INF_COST)
    g = da.graph()
    assert np.array_equal(g.col, cur[1])
    da.close()


def test_dastar_compaction(ctx):
    """Log compaction (merge_graph / mergeDiff, DA_star/DA_dyn_insert_del_nocycle.cu:940-980,
    2236-2284): with a small threshold the log is folded into the rows every
    few batches, and an explicit compaction runs in between.  Every batch:
    repair == recompute == the oracle's Dijkstra on the oracle-applied graph,
    the path is real, and the exported graph equals the oracle's edge for edge
    (GA_merge.cu's merged rows) whether or not the batch compacted."""
    hg = ddp_amd.HostGraph.random(5000, 5, 1, 80, 81)
    src, dst = hg.reachable_pairs(1, 82)
    s, t = int(src[0]), int(dst[0])
    upd = hg.gen_updates(24, 120, 120, 1, 80, 83).batches
    da = ddp_amd.DAStar(ctx, hg, s, t, K=1024)
    rc = ddp_amd.DAStar(ctx, hg, s, t, K=1024)
    da.set_compaction(300)  # every third batch or so
    cur = (hg.off, hg.col, hg.w)
    compacted = 0
    for bi, lines in enumerate(upd):
        cost, st = da.apply(lines, mode=ddp_amd.DAStar.REPAIR)
        cost_rc, _ = rc.apply(lines, mode=ddp_amd.DAStar.RECOMPUTE)
        compacted += st["compacted"]
        cur = pyoracle.apply_batch(*cur, lines)[:3]
        d, _ = pyoracle.dijkstra(*cur, s, t)
        want = int(d[t]) if d[t] >= 0 else ddp_amd.INF_COST
        say("compaction batch", bi, cost, want, "compacted" if st["compacted"] else "")
        assert cost == cost_rc == want, bi
        if want != ddp_amd.INF_COST:
            p = da.path()
            assert p[0] == s and p[-1] == t and path_cost(*cur, p) == want
        g = da.graph()
        assert np.array_equal(g.off, cur[0]) and np.array_equal(g.col, cur[1]) and np.array_equal(g.w, cur[2]), bi
        if bi == 10:
            da.compact()  # explicit, mid-run
            g = da.graph()
            assert np.array_equal(g.col, cur[1]) and np.array_equal(g.w, cur[2])
    assert compacted >= 4
    da.close()
    rc.close()


@pytest.mark.parametrize("seed", [101, 102])
def test_dastar_pile_mode(ctx, seed):
    """DA* with every search on one workgroup's pile engine (DDP_ASTAR_PILES):
    the open set left by each search -- entries that could not beat f(t) kept
    aside -- is resumed after the next batch's repair.  Every batch: the same
    cost as the grid-mode session and the oracle's Dijkstra, a real path, and
    recompute in pile mode agrees; a compaction in the middle changes nothing."""
    hg = ddp_amd.HostGraph.random(20000, 6, 1, 100, seed)
    src, dst = hg.reachable_pairs(1, seed + 1)
    s, t = int(src[0]), int(dst[0])
    upd = hg.gen_updates(12, 400, 400, 1, 100, seed + 2).batches
    pl = ddp_amd.DAStar(ctx, hg, s, t, K=1024, mode=ddp_amd.MODE_PILES)
    gr = ddp_amd.DAStar(ctx, hg, s, t, K=1024)
    rc = ddp_amd.DAStar(ctx, hg, s, t, K=1024, mode=ddp_amd.MODE_PILES)
    assert pl.cost == gr.cost == rc.cost
    cur = (hg.off, hg.col, hg.w)
    for bi, lines in enumerate(upd):
        c_pl, st = pl.apply(lines, mode=ddp_amd.DAStar.REPAIR)
        c_gr, _ = gr.apply(lines, mode=ddp_amd.DAStar.REPAIR)
        c_rc, _ = rc.apply(lines, mode=ddp_amd.DAStar.RECOMPUTE)
        cur = pyoracle.apply_batch(*cur, lines)[:3]
        d, _ = pyoracle.dijkstra(*cur, s, t)
        want = int(d[t]) if d[t] >= 0 else ddp_amd.INF_COST
        say("pile DA* batch", bi, c_pl, c_gr, c_rc, want, "rounds", st["iterations"], "invalidated", st["invalidated"])
        assert c_pl == c_gr == c_rc == want, bi
        if want != ddp_amd.INF_COST:
            p = pl.path()
            assert p[0] == s and p[-1] == t and path_cost(*cur, p) == want
        if bi == 5:
            pl.compact()
    pl.close()
    gr.close()
    rc.close()
