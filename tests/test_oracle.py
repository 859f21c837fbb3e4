"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

* golden vectors printed by the reference's own A_star_cpu.cpp
  (tests/golden/astar_cpu_ref.json), ties and the row-end quirk included;
* live runs of the reference's tools compiled into oracle/_ref (A_star_cpu,
  GenerateCSR, genOpCSR, genCSRwholegraph) when they are present;
* brute force (Floyd-Warshall) and closed forms for Dijkstra and BFS.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import full_offsets, random_graph
from oracle import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "astar_cpu_ref.json")


def floyd(off, col, w, unit=False):
    N = len(off)
    INF = np.iinfo(np.int64).max // 4
    D = np.full((N, N), INF, np.int64)
    np.fill_diagonal(D, 0)
    o = full_offsets(off, len(col))
    for u in range(N):
        for j in range(o[u], o[u + 1]):
            c = col[j]
            if c >= 0:
                D[u, c] = min(D[u, c], 1 if unit else int(w[j]))
    for k in range(N):
        D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
    D[D >= INF] = -1
    return D


def test_astar_matches_reference_golden_vectors():
    data = json.load(open(GOLDEN))
    assert len(data["cases"]) >= 100
    for c in data["cases"]:
        off, col = np.array(c["off"], np.int32), np.array(c["col"], np.int32)
        w, h = np.array(c["w"], np.uint32), np.array(c["h"], np.int32)
        found, cost, parent, _ = pyoracle.astar_ref(off, col, w, h, c["s"], c["t"], quirk=True)
        assert found == c["found"], c
        if found:
            assert cost == c["cost"]
            assert pyoracle.chain_from_parent(parent, c["t"]) == c["path_t_to_s"]


@pytest.mark.skipif(not pyoracle.ref_tool("a_star_cpu"), reason="oracle/_ref/a_star_cpu not built")
def test_astar_matches_reference_binary_live():
    rng = np.random.default_rng(7)
    with tempfile.TemporaryDirectory() as d:
        for k in range(60):
            N = int(rng.integers(2, 60))
            off, col, w = random_graph(rng, N, max_deg=4, wmax=int(rng.choice([1, 2, 5, 100])))
            s, t = int(rng.integers(0, N)), int(rng.integers(0, N))
            h = np.zeros(N, np.int32)
            if k % 2:
                roff, rcol, rw = pyoracle.transpose(off, col, w)
                hops, _ = pyoracle.bfs_hops(roff, rcol, t)
                h = np.where(hops < 0, -1, hops * int(w.min(initial=1))).astype(np.int32)
            pyoracle.write_graph_txt(os.path.join(d, "graph.txt"), off, col, w)
            pyoracle.write_hx_txt(os.path.join(d, "Hx.txt"), h)
            found, cost, path = pyoracle.run_ref_astar(d, s, t)
            f2, c2, parent, _ = pyoracle.astar_ref(off, col, w, h, s, t, quirk=True)
            assert found == f2
            if found:
                assert cost == c2
                assert pyoracle.chain_from_parent(parent, t) == path


def test_dijkstra_matches_floyd_warshall():
    rng = np.random.default_rng(11)
    for _ in range(80):
        N = int(rng.integers(1, 25))
        off, col, w = random_graph(rng, N, max_deg=4, wmax=9, tombstones=0.1)
        D = floyd(off, col, w)
        for s in range(N):
            dist, _ = pyoracle.dijkstra(off, col, w, s)
            assert np.array_equal(dist, D[s])


def test_dijkstra_closed_forms():
    # open W x H grid with unit weights: distance = Manhattan distance
    Wd, H = 7, 5
    us, vs = [], []
    for r in range(H):
        for c in range(Wd):
            v = r * Wd + c
            for rr, cc in ((r - 1, c), (r + 1, c), (r, c + 1), (r, c - 1)):
                if 0 <= rr < H and 0 <= cc < Wd:
                    us.append(v)
                    vs.append(rr * Wd + cc)
    off, col, w = pyoracle.csr_from_edges(us, vs, np.ones(len(us), np.uint32))
    dist, _ = pyoracle.dijkstra(off, col, w, 0)
    for v in range(Wd * H):
        assert dist[v] == (v // Wd) + (v % Wd)
    # chain 0 -> 1 -> ... -> n-1 with weights 1..n-1: prefix sums; reverse unreachable
    n = 9
    off, col, w = pyoracle.csr_from_edges(list(range(n - 1)), list(range(1, n)), np.arange(1, n, dtype=np.uint32))
    dist, _ = pyoracle.dijkstra(off, col, w, 0)
    assert dist.tolist() == [k * (k + 1) // 2 for k in range(n)]
    dist, _ = pyoracle.dijkstra(off, col, w, n - 1)
    assert dist.tolist() == [-1] * (n - 1) + [0]


def test_astar_without_quirk_is_dijkstra_with_consistent_h():
    rng = np.random.default_rng(3)
    for _ in range(150):
        N = int(rng.integers(2, 50))
        off, col, w = random_graph(rng, N, max_deg=5, wmax=30)
        s, t = int(rng.integers(0, N)), int(rng.integers(0, N))
        roff, rcol, rw = pyoracle.transpose(off, col, w)
        hops, _ = pyoracle.bfs_hops(roff, rcol, t)
        h = np.where(hops < 0, -1, hops * int(w.min(initial=1))).astype(np.int32)
        dist, _ = pyoracle.dijkstra(off, col, w, s)
        found, cost, parent, _ = pyoracle.astar_ref(off, col, w, h, s, t, quirk=False)
        assert found == (dist[t] >= 0)
        if found:
            assert cost == dist[t] + h[t]


def test_indexed_heap_variant_matches_faithful_astar():
    rng = np.random.default_rng(17)
    for _ in range(150):
        N = int(rng.integers(2, 60))
        off, col, w = random_graph(rng, N, max_deg=5, wmax=int(rng.choice([1, 3, 40])))
        s, t = int(rng.integers(0, N)), int(rng.integers(0, N))
        roff, rcol, rw = pyoracle.transpose(off, col, w)
        hops, _ = pyoracle.bfs_hops(roff, rcol, t)
        h = np.where(hops < 0, -1, hops * int(w.min(initial=1))).astype(np.int32)
        f1, c1, _, _ = pyoracle.astar_ref(off, col, w, h, s, t, quirk=False)
        f2, c2, _ = pyoracle.astar_indexed(off, col, w, h, s, t)
        assert f1 == f2 and (not f1 or c1 == c2)


def test_row_end_quirk_drops_only_the_last_edge():
    off = np.array([0, 1, 2], np.int32)
    col = np.array([2, 1, 0], np.int32)
    w = np.array([5, 1, 7], np.uint32)
    h = np.zeros(3, np.int32)
    assert pyoracle.astar_ref(off, col, w, h, 2, 0, quirk=True)[0] is False
    f, c, _, _ = pyoracle.astar_ref(off, col, w, h, 2, 0, quirk=False)
    assert f and c == 7


def test_bfs_matches_brute_force_hops():
    rng = np.random.default_rng(5)
    for _ in range(60):
        N = int(rng.integers(1, 30))
        off, col, w = random_graph(rng, N, max_deg=3, tombstones=0.1)
        D = floyd(off, col, w, unit=True)
        for t in range(N):
            hops, levels = pyoracle.bfs_hops(off, col, t)
            assert np.array_equal(hops, D[t])
            assert levels == int(D[t].max()) + 1


def test_transpose_and_csr_closed_forms():
    # star: 0 -> 1..5 ; transpose has rows 1..5 -> 0
    off, col, w = pyoracle.csr_from_edges([0] * 5, [1, 2, 3, 4, 5], [9, 8, 7, 6, 5])
    roff, rcol, rw = pyoracle.transpose(off, col, w)
    assert roff.tolist() == [0, 0, 1, 2, 3, 4] and rcol.tolist() == [0] * 5 and rw.tolist() == [9, 8, 7, 6, 5]
    # applying a batch twice in opposite directions restores edge multisets
    o2, c2, w2, nd = pyoracle.apply_batch(off, col, w, [(0, 0, 3, 0), (1, 0, 3, 4)])
    assert nd == 1 and c2.tolist() == [1, 2, 4, 5, 3] and w2.tolist() == [9, 8, 6, 5, 4]
    o3, c3, w3, nd = pyoracle.apply_batch(o2, c2, w2, [(0, 0, 9, 0)])  # absent: no-op
    assert nd == 0 and c3.tolist() == c2.tolist()


def _run(exe, cwd, stdin=""):
    return subprocess.run([exe], input=stdin, capture_output=True, text=True, cwd=cwd, timeout=60, check=True).stdout


@pytest.mark.skipif(not pyoracle.ref_tool("gen_csr"), reason="oracle/_ref/gen_csr not built")
def test_csr_from_edges_matches_reference_GenerateCSR():
    rng = np.random.default_rng(9)
    with tempfile.TemporaryDirectory() as d:
        for _ in range(10):
            M = int(rng.integers(1, 200))
            us = rng.integers(0, 40, M)
            vs = rng.integers(0, 40, M)
            ws = rng.integers(1, 100, M)
            txt = "".join(f"{a} {b} {c}\n" for a, b, c in zip(us, vs, ws))
            out = _run(pyoracle.ref_tool("gen_csr"), d, txt)
            off, col, w = pyoracle.csr_from_edges(us, vs, ws)
            p = os.path.join(d, "mine.txt")
            pyoracle.write_graph_txt(p, off, col, w)
            assert open(p).read() == out


@pytest.mark.skipif(not pyoracle.ref_tool("gen_op_csr"), reason="oracle/_ref/gen_op_csr not built")
def test_transpose_matches_reference_genOpCSR():
    rng = np.random.default_rng(10)
    with tempfile.TemporaryDirectory() as d:
        for _ in range(10):
            N = int(rng.integers(2, 60))
            off, col, w = random_graph(rng, N, max_deg=5, wmax=50)
            pyoracle.write_graph_txt(os.path.join(d, "graph.txt"), off, col, w)
            _run(pyoracle.ref_tool("gen_op_csr"), d)
            ref = open(os.path.join(d, "graph_op.txt")).read()
            roff, rcol, rw = pyoracle.transpose(off, col, w)
            p = os.path.join(d, "mine.txt")
            pyoracle.write_graph_txt(p, roff, rcol, rw)
            assert open(p).read() == ref


@pytest.mark.skipif(not pyoracle.ref_tool("gen_csr_whole"), reason="oracle/_ref/gen_csr_whole not built")
def test_apply_batch_matches_reference_genCSRwholegraph():
    rng = np.random.default_rng(12)
    with tempfile.TemporaryDirectory() as d:
        for _ in range(10):
            N = int(rng.integers(3, 40))
            off, col, w = random_graph(rng, N, max_deg=4, wmax=50)
            pyoracle.write_graph_txt(os.path.join(d, "graph.txt"), off, col, w)
            batches = []
            cur = (off, col, w)
            for b in range(int(rng.integers(1, 4))):
                lines = []
                co = full_offsets(cur[0], len(cur[1]))
                for _ in range(int(rng.integers(0, 6))):  # deletes first (some absent)
                    u = int(rng.integers(0, N))
                    if co[u + 1] > co[u] and rng.random() < 0.8:
                        v = int(cur[1][rng.integers(co[u], co[u + 1])])
                    else:
                        v = int(rng.integers(0, N))
                    lines.append((0, u, v, 0))
                for _ in range(int(rng.integers(0, 6))):
                    lines.append((1, int(rng.integers(0, N)), int(rng.integers(0, N)), int(rng.integers(1, 50))))
                batches.append(lines)
                o2, c2, w2, _ = pyoracle.apply_batch(*cur, lines)
                cur = (o2, c2, w2)
            with open(os.path.join(d, "Updates.txt"), "w") as f:
                for lines in batches:
                    f.write(f"{len(lines)}\n" + "".join(f"{a} {b} {c} {e}\n" for a, b, c, e in lines))
            _run(pyoracle.ref_tool("gen_csr_whole"), d)
            roff, rcol, rw = pyoracle.read_graph_txt(os.path.join(d, "graph_cg.txt"))
            # graph_cg.txt keeps deleted edges as -1 entries; drop them
            ro = full_offsets(roff, len(rcol))
            keep = rcol >= 0
            newoff = np.array([keep[:ro[u]].sum() for u in range(N)], np.int32)
            assert newoff.tolist() == cur[0].tolist()
            assert rcol[keep].tolist() == cur[1].tolist()
            assert rw[keep].tolist() == cur[2].tolist()


def test_threaded_euclid_batch_matches_faithful_astar():
    """orc_astar_batch_euclid (thread pool, h computed on demand) gives the
    same costs as orc_astar_ref run with a per-target floor(Euclid) array
    computed here independently, and as Dijkstra (the heuristic is consistent:
    edge weights are ceil(Euclid))."""
    rng = np.random.default_rng(41)
    n = 600
    xy = rng.integers(0, 5000, size=(n, 2)).astype(np.int64)
    us, vs, ws = [], [], []
    for u in range(n):
        d2 = ((xy - xy[u]) ** 2).sum(1)
        for v in np.argsort(d2)[1:5]:
            r = int(np.ceil(np.sqrt(float(d2[v]))))
            while (r - 1) * (r - 1) >= d2[v]:
                r -= 1
            us.append(u); vs.append(int(v)); ws.append(max(1, r))
    off, col, w = pyoracle.csr_from_edges(np.array(us), np.array(vs), np.array(ws))
    src = rng.integers(0, n, 40).astype(np.int32)
    dst = rng.integers(0, n, 40).astype(np.int32)
    costs, ex, started = pyoracle.astar_batch_euclid(off, col, w, xy.reshape(-1).astype(np.int32), src, dst, 4)
    assert started == 4 and ex > 0
    for i in range(40):
        d2 = ((xy - xy[dst[i]]) ** 2).sum(1)
        h = np.floor(np.sqrt(d2.astype(np.float64))).astype(np.int64)
        h -= (h * h > d2)
        h += ((h + 1) * (h + 1) <= d2)
        found, cost, _, _ = pyoracle.astar_ref(off, col, w, h.astype(np.int32), int(src[i]), int(dst[i]), quirk=False)
        dist, _ = pyoracle.dijkstra(off, col, w, int(src[i]))
        if found:
            assert costs[i] == cost == dist[dst[i]]
        else:
            assert costs[i] == 2 ** 31 - 1 and dist[dst[i]] < 0
