"""The command-line drivers honour the reference's process contract: stdin
arguments, graph.txt / Hx.txt / Updates.txt in the working directory, the
reference's stdout grammar.  Results are checked against the oracle and the
reference's own A_star_cpu where it is built."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import ddp_amd
from helpers import path_cost
from oracle import pyoracle

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ddp_amd", "bin")


def run(exe, cwd, stdin, *args):
    print("run", exe, args, flush=True)
    sys.stdout.flush()
    return subprocess.run([os.path.join(BIN, exe), *args], input=stdin, capture_output=True, text=True, cwd=cwd,
                          timeout=120, check=True).stdout


def test_bfs_hx_and_ga_star_contract():
    hg = ddp_amd.HostGraph.random(3000, 4, 1, 30, 21)
    src, dst = hg.reachable_pairs(4, 22)
    with tempfile.TemporaryDirectory() as d:
        for s, t in zip(src, dst):
            s, t = int(s), int(t)
            # bfs_hx reads the transposed graph as graph.txt (GA_star/bfs_hx.cu:45)
            roff, rcol, rw = pyoracle.transpose(hg.off, hg.col, hg.w)
            pyoracle.write_graph_txt(os.path.join(d, "graph.txt"), roff, rcol, rw)
            out = run("bfs_hx", d, f"{t}\n")
            hops, _ = pyoracle.bfs_hops(roff, rcol, t)
            got = np.array(open(os.path.join(d, "Hx.txt")).read().split(), np.int32)
            assert got.tolist() == hops.tolist()
            assert f"unreachable {int((hops < 0).sum())} ," in out
            assert f"max-level: {int(hops.max())}" in out
            # ga_star on the forward graph with that heuristic
            hg.write_txt(os.path.join(d, "graph.txt"))
            out = run("ga_star", d, f"1024\n{s} {t}\n").splitlines()
            assert out[0] == "completed input"
            cost = int(out[1].split()[1])
            path_rev = [int(x) for x in out[2].split()]
            found, want, _, _ = pyoracle.astar_ref(hg.off, hg.col, hg.w, got, s, t, quirk=False)
            assert found and cost == want
            assert path_rev[0] == t and path_rev[-1] == s
            assert path_cost(hg.off, hg.col, hg.w, path_rev[::-1]) == cost
            if pyoracle.ref_tool("a_star_cpu"):
                f2, c2, _ = pyoracle.run_ref_astar(d, s, t)  # the reference's own CPU A*
                assert f2 and c2 >= cost  # it drops edge E-1; equal unless that edge is on every shortest path
                if hg.col[-1] not in path_rev:
                    assert c2 == cost


def test_da_star_contract():
    hg = ddp_amd.HostGraph.random(4000, 6, 1, 100, 31)
    upd = hg.gen_updates(4, 100, 100, 1, 100, 32)
    src, dst = hg.reachable_pairs(1, 33)
    s, t = int(src[0]), int(dst[0])
    with tempfile.TemporaryDirectory() as d:
        hg.write_txt(os.path.join(d, "graph.txt"))
        pyoracle.write_hx_txt(os.path.join(d, "Hx.txt"), np.zeros(hg.N, np.int32))
        upd.write_txt(os.path.join(d, "Updates.txt"))
        outs = [run("da_star", d, f"1024\n{s} {t}\n", *a) for a in ((), ("--recompute",))]
        costs = [[int(l.split()[-1]) for l in o.splitlines() if l.startswith("[OUT] Cost:")] for o in outs]
        assert costs[0] == costs[1] and len(costs[0]) == 1 + len(upd.batches)
        cur = (hg.off, hg.col, hg.w)
        want = []
        for b in [None] + upd.batches:
            if b is not None:
                cur = pyoracle.apply_batch(*cur, b)[:3]
            dist, _ = pyoracle.dijkstra(*cur, s, t)
            want.append(int(dist[t]) if dist[t] >= 0 else ddp_amd.INF_COST)
        assert costs[0] == want
        paths = [l for l in outs[0].splitlines() if l.startswith("[OUT] Path(in reverse):")]
        assert len(paths) == len(want)
        assert "[INFO] RUNTIME:" in outs[0]


def test_k_short_contract():
    """k_short (k-short/GA_merge.cu): after the first search, each further path
    deletes the edge into t on the current best path and re-solves; every cost
    equals Dijkstra on the graph with the same edges deleted (first live
    instance, as the oracle's batch application does)."""
    hg = ddp_amd.HostGraph.random(3000, 5, 1, 50, 41)
    src, dst = hg.reachable_pairs(1, 42)
    s, t = int(src[0]), int(dst[0])
    with tempfile.TemporaryDirectory() as d:
        hg.write_txt(os.path.join(d, "graph.txt"))
        pyoracle.write_hx_txt(os.path.join(d, "Hx.txt"), np.zeros(hg.N, np.int32))
        out = run("k_short", d, f"1024\n{s} {t}\n6\n").splitlines()
    costs = [int(l.split()[-1]) for l in out if l.startswith("[OUT] Cost:")]
    paths = [[] if "not found" in l else [int(x) for x in l.split(":", 1)[1].split()]
             for l in out if l.startswith("[OUT] Path(in reverse):")]
    assert len(costs) == len(paths) >= 2 and out[-1].startswith("[INFO] total run time")
    cur = (hg.off, hg.col, hg.w)
    for k, (cost, prev) in enumerate(zip(costs, [None] + paths[:-1])):
        if prev is not None:  # the previous path, t first: its first edge in reverse is parent(t) -> t
            cur = pyoracle.apply_batch(*cur, [(0, prev[1], prev[0], 0)])[:3]
        dist, _ = pyoracle.dijkstra(*cur, s, t)
        want = int(dist[t]) if dist[t] >= 0 else ddp_amd.INF_COST
        assert cost == want, k
        if cost != ddp_amd.INF_COST:
            assert paths[k][0] == t and paths[k][-1] == s
            assert path_cost(*cur, paths[k][::-1]) == cost
    assert costs == sorted(costs)  # each deletion can only lengthen the best path


def _quirk_graph(off, col, w):
    """The graph A_star_cpu.cpp searches: node N-1's row ends at E-1
    (A_star_cpu.cpp:169-172), so edge E-1 is dropped when it is in that row."""
    col = np.array(col, np.int32)
    if len(col) and off[-1] <= len(col) - 1:
        col[-1] = -1
    return col


def _unique_shortest_path(off, col, w, s, t):
    """True iff exactly one node sequence s..t has the shortest cost (so every
    correct program prints the same path)."""
    dist, _ = pyoracle.dijkstra(off, col, w, s)
    if dist[t] < 0:
        return True
    o = np.append(np.asarray(off, np.int64), len(col))
    preds = {}
    for u in range(len(off)):
        if dist[u] < 0:
            continue
        for j in range(o[u], o[u + 1]):
            v = int(col[j])
            if v >= 0 and dist[v] >= 0 and dist[u] + int(w[j]) == dist[v]:
                preds.setdefault(v, set()).add(u)
    order = np.argsort(dist, kind="stable")
    count = {s: 1}
    for v in order:
        v = int(v)
        if dist[v] < 0 or v == s:
            continue
        count[v] = min(2, sum(count.get(u, 0) for u in preds.get(v, ())))
    return count.get(t, 0) == 1


def test_a_star_cpu_contract_golden():
    """a_star_cpu (GA_star/A_star_cpu.cpp:100-204) on the GPU: with
    --row-end-quirk its stdout equals the reference program's output held in
    tests/golden/astar_cpu_ref.json, byte for byte where the shortest path is
    unique (otherwise the same cost line and a valid path of that cost);
    without the switch the cost is the shortest distance on the whole graph."""
    import json
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "astar_cpu_ref.json")))["cases"]
    exact = 0
    with tempfile.TemporaryDirectory() as d:
        for i, c in enumerate(gold):
            off, col, w = (np.array(c["off"], np.int32), np.array(c["col"], np.int32), np.array(c["w"], np.uint32))
            h = np.array(c["h"], np.int32)
            s, t = c["s"], c["t"]
            pyoracle.write_graph_txt(os.path.join(d, "graph.txt"), off, col, w)
            pyoracle.write_hx_txt(os.path.join(d, "Hx.txt"), h)
            print("golden case", i, flush=True)
            out = run("a_star_cpu", d, f"{s} {t}\n", "--row-end-quirk")
            qcol = _quirk_graph(off, col, w)
            if c["found"]:
                want = "completed input\ncost %d\n%s \n" % (c["cost"], " ".join(map(str, c["path_t_to_s"])))
            else:
                want = "completed input\nunreachable !!\n"
            if _unique_shortest_path(off, qcol, w, s, t):
                assert out == want, (i, out, want)
                exact += 1
            else:
                lines = out.splitlines()
                assert lines[:2] == want.splitlines()[:2], (i, out, want)
                path = [int(x) for x in lines[2].split()][::-1]
                assert path[0] == s and path[-1] == t and path_cost(off, qcol, w, path) == c["cost"]
            if i % 6 == 0:  # without the switch: the whole graph
                out = run("a_star_cpu", d, f"{s} {t}\n").splitlines()
                dist, _ = pyoracle.dijkstra(off, col, w, s)
                if dist[t] < 0:
                    assert out[1] == "unreachable !!"
                else:
                    assert out[1] == f"cost {int(dist[t])}"
                    path = [int(x) for x in out[2].split()][::-1]
                    assert path_cost(off, col, w, path) == int(dist[t])
    assert exact >= len(gold) // 2  # most cases pin the printed path too


def test_a_star_cpu_contract_live():
    """a_star_cpu against the reference's own A_star_cpu.cpp (compiled into
    oracle/_ref) on larger graphs with BFS heuristics: identical stdout when
    the shortest path is unique, else identical cost line."""
    if not pyoracle.ref_tool("a_star_cpu"):
        pytest.skip("oracle/_ref/a_star_cpu not built")
    hg = ddp_amd.HostGraph.random(5000, 4, 1, 1000, 51)
    src, dst = hg.reachable_pairs(6, 52)
    with tempfile.TemporaryDirectory() as d:
        hg.write_txt(os.path.join(d, "graph.txt"))
        for s, t in zip(src, dst):
            s, t = int(s), int(t)
            roff, rcol, rw = pyoracle.transpose(hg.off, hg.col, hg.w)
            hops, _ = pyoracle.bfs_hops(roff, rcol, t)
            pyoracle.write_hx_txt(os.path.join(d, "Hx.txt"), np.where(hops < 0, -1, hops * int(hg.w.min())))
            ours = run("a_star_cpu", d, f"{s} {t}\n", "--row-end-quirk")
            ref = subprocess.run([pyoracle.ref_tool("a_star_cpu")], input=f"{s} {t}\n", capture_output=True,
                                 text=True, cwd=d, timeout=120, check=True).stdout
            if _unique_shortest_path(hg.off, _quirk_graph(hg.off, hg.col, hg.w), hg.w, s, t):
                assert ours == ref
            else:
                assert ours.splitlines()[:2] == ref.splitlines()[:2]
