"""Writes tests/golden/astar_cpu_ref.json: input/output vectors of the
reference's CPU A* (GA_star/A_star_cpu.cpp), produced by running the
reference's own program, compiled by oracle/Makefile into oracle/_ref/, on
seeded small graphs.  The file holds data only (graphs, heuristics, s, t and
the program's printed cost and path).

Run: python tests/golden/make_golden.py   (needs oracle/_ref/a_star_cpu)
"""
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from helpers import random_graph  # noqa: E402
from oracle import pyoracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "astar_cpu_ref.json")


def bfs_h(off, col, w, t, scale_minw=True):
    roff, rcol, rw = pyoracle.transpose(off, col, w)
    hops, _ = pyoracle.bfs_hops(roff, rcol, t)
    minw = int(w.min()) if len(w) and scale_minw else 1
    return np.where(hops < 0, -1, hops * minw).astype(np.int32)


def main():
    if not pyoracle.ref_tool("a_star_cpu"):
        sys.exit("oracle/_ref/a_star_cpu missing: run make -C oracle")
    rng = np.random.default_rng(20261202)
    cases = []
    for k in range(120):
        N = int(rng.integers(2, 40))
        off, col, w = random_graph(rng, N, max_deg=int(rng.integers(1, 5)), wmax=int(rng.choice([2, 3, 10, 50])))
        s, t = int(rng.integers(0, N)), int(rng.integers(0, N))
        dist, _ = pyoracle.dijkstra(off, col, w, s)
        reach = np.flatnonzero(dist > 0)
        if len(reach) and rng.random() < 0.85:  # mostly reachable targets
            t = int(rng.choice(reach))
        kind = ["bfs", "zero", "bfs"][k % 3]
        h = bfs_h(off, col, w, t) if kind == "bfs" else np.zeros(N, np.int32)
        cases.append((off, col, w, h, s, t, kind))
    # the row-end quirk: the only edge into t is the last edge, leaving node N-1
    off = np.array([0, 1, 2], np.int32)
    col = np.array([2, 1, 0], np.int32)  # 0->2, 1->1, 2->0 : edge E-1 = 2->0
    w = np.array([5, 1, 7], np.uint32)
    cases.append((off, col, w, np.zeros(3, np.int32), 2, 0, "quirk"))
    cases.append((off, col, w, np.zeros(3, np.int32), 0, 2, "quirk"))
    out = []
    with tempfile.TemporaryDirectory() as d:
        for off, col, w, h, s, t, kind in cases:
            pyoracle.write_graph_txt(os.path.join(d, "graph.txt"), off, col, w)
            pyoracle.write_hx_txt(os.path.join(d, "Hx.txt"), h)
            found, cost, path = pyoracle.run_ref_astar(d, s, t)
            out.append({"off": off.tolist(), "col": col.tolist(), "w": w.tolist(), "h": h.tolist(), "s": s,
                        "t": t, "kind": kind, "found": found, "cost": cost, "path_t_to_s": path})
    with open(OUT, "w") as f:
        json.dump({"source": "GA_star/A_star_cpu.cpp compiled from the reference sources", "cases": out}, f)
    print(f"wrote {len(out)} cases to {OUT}")


if __name__ == "__main__":
    main()
