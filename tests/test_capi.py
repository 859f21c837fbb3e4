"""CPU-side checks of the product library: it loads, exports every symbol the
headers declare, and its host code (file formats, transpose, update batches,
generators) matches the oracle and the reference's own tools.  No GPU calls."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import ddp_amd
from ddp_amd import _capi
from helpers import full_offsets, random_graph
from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_symbols():
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ddp_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    names = declared_symbols()
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing
    for n in names:
        getattr(L, n)
    # the ctypes table binds exactly what the headers declare
    assert names == set(_capi.SIGNATURES), (names ^ set(_capi.SIGNATURES))


def test_context_without_gpu_fails_cleanly():
    try:
        c = ddp_amd.Context(0)
    except ddp_amd.DDPError as e:
        assert e.code in (2, 9)  # DDP_ERR_HIP / DDP_ERR_NODEVICE, never a crash
    else:
        c.close()
        pytest.skip("a GPU is present")


def test_graph_txt_writer_matches_reference_layout():
    rng = np.random.default_rng(1)
    with tempfile.TemporaryDirectory() as d:
        for _ in range(5):
            off, col, w = random_graph(rng, int(rng.integers(1, 50)), max_deg=4, tombstones=0.1)
            a, b = os.path.join(d, "a.txt"), os.path.join(d, "b.txt")
            ddp_amd.HostGraph(off, col, w).write_txt(a)
            pyoracle.write_graph_txt(b, off, col, w)
            assert open(a).read() == open(b).read()
            g = ddp_amd.HostGraph.read_txt(b)
            assert g.off.tolist() == off.tolist() and g.col.tolist() == col.tolist() and g.w.tolist() == w.tolist()


@pytest.mark.skipif(not pyoracle.ref_tool("gen_csr"), reason="oracle/_ref/gen_csr not built")
def test_reader_parses_reference_GenerateCSR_output():
    rng = np.random.default_rng(2)
    with tempfile.TemporaryDirectory() as d:
        M = 300
        us, vs, ws = rng.integers(0, 60, M), rng.integers(0, 60, M), rng.integers(1, 100, M)
        txt = "".join(f"{a} {b} {c}\n" for a, b, c in zip(us, vs, ws))
        out = subprocess.run([pyoracle.ref_tool("gen_csr")], input=txt, capture_output=True, text=True,
                             cwd=d, check=True).stdout
        p = os.path.join(d, "graph.txt")
        open(p, "w").write(out)
        g = ddp_amd.HostGraph.read_txt(p)
        off, col, w = pyoracle.csr_from_edges(us, vs, ws)
        assert g.off.tolist() == off.tolist() and g.col.tolist() == col.tolist() and g.w.tolist() == w.tolist()


def test_reader_errors():
    with tempfile.TemporaryDirectory() as d:
        with pytest.raises(ddp_amd.DDPError) as e:
            ddp_amd.HostGraph.read_txt(os.path.join(d, "missing.txt"))
        assert e.value.code == 5
        p = os.path.join(d, "t.txt")
        open(p, "w").write("3 4\n1 2 0 1\n0 2 ")  # truncated
        with pytest.raises(ddp_amd.DDPError) as e:
            ddp_amd.HostGraph.read_txt(p)
        assert e.value.code == 6
        open(p, "w").write("3 2\n1 7\n0 1 1\n5 5\n")  # column 7 >= N
        with pytest.raises(ddp_amd.DDPError) as e:
            ddp_amd.HostGraph.read_txt(p)
        assert e.value.code == 6
        open(p, "w").write("3 2\n1 x\n0 1 1\n5 5\n")
        with pytest.raises(ddp_amd.DDPError):
            ddp_amd.HostGraph.read_txt(p)


def test_hx_and_updates_roundtrip():
    with tempfile.TemporaryDirectory() as d:
        h = np.array([3, -1, 0, 7], np.int32)
        p = os.path.join(d, "Hx.txt")
        ddp_amd.write_hx(p, h)
        assert open(p).read() == "3 -1 0 7 "
        assert ddp_amd.read_hx(p, 4).tolist() == h.tolist()
        with pytest.raises(ddp_amd.DDPError):
            ddp_amd.read_hx(p, 5)
        u = ddp_amd.Updates([[(0, 1, 2, 0), (1, 3, 4, 9)], [], [(1, 0, 1, 5)]])
        q = os.path.join(d, "Updates.txt")
        u.write_txt(q)
        assert ddp_amd.Updates.read_txt(q).batches == u.batches
        open(q, "w").write("2\n1 0 1 5\n")
        with pytest.raises(ddp_amd.DDPError):
            ddp_amd.Updates.read_txt(q)


def test_transpose_and_apply_batch_match_oracle():
    rng = np.random.default_rng(3)
    for _ in range(30):
        N = int(rng.integers(2, 60))
        off, col, w = random_graph(rng, N, max_deg=5, wmax=40)
        g = ddp_amd.HostGraph(off, col, w)
        r = g.transpose()
        roff, rcol, rw = pyoracle.transpose(off, col, w)
        assert r.off.tolist() == roff.tolist() and r.col.tolist() == rcol.tolist() and r.w.tolist() == rw.tolist()
        o = full_offsets(off, len(col))
        lines = []
        for _ in range(int(rng.integers(0, 10))):
            u = int(rng.integers(0, N))
            v = int(col[rng.integers(o[u], o[u + 1])]) if o[u + 1] > o[u] and rng.random() < 0.7 else int(
                rng.integers(0, N))
            lines.append((0, u, v, 0))
        for _ in range(int(rng.integers(0, 10))):
            lines.append((1, int(rng.integers(0, N)), int(rng.integers(0, N)), int(rng.integers(1, 40))))
        rng.shuffle(lines)
        g2, nd = g.apply_batch(lines)
        o2, c2, w2, nd2 = pyoracle.apply_batch(off, col, w, lines)
        assert nd == nd2 and g2.off.tolist() == o2.tolist() and g2.col.tolist() == c2.tolist()
        assert g2.w.tolist() == w2.tolist()


def test_generators_are_seeded_and_well_formed():
    a = ddp_amd.HostGraph.random(5000, 6, 1, 100, 1)
    b = ddp_amd.HostGraph.random(5000, 6, 1, 100, 1)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.w, b.w)
    assert a.E == 30000 and np.all(np.diff(full_offsets(a.off, a.E)) == 6)
    src = np.repeat(np.arange(a.N), 6)
    assert not np.any(a.col == src) and a.w.min() >= 1 and a.w.max() <= 100
    m = ddp_amd.HostGraph.grid_maze(50, 40, 300, 2)
    d, _ = pyoracle.dijkstra(m.off, m.col, m.w, 0)
    assert d[50 * 40 - 1] >= 50 + 40 - 2  # the carved corridor connects the corners
    assert set(np.unique(m.w).tolist()) <= {1}
    geo = ddp_amd.HostGraph.geometric(20000, 10, 100000, 4)
    assert 8.5 < geo.E / geo.N < 11.5
    o = full_offsets(geo.off, geo.E)
    src = np.repeat(np.arange(geo.N), np.diff(o))
    fwd = set(zip(src.tolist(), geo.col.tolist()))
    assert all((v, u) in fwd for u, v in list(fwd)[:2000])  # both directions kept
    xy = geo.xy.reshape(-1, 2).astype(np.int64)
    d2 = ((xy[src] - xy[geo.col]) ** 2).sum(1)
    ceil = np.ceil(np.sqrt(d2.astype(np.float64))).astype(np.int64)
    assert np.array_equal(np.maximum(ceil, 1), geo.w.astype(np.int64))
    r = ddp_amd.HostGraph.rmat(10, 16384, 0.57, 0.19, 0.19, 1, 1000, 5)
    assert r.N == 1024 and r.E == 16384
    s, t = geo.geo_pairs(20, 100, 100000, 1)
    assert np.all(s != t)
    u = a.gen_updates(3, 50, 40, 1, 1000, 3)
    assert len(u.batches) == 3 and all(len(bt) == 90 for bt in u.batches)
    assert all(l[0] == 0 for l in u.batches[0][:40]) and all(l[0] == 1 for l in u.batches[0][40:])


def test_hub_pairs_are_connected():
    hg = ddp_amd.HostGraph.rmat(12, 8 << 12, 0.57, 0.19, 0.19, 1, 100, 7)
    src, dst = hg.hub_pairs(40, 8)
    assert len(set(zip(src.tolist(), dst.tolist()))) > 30
    for s, t in zip(src, dst):
        assert s != t
        dist, _ = pyoracle.dijkstra(hg.off, hg.col, hg.w, int(s))
        assert dist[t] >= 0


def test_binary_graph_cache():
    """The cache returns exactly what the text parser returns, is reused while
    the text is unchanged, and is rebuilt when the text changes."""
    hg = ddp_amd.HostGraph.random(20000, 5, 1, 1000, 61)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "graph.txt")
        hg.write_txt(p)
        a = ddp_amd.HostGraph.read_txt(p, cache=True)  # parses, writes graph.txt.ddpbin
        assert os.path.exists(p + ".ddpbin")
        b = ddp_amd.HostGraph.read_txt(p, cache=True)  # from the cache
        for g in (a, b):
            assert g.N == hg.N and g.E == hg.E
            assert np.array_equal(g.off, hg.off) and np.array_equal(g.col, hg.col) and np.array_equal(g.w, hg.w)
        # a cache that does not belong to the text file is ignored
        other = ddp_amd.HostGraph.random(500, 3, 1, 9, 62)
        other.write_txt(p)
        c = ddp_amd.HostGraph.read_txt(p, cache=True)
        assert c.N == 500 and np.array_equal(c.col, other.col)
        # a corrupt cache is ignored too
        with open(p + ".ddpbin", "r+b") as f:
            f.truncate(100)
        e = ddp_amd.HostGraph.read_txt(p, cache=True)
        assert np.array_equal(e.col, other.col)
