"""Small-graph builders and result validators shared by the tests."""
import numpy as np


def random_graph(rng, N, max_deg=4, wmax=20, allow_self=True, tombstones=0.0):
    """CSR (reference layout) with random out-degrees, parallel edges allowed."""
    degs = rng.integers(0, max_deg + 1, size=N)
    E = int(degs.sum())
    off = np.zeros(N, np.int32)
    off[1:] = np.cumsum(degs)[:-1]
    col = rng.integers(0, N, size=E).astype(np.int32)
    if not allow_self:
        for u in range(N):
            for j in range(off[u], off[u] + degs[u]):
                if col[j] == u:
                    col[j] = (u + 1) % N
    w = rng.integers(1, wmax + 1, size=E).astype(np.uint32)
    if tombstones > 0 and E:
        dead = rng.random(E) < tombstones
        col[dead] = -1
    return off, col, w


def full_offsets(off, E):
    return np.append(np.asarray(off, np.int64), E)


def edge_min_weight(off, col, w, a, b, o=None):
    """Weight of the cheapest a -> b edge, None if there is none.  o: the
    full offsets (N + 1 entries), when the caller already has them."""
    if o is None:
        o = full_offsets(off, len(col))
    best = None
    for j in range(o[a], o[a + 1]):
        if col[j] == b and (best is None or w[j] < best):
            best = int(w[j])
    return best


def path_cost(off, col, w, path):
    """Sum over consecutive nodes of the cheapest parallel edge; None if an
    edge is missing."""
    o = full_offsets(off, len(col))
    total = 0
    for a, b in zip(path[:-1], path[1:]):
        x = edge_min_weight(off, col, w, int(a), int(b), o)
        if x is None:
            return None
        total += x
    return total


def transpose_np(off, col, w):
    N, E = len(off), len(col)
    o = full_offsets(off, E)
    src = np.repeat(np.arange(N), np.diff(o))
    keep = col >= 0
    src, dst, ww = src[keep], col[keep], w[keep]
    order = np.lexsort((np.arange(len(dst)), dst))
    roff = np.zeros(N, np.int32)
    cnt = np.bincount(dst, minlength=N)
    roff[1:] = np.cumsum(cnt)[:-1]
    return roff, src[order].astype(np.int32), ww[order].astype(np.uint32)


def euclid_ceil(p, q):
    """ceil(sqrt(|p - q|^2)) in exact integer arithmetic."""
    d2 = int(((np.asarray(p, np.int64) - np.asarray(q, np.int64)) ** 2).sum())
    r = int(np.floor(np.sqrt(float(d2))))
    while r * r > d2:
        r -= 1
    while (r + 1) * (r + 1) <= d2:
        r += 1
    return r if r * r == d2 else r + 1


def geometric_updates(hg, nb, per, seed, extra=1000):
    """Update batches for a graph with coordinates that keep the Euclidean
    heuristic admissible: deletes pick existing edges, each insert u->v
    weighs ceil(dist(u, v)) plus up to `extra`."""
    rng = np.random.default_rng(seed)
    p = hg.xy.reshape(-1, 2)
    o = full_offsets(hg.off, hg.E)
    out = []
    for _ in range(nb):
        lines = []
        for u in rng.integers(0, hg.N, per):
            if o[u + 1] > o[u]:
                lines.append((0, int(u), int(hg.col[rng.integers(o[u], o[u + 1])]), 0))
        for u, v in zip(rng.integers(0, hg.N, per), rng.integers(0, hg.N, per)):
            lines.append((1, int(u), int(v), max(1, euclid_ceil(p[u], p[v])) + int(rng.integers(0, extra))))
        out.append(lines)
    return out


def rmat_updates(hg, nb, per, seed):
    """nb update batches of `per` deletes of existing edges (sources drawn
    uniformly among nodes with out-edges) and `per` random inserts with
    w uniform in [1, 1000]: the C5 serving mix."""
    rng = np.random.default_rng(seed)
    o = full_offsets(hg.off, hg.E)
    deg = np.diff(o)
    out = []
    for _ in range(nb):
        u = rng.integers(0, hg.N, per)
        u = u[deg[u] > 0]
        j = o[u] + (rng.random(len(u)) * deg[u]).astype(np.int64)
        dels = [(0, int(a), int(b), 0) for a, b in zip(u, hg.col[j])]
        iu, iv, iw = rng.integers(0, hg.N, per), rng.integers(0, hg.N, per), rng.integers(1, 1001, per)
        out.append(dels + [(1, int(a), int(b), int(c)) for a, b, c in zip(iu, iv, iw)])
    return out
