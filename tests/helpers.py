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


def edge_min_weight(off, col, w, a, b):
    o = full_offsets(off, len(col))
    best = None
    for j in range(o[a], o[a + 1]):
        if col[j] == b and (best is None or w[j] < best):
            best = int(w[j])
    return best


def path_cost(off, col, w, path):
    """Sum over consecutive nodes of the cheapest parallel edge; None if an
    edge is missing."""
    total = 0
    for a, b in zip(path[:-1], path[1:]):
        x = edge_min_weight(off, col, w, int(a), int(b))
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
