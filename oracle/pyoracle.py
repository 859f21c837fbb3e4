"""ctypes wrapper of oracle/liboracle.so plus helpers that run the reference's
own tools compiled into oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_int, c_int64, c_longlong, c_uint32

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_P_I = POINTER(c_int)
_P_U = POINTER(c_uint32)
_P_LL = POINTER(c_longlong)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        L.orc_astar_ref.restype = c_int
        L.orc_astar_ref.argtypes = [c_int, c_int, _P_I, _P_I, _P_U, _P_I, c_int, c_int, c_int, _P_I, _P_I, _P_LL]
        L.orc_dijkstra.restype = c_longlong
        L.orc_dijkstra.argtypes = [c_int, c_int, _P_I, _P_I, _P_U, c_int, c_int, _P_LL]
        L.orc_bfs_hops.restype = c_int
        L.orc_bfs_hops.argtypes = [c_int, c_int, _P_I, _P_I, c_int, _P_I]
        L.orc_transpose.restype = c_int
        L.orc_transpose.argtypes = [c_int, c_int, _P_I, _P_I, _P_U, c_int, _P_I, _P_I, _P_U]
        L.orc_csr_from_edges.restype = c_int
        L.orc_csr_from_edges.argtypes = [c_int, _P_I, _P_I, _P_U, _P_I, _P_I, _P_U]
        L.orc_apply_batch.restype = c_int
        L.orc_apply_batch.argtypes = [c_int, c_int, _P_I, _P_I, _P_U, c_int, _P_I, _P_I, _P_I, _P_U, _P_I, _P_I,
                                      _P_U, _P_I]
        _lib = L
    return _lib


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_P_I)


def _u(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_P_U)


def astar_ref(off, col, w, h, s, t, quirk=True):
    """Restatement of GA_star/A_star_cpu.cpp.  Returns (found, cost, parent, expansions)."""
    N, E = len(off), len(col)
    off, po = _i(off)
    col, pc = _i(col)
    w, pw = _u(w)
    h, ph = _i(h)
    parent = np.zeros(N, np.int32)
    cost = c_int()
    ex = c_longlong()
    found = lib().orc_astar_ref(N, E, po, pc, pw, ph, s, t, int(quirk), ctypes.byref(cost),
                                parent.ctypes.data_as(_P_I), ctypes.byref(ex))
    if found < 0:
        raise MemoryError("oracle allocation failed")
    return bool(found), cost.value, parent, ex.value


def dijkstra(off, col, w, s, stop_at=-1):
    """Distances from s (int64, -1 = unreachable) and the settled-node count."""
    N, E = len(off), len(col)
    off, po = _i(off)
    col, pc = _i(col)
    w, pw = _u(w)
    dist = np.zeros(N, np.int64)
    settled = lib().orc_dijkstra(N, E, po, pc, pw, s, stop_at, dist.ctypes.data_as(_P_LL))
    return dist, settled


def bfs_hops(off, col, t):
    N, E = len(off), len(col)
    off, po = _i(off)
    col, pc = _i(col)
    dist = np.zeros(N, np.int32)
    levels = lib().orc_bfs_hops(N, E, po, pc, t, dist.ctypes.data_as(_P_I))
    return dist, levels


def transpose(off, col, w):
    N, E = len(off), len(col)
    off, po = _i(off)
    col, pc = _i(col)
    w, pw = _u(w)
    roff = np.zeros(N, np.int32)
    rcol = np.zeros(max(E, 1), np.int32)
    rw = np.zeros(max(E, 1), np.uint32)
    RE = lib().orc_transpose(N, E, po, pc, pw, 1, roff.ctypes.data_as(_P_I), rcol.ctypes.data_as(_P_I),
                             rw.ctypes.data_as(_P_U))
    return roff, rcol[:RE].copy(), rw[:RE].copy()


def csr_from_edges(us, vs, ws):
    M = len(us)
    us, pu = _i(us)
    vs, pv = _i(vs)
    ws, pw = _u(ws)
    Nmax = int(max(us.max(initial=0), vs.max(initial=0))) + 1
    off = np.zeros(Nmax, np.int32)
    col = np.zeros(max(M, 1), np.int32)
    w = np.zeros(max(M, 1), np.uint32)
    N = lib().orc_csr_from_edges(M, pu, pv, pw, off.ctypes.data_as(_P_I), col.ctypes.data_as(_P_I),
                                 w.ctypes.data_as(_P_U))
    return off[:N].copy(), col[:M].copy(), w[:M].copy()


def apply_batch(off, col, w, lines):
    """One update batch (GA_merge semantics).  lines: iterable of (flag,u,v,w)."""
    N, E = len(off), len(col)
    lines = list(lines)
    fl = np.array([l[0] for l in lines], np.int32)
    us = np.array([l[1] for l in lines], np.int32)
    vs = np.array([l[2] for l in lines], np.int32)
    ws = np.array([l[3] for l in lines], np.uint32)
    nins = int((fl == 1).sum())
    off, po = _i(off)
    col, pc = _i(col)
    w, pw = _u(w)
    ooff = np.zeros(N, np.int32)
    ocol = np.zeros(E + nins + 1, np.int32)
    ow = np.zeros(E + nins + 1, np.uint32)
    nd = c_int()
    E2 = lib().orc_apply_batch(N, E, po, pc, pw, len(lines), _i(fl)[1] if len(lines) else None,
                               us.ctypes.data_as(_P_I), vs.ctypes.data_as(_P_I), ws.ctypes.data_as(_P_U),
                               ooff.ctypes.data_as(_P_I), ocol.ctypes.data_as(_P_I), ow.ctypes.data_as(_P_U),
                               ctypes.byref(nd))
    return ooff, ocol[:E2].copy(), ow[:E2].copy(), nd.value


# ---------------------------------------------------------- text formats

def write_graph_txt(path, off, col, w):
    """graph.txt in the token layout of GA_star/GenerateCSR.cpp:92-132."""
    with open(path, "w") as f:
        f.write(f"{len(off)} {len(col)}\n")
        f.write("".join(f"{int(x)} " for x in col) + "\n")
        f.write("".join(f"{int(x)} " for x in off) + "\n")
        f.write("".join(f"{int(x)} " for x in w) + "\n")


def read_graph_txt(path):
    toks = open(path).read().split()
    N, E = int(toks[0]), int(toks[1])
    vals = np.array(toks[2:2 + 2 * E + N], dtype=np.int64)
    col = vals[:E].astype(np.int32)
    off = vals[E:E + N].astype(np.int32)
    w = vals[E + N:].astype(np.uint32)
    return off, col, w


def write_hx_txt(path, h):
    with open(path, "w") as f:
        f.write("".join(f"{int(x)} " for x in h))


def ref_tool(name):
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None


def run_ref_astar(workdir, s, t, timeout=60):
    """Runs the reference's A_star_cpu.cpp binary in workdir (graph.txt,
    Hx.txt); returns (found, cost, path t..s) parsed from its stdout."""
    exe = ref_tool("a_star_cpu")
    out = subprocess.run([exe], input=f"{s} {t}\n", capture_output=True, text=True, cwd=workdir,
                         timeout=timeout, check=True).stdout
    lines = [l for l in out.splitlines() if l.strip() and l.strip() != "completed input"]
    if lines[0].startswith("unreachable"):
        return False, None, []
    cost = int(lines[0].split()[1])
    path = [int(x) for x in lines[1].split()]
    return True, cost, path


def chain_from_parent(parent, t, limit=None):
    """t, parent[t], ... until -1 (the reference's printed order)."""
    out = [t]
    v = t
    limit = limit or len(parent) + 1
    while parent[v] != -1:
        v = int(parent[v])
        out.append(v)
        if len(out) > limit:
            raise ValueError("parent cycle")
    return out


def astar_indexed(off, col, w, h, s, t):
    """Indexed-heap variant of the CPU A* (same costs, O(log) decrease-key).
    Returns (found, cost, expansions)."""
    L = lib()
    if not hasattr(L, "_ix_bound"):
        L.orc_astar_indexed.restype = c_int
        L.orc_astar_indexed.argtypes = [c_int, c_int, _P_I, _P_I, _P_U, _P_I, c_int, c_int, _P_I, _P_LL]
        L._ix_bound = True
    N, E = len(off), len(col)
    off, po = _i(off)
    col, pc = _i(col)
    w, pw = _u(w)
    h, ph = _i(h)
    cost = c_int()
    ex = c_longlong()
    found = L.orc_astar_indexed(N, E, po, pc, pw, ph, s, t, ctypes.byref(cost), ctypes.byref(ex))
    if found < 0:
        raise MemoryError("oracle allocation failed")
    return bool(found), cost.value, ex.value


def astar_batch_euclid(off, col, w, xy, src, dst, nthreads):
    """orc_astar_batch_euclid: A_star_cpu.cpp's search for every (src[i],
    dst[i]) on a pool of nthreads workers, h = floor(Euclid) to the target.
    Returns (costs, total expansions, threads started); unreachable = INT_MAX."""
    L = lib()
    if not hasattr(L, "_bx_bound"):
        L.orc_astar_batch_euclid.restype = c_int
        L.orc_astar_batch_euclid.argtypes = [c_int, c_int, _P_I, _P_I, _P_U, _P_I, c_int, _P_I, _P_I, c_int, _P_I,
                                             _P_LL]
        L._bx_bound = True
    N, E = len(off), len(col)
    off, po = _i(off)
    col, pc = _i(col)
    w, pw = _u(w)
    xy, pxy = _i(xy)
    src, ps = _i(src)
    dst, pd = _i(dst)
    costs = np.zeros(len(src), np.int32)
    ex = c_longlong()
    started = L.orc_astar_batch_euclid(N, E, po, pc, pw, pxy, len(src), ps, pd, int(nthreads),
                                       costs.ctypes.data_as(_P_I), ctypes.byref(ex))
    if started < 0:
        raise MemoryError("oracle allocation failed")
    return costs, ex.value, started
