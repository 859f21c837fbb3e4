"""MI355X-native parallel A* (GA*, reverse-BFS heuristic, DA*) over a C ABI.

Python is a thin host layer: every search runs in libddp.so's HIP kernels for
gfx950.  The class and method names of the reference's library facade are
kept where one exists (GPU_Graph, GPU_A_Star.set_huiristics / get_path,
LIB/include/a_star.cuh:12-51).
"""
import ctypes
import os
import weakref
from ctypes import byref, c_int32, c_uint32, c_void_p

import numpy as np

from . import _capi
from ._capi import (DDPError, ENGINE_AUTO, ENGINE_BIDI, ENGINE_HEAPS, ENGINE_PILES, GRAPH_REORDER, GRAPH_REVERSE, H_ARRAY, H_EUCLID,
                    H_MANHATTAN, H_ZERO, INF_COST, MODE_AUTO,
                    MODE_BIDI, MODE_GRID, MODE_PILES, MODE_WORKGROUP, AstarOpts,
                    AstarStats, BatchOpts, BatchStats, DeviceInfo, Heur, Update, as_i32, as_u32, ptr_f64, ptr_i32,
                    ptr_u32)

__all__ = [
    "DDPError", "INF_COST", "H_ZERO", "H_ARRAY", "H_MANHATTAN", "H_EUCLID", "MODE_AUTO", "MODE_GRID",
    "MODE_WORKGROUP", "MODE_PILES", "MODE_BIDI", "ENGINE_AUTO", "ENGINE_HEAPS", "ENGINE_PILES", "ENGINE_BIDI", "Context", "HostGraph",
    "DeviceGraph", "AStarResult", "astar", "Batch", "astar_batch", "Updates", "GPU_Graph", "GPU_A_Star",
    "DAStar", "Server", "alternative_paths",
]


def _check(rc, ctx=None):
    if rc != 0:
        L = _capi.lib()
        msg = L.ddp_last_error(ctx).decode() if ctx else L.ddp_create_error().decode()
        raise DDPError(rc, msg)


class Context:
    """A device plus the HIP stream every call on it uses (ddp_ctx)."""

    def __init__(self, device=0):
        L = _capi.lib()
        h = c_void_p()
        rc = L.ddp_ctx_create(device, byref(h))
        if rc:
            raise DDPError(rc, L.ddp_create_error().decode())
        self._h = h
        self._children = weakref.WeakSet()

    @property
    def handle(self):
        return self._h

    def _adopt(self, obj):
        self._children.add(obj)

    def synchronize(self):
        _check(_capi.lib().ddp_ctx_synchronize(self._h), self._h)

    def device_info(self):
        info = DeviceInfo()
        _check(_capi.lib().ddp_device_info(self._h, byref(info)), self._h)
        return {f: getattr(info, f) for f, _ in DeviceInfo._fields_}

    def close(self):
        if getattr(self, "_h", None):
            for obj in list(self._children):  # handles first: they use the context's stream
                obj.close()
            _capi.lib().ddp_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ----------------------------------------------------------------- host graph

class HostGraph:
    """CSR in the reference's layout: off has N entries (no sentinel)."""

    def __init__(self, off, col, w, xy=None):
        self.off = as_i32(off)
        self.col = as_i32(col)
        self.w = as_u32(w)
        self.N = int(self.off.shape[0])
        self.E = int(self.col.shape[0])
        self.xy = None if xy is None else as_i32(xy)

    # -- conversion to/from the library's host handle
    @classmethod
    def _from_handle(cls, h):
        L = _capi.lib()
        N, E = c_int32(), c_int32()
        po, pc, pw = _capi.P_I32(), _capi.P_I32(), _capi.P_U32()
        _check(L.ddp_hgraph_view(h, byref(N), byref(E), byref(po), byref(pc), byref(pw)))
        off = np.ctypeslib.as_array(po, shape=(N.value,)).copy()
        col = np.ctypeslib.as_array(pc, shape=(E.value,)).copy() if E.value else np.zeros(0, np.int32)
        w = np.ctypeslib.as_array(pw, shape=(E.value,)).copy() if E.value else np.zeros(0, np.uint32)
        pxy = L.ddp_hgraph_coords(h)
        xy = np.ctypeslib.as_array(pxy, shape=(2 * N.value,)).copy() if pxy else None
        L.ddp_hgraph_free(h)
        return cls(off, col, w, xy)

    def _handle(self):
        L = _capi.lib()
        h = c_void_p()
        _check(L.ddp_hgraph_from_arrays(self.N, self.E, ptr_i32(self.off), ptr_i32(self.col), ptr_u32(self.w),
                                        byref(h)))
        return h

    def offsets_full(self):
        """N+1 offsets (sentinel E appended)."""
        return np.append(self.off, np.int32(self.E))

    # -- files
    @classmethod
    def read_txt(cls, path, cache=False, cache_path=None):
        """graph.txt (reference layout).  cache=True reads through a binary
        cache (path + ".ddpbin" unless cache_path is given), written on the
        first read and reused while the text file is unchanged."""
        L = _capi.lib()
        h = c_void_p()
        err = ctypes.create_string_buffer(512)
        if cache or cache_path:
            rc = L.ddp_read_graph_cached(os.fsencode(path), None if cache_path is None else os.fsencode(cache_path),
                                         byref(h), err, 512)
        else:
            rc = L.ddp_read_graph_txt(os.fsencode(path), byref(h), err, 512)
        if rc:
            raise DDPError(rc, err.value.decode())
        return cls._from_handle(h)

    def write_txt(self, path):
        L = _capi.lib()
        h = self._handle()
        try:
            _check(L.ddp_write_graph_txt(os.fsencode(path), h))
        finally:
            L.ddp_hgraph_free(h)

    def transpose(self):
        L = _capi.lib()
        h = self._handle()
        r = c_void_p()
        try:
            _check(L.ddp_hgraph_transpose(h, byref(r)))
        finally:
            L.ddp_hgraph_free(h)
        return HostGraph._from_handle(r)

    def apply_batch(self, lines):
        """Returns (G_{b+1}, effective deletes) for one update batch."""
        L = _capi.lib()
        arr = _updates_array(lines)
        h = self._handle()
        r = c_void_p()
        nd = c_int32()
        try:
            _check(L.ddp_hgraph_apply_batch(h, arr, len(lines), byref(r), byref(nd)))
        finally:
            L.ddp_hgraph_free(h)
        g = HostGraph._from_handle(r)
        g.xy = self.xy
        return g, nd.value

    # -- generators
    @classmethod
    def _gen(cls, fn, *args):
        h = c_void_p()
        _check(fn(*args, byref(h)))
        return cls._from_handle(h)

    @classmethod
    def random(cls, N, deg, wmin, wmax, seed):
        return cls._gen(_capi.lib().ddp_gen_random, N, deg, wmin, wmax, seed)

    @classmethod
    def grid_maze(cls, width, height, p_block_permille, seed):
        return cls._gen(_capi.lib().ddp_gen_grid_maze, width, height, p_block_permille, seed)

    @classmethod
    def geometric(cls, N, mean_degree, scale, seed):
        return cls._gen(_capi.lib().ddp_gen_geometric, N, float(mean_degree), scale, seed)

    @classmethod
    def rmat(cls, scale, E, a, b, c, wmin, wmax, seed):
        return cls._gen(_capi.lib().ddp_gen_rmat, scale, E, a, b, c, wmin, wmax, seed)

    def geo_pairs(self, nq, dmin, dmax, seed):
        L = _capi.lib()
        src = np.zeros(nq, np.int32)
        dst = np.zeros(nq, np.int32)
        h = self._with_coords_handle()
        try:
            _check(L.ddp_gen_geo_pairs(h, nq, float(dmin), float(dmax), seed, ptr_i32(src), ptr_i32(dst)))
        finally:
            L.ddp_hgraph_free(h)
        return src, dst

    def reachable_pairs(self, nq, seed):
        L = _capi.lib()
        src = np.zeros(nq, np.int32)
        dst = np.zeros(nq, np.int32)
        h = self._handle()
        try:
            _check(L.ddp_gen_reachable_pairs(h, nq, seed, ptr_i32(src), ptr_i32(dst)))
        finally:
            L.ddp_hgraph_free(h)
        return src, dst

    def hub_pairs(self, nq, seed):
        """Connected pairs through the largest out-degree node (ddp_gen_hub_pairs)."""
        L = _capi.lib()
        src = np.zeros(nq, np.int32)
        dst = np.zeros(nq, np.int32)
        h = self._handle()
        try:
            _check(L.ddp_gen_hub_pairs(h, nq, seed, ptr_i32(src), ptr_i32(dst)))
        finally:
            L.ddp_hgraph_free(h)
        return src, dst

    def _with_coords_handle(self):
        if self.xy is None:
            raise ValueError("graph has no coordinates")
        h = self._handle()
        _check(_capi.lib().ddp_hgraph_set_coords(h, ptr_i32(self.xy)))
        return h

    def gen_updates(self, nb, n_ins, n_del, wmin, wmax, seed):
        L = _capi.lib()
        h = self._handle()
        u = c_void_p()
        try:
            _check(L.ddp_gen_updates(h, nb, n_ins, n_del, wmin, wmax, seed, byref(u)))
        finally:
            L.ddp_hgraph_free(h)
        return Updates._from_handle(u)


def _updates_array(lines):
    arr = (Update * max(1, len(lines)))()
    for i, l in enumerate(lines):
        arr[i].flag, arr[i].u, arr[i].v, arr[i].w = int(l[0]), int(l[1]), int(l[2]), int(l[3])
    return arr


class Updates:
    """Update batches: list of batches, each a list of (flag, u, v, w)."""

    def __init__(self, batches):
        self.batches = [list(map(tuple, b)) for b in batches]

    @classmethod
    def _from_handle(cls, h):
        L = _capi.lib()
        nb = L.ddp_updates_count(h)
        out = []
        for b in range(nb):
            p = ctypes.POINTER(Update)()
            n = c_int32()
            _check(L.ddp_updates_batch(h, b, byref(p), byref(n)))
            out.append([(p[i].flag, p[i].u, p[i].v, p[i].w) for i in range(n.value)])
        L.ddp_updates_free(h)
        return cls(out)

    @classmethod
    def read_txt(cls, path):
        L = _capi.lib()
        h = c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = L.ddp_read_updates_txt(os.fsencode(path), byref(h), err, 512)
        if rc:
            raise DDPError(rc, err.value.decode())
        return cls._from_handle(h)

    def write_txt(self, path):
        with open(path, "w") as f:
            for b in self.batches:
                f.write(f"{len(b)}\n")
                for fl, u, v, w in b:
                    f.write(f"{fl} {u} {v} {w}\n")


def read_hx(path, n):
    L = _capi.lib()
    h = np.zeros(n, np.int32)
    err = ctypes.create_string_buffer(512)
    rc = L.ddp_read_hx_txt(os.fsencode(path), n, ptr_i32(h), err, 512)
    if rc:
        raise DDPError(rc, err.value.decode())
    return h


def write_hx(path, h):
    h = as_i32(h)
    _check(_capi.lib().ddp_write_hx_txt(os.fsencode(path), h.shape[0], ptr_i32(h)))


# --------------------------------------------------------------- device graph

class DeviceGraph:
    """CSR resident in HBM (ddp_graph)."""

    def __init__(self, ctx, hg, reverse=False, reorder=False):
        """reorder: relabel nodes on the device for cache locality (Hilbert
        order of the coordinates, else breadth-first); every method still
        takes and returns the host graph's node ids."""
        self.ctx = ctx
        self.N, self.E = hg.N, hg.E
        self.reordered = bool(reorder)
        h = c_void_p()
        flags = (GRAPH_REVERSE if reverse else 0) | (GRAPH_REORDER if reorder else 0)
        xy = None if hg.xy is None else as_i32(hg.xy)
        _check(_capi.lib().ddp_graph_upload_ex(ctx.handle, hg.N, hg.E, ptr_i32(hg.off), ptr_i32(hg.col),
                                               ptr_u32(hg.w), ptr_i32(xy), flags, byref(h)), ctx.handle)
        self._h = h
        ctx._adopt(self)

    @property
    def handle(self):
        return self._h

    def info(self):
        N, E, mw, md = c_int32(), c_int32(), c_uint32(), c_int32()
        _check(_capi.lib().ddp_graph_info(self._h, byref(N), byref(E), byref(mw), byref(md)))
        return {"N": N.value, "E": E.value, "min_weight": mw.value, "max_out_degree": md.value}

    def set_h(self, h):
        h = as_i32(h)
        assert h.shape[0] == self.N
        _check(_capi.lib().ddp_graph_set_h(self._h, ptr_i32(h)), self.ctx.handle)

    def set_coords(self, xy):
        xy = as_i32(xy)
        assert xy.shape[0] == 2 * self.N
        _check(_capi.lib().ddp_graph_set_coords(self._h, ptr_i32(xy)), self.ctx.handle)

    def set_grid(self, width):
        _check(_capi.lib().ddp_graph_set_grid(self._h, width), self.ctx.handle)

    def bfs_hops(self, t, reverse=False, scale=1, store_as_h=False, want=True):
        out = np.zeros(self.N, np.int32) if want else None
        lv = c_int32()
        _check(_capi.lib().ddp_bfs_hops(self.ctx.handle, self._h, t, int(reverse), scale, int(store_as_h),
                                        ptr_i32(out), byref(lv)), self.ctx.handle)
        return out, lv.value

    def close(self):
        if getattr(self, "_h", None):
            _capi.lib().ddp_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AStarResult:
    def __init__(self, cost, path, parent, stats):
        self.cost = cost
        self.path = path
        self.parent = parent
        self.stats = stats

    @property
    def found(self):
        return self.cost != INF_COST


def _heur(kind, scale):
    return Heur(int(kind), int(scale))


def astar(ctx, dg, s, t, heur=H_ARRAY, scale=1, K=0, block=0, heap_cap=0, lds_cap=0, max_iters=0,
          want_parent=False, path_cap=None, mode=MODE_AUTO):
    """One query through the single-query GA* kernel (ddp_astar); mode picks
    grid or workgroup execution (MODE_AUTO: grid)."""
    o = AstarOpts(K, block, heap_cap, lds_cap, max_iters, mode)
    st = AstarStats()
    cost = c_int32()
    par = np.zeros(dg.N, np.int32) if want_parent else None
    cap = dg.N if path_cap is None else path_cap
    path = np.zeros(max(cap, 1), np.int32)
    plen = c_int32()
    _check(_capi.lib().ddp_astar(ctx.handle, dg.handle, _heur(heur, scale), s, t, byref(o), byref(cost),
                                 ptr_i32(par), ptr_i32(path), cap, byref(plen), byref(st)), ctx.handle)
    n = plen.value
    p = path[:n].copy() if n > 0 else np.zeros(0, np.int32)
    stats = {f: getattr(st, f) for f, _ in AstarStats._fields_}
    return AStarResult(cost.value, p, par, stats)


class Batch:
    """Batched queries with slots kept between launches (ddp_batch)."""

    def __init__(self, ctx, dg, max_queries, heur=H_EUCLID, scale=1, block=0, slots=0, heap_cap=0, lds_cap=0,
                 max_iters=0, path_cap=0, schedule=0, engine=ENGINE_AUTO):
        self.ctx, self.dg = ctx, dg
        self.path_cap = path_cap
        o = BatchOpts(block, slots, heap_cap, lds_cap, max_iters, path_cap, schedule, engine)
        h = c_void_p()
        _check(_capi.lib().ddp_batch_create(ctx.handle, dg.handle, _heur(heur, scale), max_queries, byref(o),
                                            byref(h)), ctx.handle)
        self._h = h
        ctx._adopt(self)
        self.nq = 0

    def info(self):
        v = [c_int32() for _ in range(4)]
        _check(_capi.lib().ddp_batch_info(self._h, *[byref(x) for x in v]))
        return dict(zip(["slots", "block", "heap_cap", "lds_cap"], [x.value for x in v]))

    def upload(self, src, dst):
        src, dst = as_i32(src), as_i32(dst)
        _check(_capi.lib().ddp_batch_upload(self._h, src.shape[0], ptr_i32(src), ptr_i32(dst)), self.ctx.handle)
        self.nq = src.shape[0]

    def launch(self):
        _check(_capi.lib().ddp_batch_launch(self._h), self.ctx.handle)

    def wait(self):
        st = BatchStats()
        _check(_capi.lib().ddp_batch_wait(self._h, byref(st)), self.ctx.handle)
        return {f: getattr(st, f) for f, _ in BatchStats._fields_}

    def download(self, paths=False):
        n = self.nq
        cost = np.zeros(n, np.int32)
        plen = np.zeros(n, np.int32)
        qerr = np.zeros(n, np.uint32)
        lat = np.zeros(n, np.float64)
        pb = np.zeros(n * self.path_cap, np.int32) if (paths and self.path_cap) else None
        _check(_capi.lib().ddp_batch_download(self._h, ptr_i32(cost), ptr_i32(plen), ptr_i32(pb), ptr_u32(qerr),
                                              ptr_f64(lat)), self.ctx.handle)
        out = {"cost": cost, "path_len": plen, "qerr": qerr, "latency_ms": lat}
        if pb is not None:
            out["paths"] = [pb[i * self.path_cap:i * self.path_cap + plen[i]].copy() if plen[i] > 0
                            else np.zeros(0, np.int32) for i in range(n)]
        return out

    def close(self):
        if getattr(self, "_h", None):
            _capi.lib().ddp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def astar_batch(ctx, dg, src, dst, heur=H_EUCLID, scale=1, block=0, slots=0, heap_cap=0, lds_cap=0,
                path_cap=0, schedule=0, engine=ENGINE_AUTO):
    src, dst = as_i32(src), as_i32(dst)
    n = src.shape[0]
    o = BatchOpts(block, slots, heap_cap, lds_cap, 0, path_cap, schedule, engine)
    cost = np.zeros(n, np.int32)
    plen = np.zeros(n, np.int32)
    pb = np.zeros(max(1, n * path_cap), np.int32) if path_cap else None
    st = BatchStats()
    _check(_capi.lib().ddp_astar_batch(ctx.handle, dg.handle, _heur(heur, scale), n, ptr_i32(src), ptr_i32(dst),
                                       byref(o), ptr_i32(cost), ptr_i32(plen), ptr_i32(pb), byref(st)), ctx.handle)
    paths = None
    if pb is not None:
        paths = [pb[i * path_cap:i * path_cap + plen[i]].copy() if plen[i] > 0 else np.zeros(0, np.int32)
                 for i in range(n)]
    return cost, paths, {f: getattr(st, f) for f, _ in BatchStats._fields_}


class DAStar:
    """Dynamic A* for one (s, t) query over a graph that receives update
    batches (ddp_dastar).  apply() returns the new cost and per-batch stats."""

    REPAIR, RECOMPUTE = 0, 1

    def __init__(self, ctx, hg, s, t, h=None, K=0, block=0, heap_cap=0, lds_cap=0, mode=MODE_AUTO):
        """mode: MODE_PILES runs every search (initial, resumed, recomputed) on
        one workgroup's pile engine; anything else the GA* grid kernel."""
        self.ctx, self.N, self.s, self.t = ctx, hg.N, s, t
        o = AstarOpts(K, block, heap_cap, lds_cap, 0, mode)
        hh = None if h is None else as_i32(h)
        handle = c_void_p()
        cost = c_int32()
        _check(_capi.lib().ddp_dastar_open(ctx.handle, hg.N, hg.E, ptr_i32(hg.off), ptr_i32(hg.col), ptr_u32(hg.w),
                                           ptr_i32(hh), s, t, byref(o), byref(handle), byref(cost)), ctx.handle)
        self._h = handle
        ctx._adopt(self)
        self.cost = cost.value

    def apply(self, lines, mode=0):
        arr = _updates_array(lines)
        cost = c_int32()
        st = _capi.DAStats()
        _check(_capi.lib().ddp_dastar_apply(self._h, arr, len(lines), mode, byref(cost), byref(st)), self.ctx.handle)
        self.cost = cost.value
        return cost.value, {f: getattr(st, f) for f, _ in _capi.DAStats._fields_}

    def compact(self):
        """Fold the inserted-edge log into the CSR rows (ddp_dastar_compact)."""
        _check(_capi.lib().ddp_dastar_compact(self._h), self.ctx.handle)

    def set_compaction(self, log_entries):
        """Compact after any batch that leaves more than log_entries in the log (0: never)."""
        _check(_capi.lib().ddp_dastar_set_compaction(self._h, log_entries), self.ctx.handle)

    def path(self):
        buf = np.zeros(self.N, np.int32)
        n = c_int32()
        _check(_capi.lib().ddp_dastar_path(self._h, ptr_i32(buf), self.N, byref(n)), self.ctx.handle)
        return buf[:max(n.value, 0)].copy()

    def graph(self):
        E = c_int32()
        _check(_capi.lib().ddp_dastar_num_edges(self._h, byref(E)))
        off = np.zeros(self.N, np.int32)
        col = np.zeros(max(E.value, 1), np.int32)
        w = np.zeros(max(E.value, 1), np.uint32)
        _check(_capi.lib().ddp_dastar_export(self._h, ptr_i32(off), ptr_i32(col), ptr_u32(w)))
        return HostGraph(off, col[:E.value], w[:E.value])

    def close(self):
        if getattr(self, "_h", None):
            _capi.lib().ddp_dastar_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def alternative_paths(ctx, hg, s, t, num_paths, h=None, K=0):
    """The loop of k-short/GA_merge.cu:702-945: solve, then num_paths times
    delete the edge into t on the current best path (parent(t) -> t, first
    live instance) and re-solve, here by DA* repair.  Returns [(cost, path
    s..t)], stopping after the first unreachable result."""
    d = DAStar(ctx, hg, s, t, h=h, K=K)
    try:
        out = []
        cost = d.cost
        for k in range(num_paths + 1):
            path = d.path() if cost != INF_COST else np.zeros(0, np.int32)
            out.append((cost, path))
            if k == num_paths or len(path) < 2:
                break
            cost, _ = d.apply([(0, int(path[-2]), int(t), 0)], mode=DAStar.REPAIR)
        return out
    finally:
        d.close()


def _result_dict(r):
    return {f: getattr(r, f) for f, _ in _capi.ServerResult._fields_}


class Server:
    """Continuous-batching A* server on one GPU (ddp_server)."""

    QUERY, UPDATE = 0, 1

    def __init__(self, ctx, dg, heur=H_ZERO, scale=1, block=0, slots=0, heap_cap=0, ring=0, max_inserts=0,
                 idle_ms=0, engine=ENGINE_AUTO, grid_k=0):
        self.ctx, self.dg = ctx, dg
        o = _capi.ServerOpts(block, slots, heap_cap, ring, max_inserts, idle_ms, engine, grid_k)
        h = c_void_p()
        _check(_capi.lib().ddp_server_create(ctx.handle, dg.handle, _heur(heur, scale), byref(o), byref(h)),
               ctx.handle)
        self._h = h
        ctx._adopt(self)

    def info(self):
        v = [c_int32() for _ in range(3)]
        _check(_capi.lib().ddp_server_info(self._h, *[byref(x) for x in v]))
        return dict(zip(["slots", "block", "version"], [x.value for x in v]))

    def stats(self):
        st = _capi.ServerStats()
        _check(_capi.lib().ddp_server_stats(self._h, byref(st)))
        return {f: getattr(st, f) for f, _ in _capi.ServerStats._fields_}

    def submit_query(self, s, t, tag):
        return _capi.lib().ddp_server_submit_query(self._h, s, t, tag)

    def submit_update(self, lines, tag):
        arr = _updates_array(lines)
        return _capi.lib().ddp_server_submit_update(self._h, arr, len(lines), tag)

    def poll(self, max_results=4096):
        buf = (_capi.ServerResult * max_results)()
        n = c_int32()
        _check(_capi.lib().ddp_server_poll(self._h, buf, max_results, byref(n)), self.ctx.handle)
        return [_result_dict(buf[i]) for i in range(n.value)]

    def drain(self, timeout_ms=600000):
        _check(_capi.lib().ddp_server_drain(self._h, timeout_ms), self.ctx.handle)

    def compact(self):
        """Fold the inserted-edge log into CSR rows (ddp_server_compact)."""
        _check(_capi.lib().ddp_server_compact(self._h), self.ctx.handle)

    def replay(self, arrival_ms, types, src, dst, updates):
        """Open-loop replay; updates: list aligned with requests (None for
        queries).  Returns one result dict per request."""
        n = len(arrival_ms)
        arr = np.ascontiguousarray(arrival_ms, np.float64)
        ty = as_i32(types)
        s = as_i32(src)
        t = as_i32(dst)
        offs = [0]
        flat = []
        for u in updates:
            if u:
                flat.extend(u)
            offs.append(len(flat))
        uo = as_i32(offs)
        ua = _updates_array(flat)
        out = (_capi.ServerResult * max(1, n))()
        _check(_capi.lib().ddp_server_replay(self._h, n, ptr_f64(arr), ptr_i32(ty), ptr_i32(s), ptr_i32(t),
                                             ptr_i32(uo), ua, out), self.ctx.handle)
        return [_result_dict(out[i]) for i in range(n)]

    def close(self):
        if getattr(self, "_h", None):
            _capi.lib().ddp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------- facade of the reference library

class GPU_Graph(DeviceGraph):
    """LIB/include/graph.cuh:12-90: a graph copied to the device."""


class GPU_A_Star:
    """LIB/include/a_star.cuh:12-51: GPU_A_Star(graph, start, end, K),
    set_huiristics(hx), get_path() -> node list start..end (empty if
    unreachable)."""

    def __init__(self, graph, start, end, K):
        self.graph, self.start, self.end, self.K = graph, start, end, K
        self.cost = None
        self._heur = H_ZERO

    def set_huiristics(self, hx):
        self.graph.set_h(hx)
        self._heur = H_ARRAY

    def get_path(self):
        r = astar(self.graph.ctx, self.graph, self.start, self.end, heur=self._heur, K=self.K)
        self.cost = r.cost
        return [int(x) for x in r.path]
