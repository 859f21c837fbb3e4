"""ctypes binding of libddp.so (the C ABI declared in include/*.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C ddp_amd/csrc``).  Loading fails loudly when it is missing: there is
no fallback implementation of any entry point.
"""
import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libddp.so")

DDP_OK = 0
ERR_NAMES = {
    1: "DDP_ERR_ARG", 2: "DDP_ERR_HIP", 3: "DDP_ERR_NOMEM", 4: "DDP_ERR_OVERFLOW",
    5: "DDP_ERR_IO", 6: "DDP_ERR_FORMAT", 7: "DDP_ERR_TIMEOUT", 8: "DDP_ERR_STATE",
    9: "DDP_ERR_NODEVICE", 10: "DDP_ERR_CAPACITY",
}
INF_COST = 2147483647

H_ZERO, H_ARRAY, H_MANHATTAN, H_EUCLID = 0, 1, 2, 3
MODE_AUTO, MODE_GRID, MODE_WORKGROUP, MODE_PILES, MODE_BIDI = 0, 1, 2, 3, 4  # DDP_ASTAR_*
ENGINE_AUTO, ENGINE_HEAPS, ENGINE_PILES, ENGINE_BIDI = 0, 1, 2, 3  # DDP_BATCH_*
GRAPH_REVERSE = 1
GRAPH_REORDER = 2


class DDPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Heur(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("scale", c_int32)]


class AstarOpts(ctypes.Structure):
    _fields_ = [("K", c_int32), ("block", c_int32), ("heap_cap", c_int32), ("lds_cap", c_int32),
                ("max_iters", c_int32), ("mode", c_int32), ("reserved", c_int32 * 2)]


class AstarStats(ctypes.Structure):
    _fields_ = [("expansions", c_int64), ("edges", c_int64), ("improvements", c_int64),
                ("stale_pops", c_int64), ("iterations", c_int32), ("K", c_int32), ("block", c_int32),
                ("blocks", c_int32), ("kernel_ms", c_float), ("total_ms", c_float)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("compute_units", c_int32), ("lds_bytes_per_cu", c_int32), ("wave_size", c_int32),
                ("l2_bytes", c_int32), ("hbm_total_bytes", c_int64), ("hbm_free_bytes", c_int64),
                ("arch", c_char * 64)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("block", c_int32), ("slots", c_int32), ("heap_cap", c_int32), ("lds_cap", c_int32),
                ("max_iters", c_int32), ("path_cap", c_int32), ("schedule", c_int32), ("engine", c_int32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("expansions", c_int64), ("edges", c_int64), ("improvements", c_int64),
                ("stale_pops", c_int64), ("slots", c_int32), ("block", c_int32), ("errors", c_int32),
                ("kernel_ms", c_float), ("total_ms", c_float), ("pile_rebuilds", c_int64),
                ("pile_refills", c_int64), ("pile_scanned", c_int64)]


class Update(ctypes.Structure):
    _fields_ = [("flag", c_int32), ("u", c_int32), ("v", c_int32), ("w", c_uint32)]


class DAStats(ctypes.Structure):
    _fields_ = [("inserts", c_int32), ("deletes", c_int32), ("deletes_effective", c_int32), ("roots", c_int32),
                ("invalidated", c_int32), ("repulled", c_int32), ("pushes", c_int32), ("iterations", c_int32),
                ("expansions", c_int64), ("edges", c_int64), ("improvements", c_int64), ("host_ms", c_float),
                ("repair_ms", c_float), ("search_ms", c_float), ("total_ms", c_float), ("compacted", c_int32),
                ("compact_ms", c_float), ("pile_refills", c_int64), ("pile_scanned", c_int64)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("block", c_int32), ("slots", c_int32), ("heap_cap", c_int32), ("ring", c_int32),
                ("max_inserts", c_int32), ("idle_ms", c_int32), ("engine", c_int32), ("grid_k", c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", c_int64), ("grid_queries", c_int64), ("direct_updates", c_int64), ("launches", c_int32),
                ("compactions", c_int32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("tag", c_uint64), ("type", c_int32), ("cost", c_int32), ("version", c_int32), ("status", c_int32),
                ("submit_ms", c_double), ("done_ms", c_double), ("device_ms", c_double)]


P_I32 = POINTER(c_int32)
P_U32 = POINTER(c_uint32)
P_F64 = POINTER(c_double)
PP = POINTER(c_void_p)

# name -> (restype, argtypes); every symbol declared in include/*.h appears here.
SIGNATURES = {
    # ddp_capi.h
    "ddp_ctx_create": (c_int, [c_int, PP]),
    "ddp_ctx_destroy": (None, [c_void_p]),
    "ddp_last_error": (c_char_p, [c_void_p]),
    "ddp_create_error": (c_char_p, []),
    "ddp_device_info": (c_int, [c_void_p, POINTER(DeviceInfo)]),
    "ddp_ctx_synchronize": (c_int, [c_void_p]),
    "ddp_copy_bandwidth": (c_int, [c_void_p, c_int64, c_int32, POINTER(c_float), POINTER(c_float)]),
    "ddp_graph_upload": (c_int, [c_void_p, c_int32, c_int32, P_I32, P_I32, P_U32, c_uint32, PP]),
    "ddp_graph_upload_ex": (c_int, [c_void_p, c_int32, c_int32, P_I32, P_I32, P_U32, P_I32, c_uint32, PP]),
    "ddp_graph_destroy": (None, [c_void_p]),
    "ddp_graph_info": (c_int, [c_void_p, P_I32, P_I32, P_U32, P_I32]),
    "ddp_graph_set_h": (c_int, [c_void_p, P_I32]),
    "ddp_graph_set_coords": (c_int, [c_void_p, P_I32]),
    "ddp_graph_set_grid": (c_int, [c_void_p, c_int32]),
    "ddp_bfs_hops": (c_int, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, P_I32, P_I32]),
    "ddp_bfs_hops_noalloc": (c_int, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p, P_I32]),
    "ddp_astar": (c_int, [c_void_p, c_void_p, Heur, c_int32, c_int32, POINTER(AstarOpts), P_I32, P_I32,
                          P_I32, c_int32, P_I32, POINTER(AstarStats)]),
    # ddp_batch.h
    "ddp_batch_create": (c_int, [c_void_p, c_void_p, Heur, c_int32, POINTER(BatchOpts), PP]),
    "ddp_batch_destroy": (None, [c_void_p]),
    "ddp_batch_info": (c_int, [c_void_p, P_I32, P_I32, P_I32, P_I32]),
    "ddp_batch_upload": (c_int, [c_void_p, c_int32, P_I32, P_I32]),
    "ddp_batch_launch": (c_int, [c_void_p]),
    "ddp_batch_wait": (c_int, [c_void_p, POINTER(BatchStats)]),
    "ddp_batch_download": (c_int, [c_void_p, P_I32, P_I32, P_I32, P_U32, P_F64]),
    "ddp_astar_batch": (c_int, [c_void_p, c_void_p, Heur, c_int32, P_I32, P_I32, POINTER(BatchOpts), P_I32,
                                P_I32, P_I32, POINTER(BatchStats)]),
    # ddp_dastar.h
    "ddp_dastar_open": (c_int, [c_void_p, c_int32, c_int32, P_I32, P_I32, P_U32, P_I32, c_int32, c_int32,
                                POINTER(AstarOpts), PP, P_I32]),
    "ddp_dastar_apply": (c_int, [c_void_p, POINTER(Update), c_int32, c_int32, P_I32, POINTER(DAStats)]),
    "ddp_dastar_path": (c_int, [c_void_p, P_I32, c_int32, P_I32]),
    "ddp_dastar_num_edges": (c_int, [c_void_p, P_I32]),
    "ddp_dastar_compact": (c_int, [c_void_p]),
    "ddp_dastar_set_compaction": (c_int, [c_void_p, c_int64]),
    "ddp_dastar_export": (c_int, [c_void_p, P_I32, P_I32, P_U32]),
    "ddp_dastar_close": (None, [c_void_p]),
    # ddp_server.h
    "ddp_server_create": (c_int, [c_void_p, c_void_p, Heur, POINTER(ServerOpts), PP]),
    "ddp_server_info": (c_int, [c_void_p, P_I32, P_I32, P_I32]),
    "ddp_server_stats": (c_int, [c_void_p, c_void_p]),
    "ddp_server_submit_query": (c_int, [c_void_p, c_int32, c_int32, c_uint64]),
    "ddp_server_submit_update": (c_int, [c_void_p, POINTER(Update), c_int32, c_uint64]),
    "ddp_server_poll": (c_int, [c_void_p, POINTER(ServerResult), c_int32, P_I32]),
    "ddp_server_drain": (c_int, [c_void_p, c_int32]),
    "ddp_server_compact": (c_int, [c_void_p]),
    "ddp_server_destroy": (None, [c_void_p]),
    "ddp_server_replay": (c_int, [c_void_p, c_int32, P_F64, P_I32, P_I32, P_I32, P_I32, POINTER(Update),
                                  POINTER(ServerResult)]),
    # ddp_tune.h
    "ddp_tune_batch": (c_int, [c_void_p, c_void_p, Heur, c_int32, P_I32, P_I32, c_int32, POINTER(BatchOpts),
                               POINTER(c_float)]),
    "ddp_tune_astar": (c_int, [c_void_p, c_void_p, Heur, c_int32, P_I32, P_I32, c_int32, POINTER(AstarOpts),
                               POINTER(c_float)]),
    # ddp_io.h
    "ddp_hgraph_from_arrays": (c_int, [c_int32, c_int32, P_I32, P_I32, P_U32, PP]),
    "ddp_hgraph_free": (None, [c_void_p]),
    "ddp_hgraph_view": (c_int, [c_void_p, P_I32, P_I32, POINTER(P_I32), POINTER(P_I32), POINTER(P_U32)]),
    "ddp_hgraph_coords": (P_I32, [c_void_p]),
    "ddp_hgraph_set_coords": (c_int, [c_void_p, P_I32]),
    "ddp_read_graph_txt": (c_int, [c_char_p, PP, c_char_p, c_int]),
    "ddp_read_graph_cached": (c_int, [c_char_p, c_char_p, PP, c_char_p, c_int]),
    "ddp_write_graph_txt": (c_int, [c_char_p, c_void_p]),
    "ddp_read_hx_txt": (c_int, [c_char_p, c_int32, P_I32, c_char_p, c_int]),
    "ddp_write_hx_txt": (c_int, [c_char_p, c_int32, P_I32]),
    "ddp_hgraph_transpose": (c_int, [c_void_p, PP]),
    "ddp_read_updates_txt": (c_int, [c_char_p, PP, c_char_p, c_int]),
    "ddp_write_updates_txt": (c_int, [c_char_p, c_void_p]),
    "ddp_updates_count": (c_int, [c_void_p]),
    "ddp_updates_batch": (c_int, [c_void_p, c_int32, POINTER(POINTER(Update)), P_I32]),
    "ddp_updates_free": (None, [c_void_p]),
    "ddp_hgraph_apply_batch": (c_int, [c_void_p, POINTER(Update), c_int32, PP, P_I32]),
    "ddp_gen_random": (c_int, [c_int32, c_int32, c_uint32, c_uint32, c_uint64, PP]),
    "ddp_gen_grid_maze": (c_int, [c_int32, c_int32, c_int32, c_uint64, PP]),
    "ddp_gen_geometric": (c_int, [c_int32, c_double, c_int32, c_uint64, PP]),
    "ddp_gen_rmat": (c_int, [c_int32, c_int64, c_double, c_double, c_double, c_uint32, c_uint32, c_uint64, PP]),
    "ddp_gen_updates": (c_int, [c_void_p, c_int32, c_int32, c_int32, c_uint32, c_uint32, c_uint64, PP]),
    "ddp_gen_geo_pairs": (c_int, [c_void_p, c_int32, c_double, c_double, c_uint64, P_I32, P_I32]),
    "ddp_gen_reachable_pairs": (c_int, [c_void_p, c_int32, c_uint64, P_I32, P_I32]),
    "ddp_gen_hub_pairs": (c_int, [c_void_p, c_int32, c_uint64, P_I32, P_I32]),
}

_lib = None


def lib():
    """Loads libddp.so once; raises ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                              f"or `make -C ddp_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr_i32(a):
    return None if a is None else a.ctypes.data_as(P_I32)


def ptr_u32(a):
    return None if a is None else a.ctypes.data_as(P_U32)


def ptr_f64(a):
    return None if a is None else a.ctypes.data_as(P_F64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)
