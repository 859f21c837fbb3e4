"""ddp_bfs_hops_noalloc (the allocation-free reverse-BFS entry point) returns
exactly what ddp_bfs_hops returns, through the C ABI, on two calls in a row
(the second on a smaller graph, so leftovers of the first would show), and
refuses a graph larger than its static device memory."""
import ctypes
import os
from ctypes import byref, c_int32, c_void_p

import numpy as np
import pytest

import ddp_amd
from ddp_amd import _capi

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.skip("HIP runtime library not found")


class DevBuf:
    """Device memory owned by the test (the entry point under test allocates none)."""

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes, self.p = hip, nbytes, c_void_p()
        assert hip.hipMalloc(byref(self.p), ctypes.c_size_t(nbytes)) == 0

    def to_host_i32(self, n):
        out = np.empty(n, np.int32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(c_void_p), self.p, ctypes.c_size_t(4 * n), 2) == 0  # D2H
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(ctx, dg, t, reverse, scale, buf):
    lv = c_int32()
    rc = _capi.lib().ddp_bfs_hops_noalloc(ctx.handle, dg.handle, t, reverse, scale, buf.p, byref(lv))
    return rc, lv.value


def test_bfs_noalloc_matches_bfs_hops(ctx):
    hip = _hip()
    cases = [
        # (graph, reverse, scale, relabelled), the second one smaller
        (ddp_amd.HostGraph.random(60000, 5, 1, 40, 91), 1, 3, False),
        (ddp_amd.HostGraph.geometric(20000, 8, 100000, 92), 1, 1, True),
    ]
    buf = DevBuf(hip, 4 * 60000)
    try:
        for hg, reverse, scale, reorder in cases:
            dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=reorder)
            for t in (0, hg.N // 3, hg.N - 1):
                print("noalloc case N", hg.N, "t", t, "relabelled", reorder, flush=True)
                want, wl = dg.bfs_hops(t, reverse=bool(reverse), scale=scale)
                rc, lv = _noalloc(ctx, dg, t, reverse, scale, buf)
                assert rc == 0, rc
                got = buf.to_host_i32(hg.N)
                assert np.array_equal(got, want) and lv == wl
                assert got[t] == 0
            dg.close()
    finally:
        buf.free()


def test_bfs_noalloc_refuses_oversized_graph(ctx):
    hip = _hip()
    n = (1 << 21) + 1
    off = np.ones(n, np.int32)  # one edge 0 -> 1, every other row empty
    off[0] = 0
    col = np.array([1], np.int32)
    w = np.ones(1, np.uint32)
    dg = ddp_amd.DeviceGraph(ctx, ddp_amd.HostGraph(off, col, w), reverse=True)
    buf = DevBuf(hip, 4 * n)
    try:
        rc, _ = _noalloc(ctx, dg, 1, 1, 1, buf)
        assert rc == 10  # DDP_ERR_CAPACITY
        assert b"2097152" in _capi.lib().ddp_last_error(ctx.handle)
        # the ordinary entry point still answers it
        h, _ = dg.bfs_hops(1, reverse=True, scale=1)
        assert h[0] == 1 and h[1] == 0 and (h[2:] == -1).all()
    finally:
        buf.free()
        dg.close()
