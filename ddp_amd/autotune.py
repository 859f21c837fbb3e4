"""Launch-shape autotuning (ddp_tune_batch / ddp_tune_astar in libddp.so).

The tuner times candidate shapes on a sample of the caller's own queries and
caches the winner per graph and device; these wrappers return plain dicts.
"""
from ctypes import byref, c_float

from . import _capi
from ._capi import AstarOpts, BatchOpts, Heur, as_i32, ptr_i32


def _check(rc, ctx):
    if rc:
        raise _capi.DDPError(rc, _capi.lib().ddp_last_error(ctx.handle).decode())


def tune_batch(ctx, dg, src, dst, heur, scale=1, sample=1024):
    src, dst = as_i32(src), as_i32(dst)
    o = BatchOpts()
    q = c_float()
    _check(_capi.lib().ddp_tune_batch(ctx.handle, dg.handle, Heur(heur, scale), src.shape[0], ptr_i32(src),
                                      ptr_i32(dst), sample, byref(o), byref(q)), ctx)
    return {"block": o.block, "engine": o.engine, "heap_cap": o.heap_cap, "lds_cap": o.lds_cap,
            "queries_per_s": q.value}


def pick_batch_block(ctx, dg, src, dst, heur, scale=1, sample=1024):
    return tune_batch(ctx, dg, src, dst, heur, scale, sample)["block"]


def tune_astar(ctx, dg, src, dst, heur, scale=1, sample=16):
    src, dst = as_i32(src), as_i32(dst)
    o = AstarOpts()
    ms = c_float()
    _check(_capi.lib().ddp_tune_astar(ctx.handle, dg.handle, Heur(heur, scale), src.shape[0], ptr_i32(src),
                                      ptr_i32(dst), sample, byref(o), byref(ms)), ctx)
    return {"K": o.K, "block": o.block, "mode": o.mode, "heap_cap": o.heap_cap, "lds_cap": o.lds_cap,
            "mean_ms": ms.value}
