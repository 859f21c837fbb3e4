"""The autotuner (ddp_tune_batch / ddp_tune_astar) varies block, engine and
heap depth (heap_cap, lds_cap); whatever it picks returns the same costs as
the library's defaults, and its cache entry names and stores every option."""
import os

import numpy as np
import pytest

import ddp_amd
from ddp_amd import autotune

pytestmark = pytest.mark.gpu


def test_tuned_batch_options_keep_costs(ctx, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("DDP_TUNE_CACHE", str(cache))
    hg = ddp_amd.HostGraph.geometric(200000, 10, 100000, 4)  # C4-shaped, smaller
    src, dst = hg.geo_pairs(256, 100, 100000, 4)
    dg = ddp_amd.DeviceGraph(ctx, hg, reverse=True, reorder=True)
    best = autotune.tune_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, sample=256)
    print("tuned", best, flush=True)
    assert best["engine"] in (ddp_amd.ENGINE_PILES, ddp_amd.ENGINE_HEAPS, ddp_amd.ENGINE_BIDI)
    assert best["queries_per_s"] > 0
    base, _, _ = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID)
    tuned, _, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=best["block"],
                                       engine=best["engine"], heap_cap=best["heap_cap"], lds_cap=best["lds_cap"])
    assert st["errors"] == 0 and np.array_equal(base, tuned)
    # every heap-depth candidate returns the same costs too
    for cap, lcap in ((0, 8), (4096, 16)):
        c, _, st = ddp_amd.astar_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, block=512,
                                       engine=ddp_amd.ENGINE_HEAPS, heap_cap=cap, lds_cap=lcap)
        assert st["errors"] == 0 and np.array_equal(base, c)
    line = [l for l in cache.read_text().splitlines() if l.startswith("batch:block,engine,heap_cap,lds_cap:")]
    assert len(line) == 1
    vals = [int(x) for x in line[0].split()[1:]]
    assert vals[:4] == [best["block"], best["engine"], best["heap_cap"], best["lds_cap"]]
    again = autotune.tune_batch(ctx, dg, src, dst, heur=ddp_amd.H_EUCLID, sample=256)  # from the cache
    assert {k: again[k] for k in ("block", "engine", "heap_cap", "lds_cap")} == \
        {k: best[k] for k in ("block", "engine", "heap_cap", "lds_cap")}
    dg.close()


def test_tuned_astar_options_keep_costs(ctx, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("DDP_TUNE_CACHE", str(cache))
    hg = ddp_amd.HostGraph.random(30000, 6, 1, 100, 21)
    src, dst = hg.reachable_pairs(4, 22)
    dg = ddp_amd.DeviceGraph(ctx, hg)
    best = autotune.tune_astar(ctx, dg, src, dst, heur=ddp_amd.H_ZERO, sample=4)
    print("tuned", best, flush=True)
    for s, t in zip(src, dst):
        r0 = ddp_amd.astar(ctx, dg, int(s), int(t), heur=ddp_amd.H_ZERO)
        r1 = ddp_amd.astar(ctx, dg, int(s), int(t), heur=ddp_amd.H_ZERO, K=best["K"], block=best["block"],
                           mode=best["mode"], heap_cap=best["heap_cap"], lds_cap=best["lds_cap"])
        assert r0.cost == r1.cost
    line = [l for l in cache.read_text().splitlines() if l.startswith("astar:K,block,mode,heap_cap,lds_cap:")]
    assert len(line) == 1 and len(line[0].split()) == 7
    dg.close()
