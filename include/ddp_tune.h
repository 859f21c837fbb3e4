/*
 * ddp_tune.h -- picks GA* launch shapes per graph by timing candidates.
 *
 * No reference counterpart: the reference hard-codes 512-thread blocks
 * (GA_star/GA_v2.cu:483) and takes K from stdin.  Results are cached per
 * (graph size, degree, weights, heuristic, device) in a text file
 * ($DDP_TUNE_CACHE, default $HOME/.cache/ddp_amd/tune.txt) when writable.
 */
#ifndef DDP_TUNE_H
#define DDP_TUNE_H

#include "ddp_batch.h"
#include "ddp_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Times the batched kernel on up to `sample` of the given queries (spread
 * over the schedule order): the pile engine at blocks 256/512/1024 and, at
 * 1024, with half its default queue memory; the heap engine at block 512 with
 * 8 or 16 LDS levels per lane and the default or 4096 entries per lane (heap
 * depth); the bidirectional engine at 512/1024 when the heuristic is a closed
 * form and the graph has transposed rows.  Writes the fastest options (block,
 * engine, heap_cap, lds_cap; 0 = default) to *best; *qps gets its measured
 * queries/s.  Cached under a key naming every option it varies. */
int ddp_tune_batch(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq, const int32_t *src,
                   const int32_t *dst, int32_t sample, ddp_batch_opts_t *best, float *qps);

/* Times single-query GA* in workgroup mode (K 256..1024), pile mode (K 512,
 * 1024), bidirectional mode (K 256..1024; closed-form heuristics, graphs with
 * transposed rows) and grid mode (K 1024..8192 at block 64/256, 16384 and
 * 32768 at block 256/1024, 65536 at block 1024) on up to `sample`
 * of the queries, then the winner with 8, 16 and 32 LDS heap levels per queue
 * (grid and workgroup modes); writes the fastest options (K, block, mode,
 * heap_cap, lds_cap) and its mean ms. */
int ddp_tune_astar(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq, const int32_t *src,
                   const int32_t *dst, int32_t sample, ddp_astar_opts_t *best, float *mean_ms);

#ifdef __cplusplus
}
#endif

#endif /* DDP_TUNE_H */
