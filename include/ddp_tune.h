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

/* Times the batched kernel (pile engine at blocks 256/512/1024, heap engine
 * at 512) on up to `sample` of the given queries (spread over the schedule
 * order) and writes the fastest options (block, engine) to *best; *qps gets
 * its measured queries/s. */
int ddp_tune_batch(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq, const int32_t *src,
                   const int32_t *dst, int32_t sample, ddp_batch_opts_t *best, float *qps);

/* Times single-query GA* in workgroup mode (K 256..1024), pile mode (K 512,
 * 1024) and grid mode (K 1024..8192, block 64/256) on up to `sample` of the
 * queries; writes the fastest options (K, block, mode) and its mean ms. */
int ddp_tune_astar(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq, const int32_t *src,
                   const int32_t *dst, int32_t sample, ddp_astar_opts_t *best, float *mean_ms);

#ifdef __cplusplus
}
#endif

#endif /* DDP_TUNE_H */
