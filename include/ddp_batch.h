/*
 * ddp_batch.h -- batched A* (many source/target queries in one launch).
 *
 * No reference counterpart exists: the reference answers one query per
 * process (GA_star/GA_v2.cu:298-577) and its applications loop over queries
 * on the host (WSN/wsn_2011.cu:355-456, k-short/GA_merge.cu:702-731).  A
 * maintainer replacing such a host loop binds these entry points; each query
 * returns exactly what ddp_astar returns for it alone.
 */
#ifndef DDP_BATCH_H
#define DDP_BATCH_H

#include "ddp_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddp_batch ddp_batch;

typedef struct {
    int32_t block;      /* lanes per query workgroup: 64..1024; 0 = 1024 (piles), 256 (heaps) */
    int32_t slots;      /* concurrent queries; 0 = what fits (occupancy and free HBM) */
    int32_t heap_cap;   /* entries per queue (LDS part + HBM spill); 0 = 2048 (heaps) or
                           max(2048, 3N / block) (piles).  A slot's queue memory is
                           heap_cap * block entries in either engine */
    int32_t lds_cap;    /* heap entries per queue kept in LDS; 0 = auto */
    int32_t max_iters;  /* per-query iteration limit; 0 = auto */
    int32_t path_cap;   /* path entries kept per query (0 = costs only) */
    int32_t schedule;   /* 0 = longest straight-line estimate first, 1 = given order */
    int32_t engine;     /* DDP_BATCH_AUTO / _HEAPS / _PILES */
} ddp_batch_opts_t;

/* Queue discipline inside a query slot.  HEAPS: one binary heap per lane (the
 * GA* structure; an iteration pops one node per lane).  PILES: a near ring of
 * entries under a moving f threshold and a far pile for the rest, both in HBM;
 * a round expands up to one near entry per lane and no heap is ever sifted.
 * Both return exact costs.  AUTO = PILES. */
#define DDP_BATCH_AUTO  0
#define DDP_BATCH_HEAPS 1
#define DDP_BATCH_PILES 2
/* BIDI: the pile engine run from both ends at once in one workgroup, a
 * forward search from s over the rows and a backward one from t over the
 * transposed rows, meeting in the middle (about half the rounds per query).
 * Needs a graph uploaded with DDP_GRAPH_REVERSE, a closed-form heuristic
 * (DDP_H_ZERO, _MANHATTAN or _EUCLID; the keys use the average potential
 * (h_t - h_s) / 2) and a block of 256, 512 or 1024 lanes; a slot then holds
 * two (f, parent) arrays and twice the queue memory.  Same exact costs. */
#define DDP_BATCH_BIDI  3

typedef struct {
    int64_t expansions, edges, improvements, stale_pops;
    int32_t slots, block;
    int32_t errors;     /* queries whose flags are non-zero */
    float kernel_ms;    /* launch-to-finish device time (HIP events on the ctx stream) */
    float total_ms;
    int64_t pile_rebuilds;  /* pile engine: far piles rebuilt after an overflow */
    int64_t pile_refills;   /* pile engine: near-ring refills from the far pile */
    int64_t pile_scanned;   /* pile engine: far-pile entries read by those refills */
} ddp_batch_stats_t;

/* Plan: allocates the query slots ((f,parent) arrays, touched logs, heaps) once
 * and clears them; slot count is limited by occupancy and by free HBM. */
int ddp_batch_create(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t max_queries,
                     const ddp_batch_opts_t *opts, ddp_batch **out);
void ddp_batch_destroy(ddp_batch *b);
int ddp_batch_info(const ddp_batch *b, int32_t *slots, int32_t *block, int32_t *heap_cap,
                   int32_t *lds_cap);
/* Copies the query arrays to HBM and fixes the schedule order.  With
 * DDP_H_ARRAY every query must have the target the array was computed for:
 * DDP_ERR_ARG if ddp_bfs_hops(store_as_h) computed it for another one (an
 * array given to ddp_graph_set_h carries no target; the caller keeps this
 * rule). */
int ddp_batch_upload(ddp_batch *b, int32_t nq, const int32_t *src, const int32_t *dst);
/* Asynchronous: replays the captured hipGraph (counter reset + persistent
 * search kernel) on the context stream. */
int ddp_batch_launch(ddp_batch *b);
int ddp_batch_wait(ddp_batch *b, ddp_batch_stats_t *stats);
/* Per query: cost (Cx of t, DDP_INF_COST if unreachable), path length
 * (negative if longer than path_cap), path s..t (nq * path_cap entries),
 * error flags (1 = heap_cap exceeded, 2 = iteration limit, 4 = bad id,
 * 8 = broken parent chain), completion time in ms since the first query
 * started.  Any pointer may be NULL. */
int ddp_batch_download(ddp_batch *b, int32_t *costs, int32_t *path_len, int32_t *paths,
                       uint32_t *qerr, double *latency_ms);
/* One-shot convenience: create + upload + launch + wait + download + destroy.
 * Fails (DDP_ERR_OVERFLOW / DDP_ERR_TIMEOUT / DDP_ERR_ARG) if any query failed. */
int ddp_astar_batch(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq,
                    const int32_t *src, const int32_t *dst, const ddp_batch_opts_t *opts,
                    int32_t *costs, int32_t *path_len, int32_t *paths,
                    ddp_batch_stats_t *stats);

#ifdef __cplusplus
}
#endif

#endif /* DDP_BATCH_H */
