/*
 * ddp_dastar.h -- dynamic A* (DA*): keep one query's answer exact while the
 * graph receives batches of edge inserts and deletes.
 *
 * Replaces the DA* driver loop of DA_star/DA_dyn_insert_del_nocycle.cu:
 * initial GA* (1388-1455), per Updates.txt batch the insert propagation
 * (propogateAdd / insert_propagate, 1578-1649), deletions (remove_edge,
 * propogateDel, fix_parent_cycle, delete_propagate, 1651-1814), heap repair and
 * resumed A* (1826-1977) and the CSR merge (2015-2102).
 *
 * Batch semantics are the reference's (DA_star/GA_merge.cu:1019-1089): every
 * delete removes the first live instance of u->v in u's row (no-op if none),
 * then inserts are appended to their source's row in line order.  After each
 * batch the returned cost equals a from-scratch search on the updated graph
 * (DDP_DASTAR_RECOMPUTE runs exactly that, as DA_star/GA_merge.cu does).
 * The heuristic must stay admissible for every graph version (for example
 * reverse-BFS hops x min weight computed on the union of all versions);
 * -1 entries are read as 0 (DA_star/DA_dyn_insert_del_nocycle.cu:1202-1210).
 */
#ifndef DDP_DASTAR_H
#define DDP_DASTAR_H

#include "ddp_capi.h"
#include "ddp_io.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddp_dastar ddp_dastar;

#define DDP_DASTAR_REPAIR 0     /* incremental repair of the retained search */
#define DDP_DASTAR_RECOMPUTE 1  /* full GA* from scratch on the updated graph */

typedef struct {
    int32_t inserts;
    int32_t deletes;            /* delete lines in the batch */
    int32_t deletes_effective;  /* deletes that matched a live edge */
    int32_t roots;              /* deleted edges that were parent edges */
    int32_t invalidated;        /* nodes whose value depended on a deleted edge */
    int32_t repulled;           /* invalidated nodes re-attached from live in-edges */
    int32_t pushes;             /* repair pushes handed to the GA* queues */
    int32_t iterations;         /* GA* iterations of the resumed search */
    int64_t expansions, edges, improvements;
    float host_ms;              /* host preparation: splitting and grouping the lines, upload */
    float repair_ms;            /* device time of the deletes, invalidation, re-pull and insert relax */
    float search_ms;            /* resumed (or recomputed) GA* kernel time */
    float total_ms;
    int32_t compacted;          /* 1 if the batch ended with a log compaction */
    float compact_ms;           /* its time (included in total_ms) */
    int64_t pile_refills;       /* pile engine: near-ring refills of the resumed search */
    int64_t pile_scanned;       /* pile engine: far-pile entries read by those refills */
} ddp_dastar_stats_t;

/* Uploads the graph (forward + transposed rows), runs the initial GA* from s
 * (opts as ddp_astar) and keeps the search state resident.  h: N ints or NULL
 * (h = 0).  *cost = Cx of t, DDP_INF_COST if unreachable. */
int ddp_dastar_open(ddp_ctx *ctx, int32_t N, int32_t E, const int32_t *off, const int32_t *col,
                    const uint32_t *w, const int32_t *h, int32_t s, int32_t t,
                    const ddp_astar_opts_t *opts, ddp_dastar **out, int32_t *cost);
/* Applies one update batch and returns the new cost of t. */
int ddp_dastar_apply(ddp_dastar *d, const ddp_update_t *lines, int32_t n, int32_t mode,
                     int32_t *cost, ddp_dastar_stats_t *stats);
/* Current path s..t (*len = 0 if unreachable, -len if cap is too small). */
int ddp_dastar_path(ddp_dastar *d, int32_t *path, int32_t cap, int32_t *len);
/* Folds the inserted-edge log back into the CSR rows on the device, as the
 * reference's merge_graph / mergeDiff rebuild merged rows after every batch
 * (DA_star/DA_dyn_insert_del_nocycle.cu:940-980, 2015-2102, 2236-2284;
 * DA_star/GA_merge.cu:1042-1089): each row becomes its surviving base edges in
 * row order followed by its live inserted edges in insert order, deleted
 * edges are dropped, the transposed rows are rebuilt and the log is emptied.
 * The graph version and the search state are unchanged, so costs, paths and
 * ddp_dastar_export read the same before and after.  ddp_dastar_apply also
 * compacts by itself after a batch that leaves more than the threshold of
 * ddp_dastar_set_compaction in the log (default max(65536, E/4) entries;
 * 0 turns that off). */
int ddp_dastar_compact(ddp_dastar *d);
int ddp_dastar_set_compaction(ddp_dastar *d, int64_t log_entries);
/* Live edge count of the current graph version. */
int ddp_dastar_num_edges(const ddp_dastar *d, int32_t *E);
/* Current graph in merged order (surviving base edges, then inserts, per
 * row): off has N entries, col/w E entries (E from ddp_dastar_num_edges). */
int ddp_dastar_export(const ddp_dastar *d, int32_t *off, int32_t *col, uint32_t *w);
void ddp_dastar_close(ddp_dastar *d);

#ifdef __cplusplus
}
#endif

#endif /* DDP_DASTAR_H */
