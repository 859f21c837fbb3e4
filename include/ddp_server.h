/*
 * ddp_server.h -- online A* serving on one GPU with continuous batching.
 *
 * No reference counterpart: the reference answers one query per process run
 * (GA_star/GA_v2.cu:298-577) and applies update batches in a blocking host
 * loop (DA_star/DA_dyn_insert_del_nocycle.cu:1490-2111).  A request loop that
 * replaces those binds these entry points.
 *
 * A persistent kernel keeps the graph and every in-flight query resident in
 * HBM.  Requests (point-to-point queries and edge-update batches) are
 * appended to a ring in pinned host memory; each workgroup slot claims the
 * next request as soon as it is free, so requests that arrive while others
 * run are admitted as slots free up and overlap on the device.  Slot count is
 * fixed at creation from free HBM (each slot owns an N-entry (f, parent) array
 * and its heaps).
 *
 * Graph versions: update batch number k (in submission order) creates
 * version k; a query submitted after k batches sees exactly version k.  Edges
 * carry the version that deleted them and inserted edges the version that
 * added them, so a query's answer depends only on (s, t, version), never on
 * what else is running.  Update semantics per batch are the reference's
 * (DA_star/GA_merge.cu:1019-1089): each delete removes the first instance of
 * u->v still live at the previous version; inserts are appended.
 */
#ifndef DDP_SERVER_H
#define DDP_SERVER_H

#include "ddp_capi.h"
#include "ddp_io.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddp_server ddp_server;

typedef struct {
    int32_t block;          /* lanes per query slot: 64..1024; 0 = 1024 (piles), 256 (heaps) */
    int32_t slots;          /* 0 = as many as occupancy and free HBM allow */
    int32_t heap_cap;       /* queue entries per lane (a slot holds heap_cap * block);
                               0 = 2048 (heaps) or max(2048, 3N / block) (piles) */
    int32_t ring;           /* request ring entries (power of two); 0 = 65536 */
    int32_t max_inserts;    /* capacity of the inserted-edge log; 0 = 1<<20.  An update
                               that does not fit triggers ddp_server_compact first */
    int32_t idle_ms;        /* the kernel exits once a slot has waited this long for work
                               while no slot runs a request; 0 = 2000 */
    int32_t engine;         /* DDP_SERVER_AUTO / _HEAPS / _PILES (as DDP_BATCH_*) */
    int32_t grid_k;         /* 0 = off.  K > 0: low-load queries are searched in grid mode
                               (ddp_astar's DDP_ASTAR_GRID, K lane heaps over the whole GPU)
                               instead of one slot: a query goes there while no update is
                               in flight on the slot path, if nothing else is in flight,
                               or, while grid mode alone keeps up with submissions (a
                               moving average of grid-mode queries' device times against
                               the mean gap between submissions), if grid queries are
                               queued or the youngest running slot query is older than
                               one grid-mode query takes.  Grid queries run one after
                               another with the GPU to themselves, after the slot queries
                               already running; requests sent to slots meanwhile start
                               after them.  Worth it where
                               one query is a large share of the graph (a full search of a
                               power-law graph), which alone on the GPU finishes many
                               times sooner than in one slot */
} ddp_server_opts_t;

/* Queue discipline of a query slot, as for batches (include/ddp_batch.h):
 * per-lane heaps, or near/far piles.  AUTO = PILES. */
#define DDP_SERVER_AUTO  0
#define DDP_SERVER_HEAPS 1
#define DDP_SERVER_PILES 2
/* BIDI: each query searched from both ends at once (DDP_BATCH_BIDI) on the
 * versioned graph: transposed rows carry their own delete versions and
 * inserted edges are listed by target too.  Needs DDP_GRAPH_REVERSE and a
 * block of 256, 512 or 1024 lanes; a slot holds two (f, parent) arrays. */
#define DDP_SERVER_BIDI  3

#define DDP_REQ_QUERY 0
#define DDP_REQ_UPDATE 1

typedef struct {
    uint64_t tag;           /* caller's id */
    int32_t type;           /* DDP_REQ_QUERY / DDP_REQ_UPDATE */
    int32_t cost;           /* queries: Cx of t, DDP_INF_COST if unreachable */
    int32_t version;        /* graph version the request saw (queries) or made (updates) */
    int32_t status;         /* 0 ok; 1 heap_cap exceeded; 2 iteration limit; 4 bad id;
                               8 a grid-mode search's device-wide barrier wait ran out;
                               16 an update's wait for the previous version ran out;
                               32 an update carried an invalid log entry.  16 and 32
                               are sticky: once an update fails, every later request
                               reports the same bits and is not run */
    double submit_ms;       /* host clock (ms since ddp_server_create) */
    double done_ms;         /* host clock when the result was collected */
    double device_ms;       /* device time from claim to completion */
} ddp_server_result_t;

int ddp_server_create(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, const ddp_server_opts_t *opts,
                      ddp_server **out);
int ddp_server_info(const ddp_server *sv, int32_t *slots, int32_t *block, int32_t *version);

typedef struct {
    int64_t requests;       /* submitted so far */
    int64_t grid_queries;   /* of them, queries searched in grid mode (grid_k) */
    int64_t direct_updates; /* of them, updates applied by a kernel of their own (grid_k):
                               no update pending on the slot path and no open launch
                               running slot queries */
    int32_t launches;       /* serve-kernel launches (the first request, and relaunches
                               after an idle exit, a grid-mode query or a compaction) */
    int32_t compactions;    /* inserted-edge log compactions */
} ddp_server_stats_t;
int ddp_server_stats(const ddp_server *sv, ddp_server_stats_t *out);
/* Non-blocking; DDP_ERR_NOMEM if the ring (or, for updates, the payload arena)
 * is full: try again after polling.  An update whose inserts do not fit in
 * the rest of the inserted-edge log first compacts it (ddp_server_compact,
 * which blocks); DDP_ERR_CAPACITY if the batch alone exceeds the log, which
 * polling does not cure. */
int ddp_server_submit_query(ddp_server *sv, int32_t s, int32_t t, uint64_t tag);
int ddp_server_submit_update(ddp_server *sv, const ddp_update_t *lines, int32_t n, uint64_t tag);
/* Folds the inserted-edge log back into CSR rows at the current version, as
 * the reference's merge_graph / mergeDiff rebuild merged rows
 * (DA_star/DA_dyn_insert_del_nocycle.cu:940-980, 2236-2284;
 * DA_star/GA_merge.cu:1042-1089): each row becomes its live base edges in row
 * order followed by its live inserted edges in insert order, deleted edges
 * are dropped and the log is emptied, so later deletes still take the first
 * live instance of the merged row.  Waits until every submitted request has
 * finished and stops the kernel first (queries read the log); results stay
 * available to ddp_server_poll.  Requests submitted afterwards relaunch it. */
int ddp_server_compact(ddp_server *sv);
/* Collects up to max finished requests in completion order of submission. */
int ddp_server_poll(ddp_server *sv, ddp_server_result_t *out, int32_t max, int32_t *n);
/* Waits until every submitted request has finished (bounded by timeout_ms). */
int ddp_server_drain(ddp_server *sv, int32_t timeout_ms);
void ddp_server_destroy(ddp_server *sv);

/* Open-loop replay: request i is submitted at arrival_ms[i] after the start
 * (queries: s[i], t[i]; updates: type[i] = DDP_REQ_UPDATE with lines
 * upd[upd_off[i] .. upd_off[i+1])), results are written to out[i], with
 * submit_ms = arrival_ms[i] (so done_ms - submit_ms includes any wait before
 * the request entered the ring) and done_ms relative to the start.  Only
 * DDP_ERR_NOMEM is retried (after polling); any other failure, such as
 * DDP_ERR_CAPACITY, is returned. */
int ddp_server_replay(ddp_server *sv, int32_t n, const double *arrival_ms, const int32_t *type,
                      const int32_t *s, const int32_t *t, const int32_t *upd_off,
                      const ddp_update_t *upd, ddp_server_result_t *out);

#ifdef __cplusplus
}
#endif

#endif /* DDP_SERVER_H */
