/*
 * ddp_capi.h -- C ABI of the MI355X (gfx950) parallel-A* library.
 *
 * Plain C types only: int32 node ids, uint32 weights, caller-owned host
 * buffers, opaque handles that own device memory.  Every entry point returns
 * an int status (DDP_OK = 0) and never exits the process; the message of the
 * last failure on a context is available from ddp_last_error().
 *
 * Graph layout accepted at the boundary is the reference's file layout
 * (GA_star/GenerateCSR.cpp:92-132, readers GA_star/GA_v2.cu:307-348):
 *   off[N]  row offsets, off[0] = 0, no sentinel (row N-1 ends at E)
 *   col[E]  target ids; a negative id is a deleted edge (GA_star/GA_v2.cu:122-125)
 *   w[E]    unsigned weights
 * Costs are the reference's Cx values: f = g + h of the target, int32,
 * INT_MAX (2147483647) when the target is unreachable
 * (DA_star/DA_dyn_insert_del_nocycle.cu:1469-1471).
 *
 * Threading: one host thread per context at a time.  Handles created from a
 * context use that context's HIP stream.
 */
#ifndef DDP_CAPI_H
#define DDP_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDP_OK            0
#define DDP_ERR_ARG       1  /* bad argument (null pointer, id out of range, ...) */
#define DDP_ERR_HIP       2  /* HIP runtime error */
#define DDP_ERR_NOMEM     3  /* device memory exhausted / admission refused */
#define DDP_ERR_OVERFLOW  4  /* a device queue or buffer hit its capacity */
#define DDP_ERR_IO        5  /* file could not be opened / written */
#define DDP_ERR_FORMAT    6  /* malformed or truncated input file */
#define DDP_ERR_TIMEOUT   7  /* device wait or iteration limit reached */
#define DDP_ERR_STATE     8  /* call not valid in the handle's current state */
#define DDP_ERR_NODEVICE  9  /* no usable GPU */
#define DDP_ERR_CAPACITY 10  /* a fixed capacity is used up and will not free itself
                                (unlike DDP_ERR_NOMEM, retrying does not help) */

#define DDP_INF_COST 2147483647

typedef struct ddp_ctx ddp_ctx;
typedef struct ddp_graph ddp_graph;

/* ------------------------------------------------------------------ context */

/* Replaces the implicit CUDA context + default stream every reference
 * driver uses (GA_star/GA_v2.cu:414-473).  Creates a HIP stream on `device`. */
int ddp_ctx_create(int device, ddp_ctx **out);
void ddp_ctx_destroy(ddp_ctx *ctx);
/* Replaces gpuErrchk's fprintf+exit (GA_star/GA_v2.cu:14-22): the text of the
 * last error on this context ("" if none). */
const char *ddp_last_error(const ddp_ctx *ctx);
/* Replaces the global error: message of the last ddp_ctx_create failure. */
const char *ddp_create_error(void);

typedef struct {
    int32_t compute_units;
    int32_t lds_bytes_per_cu;
    int32_t wave_size;
    int32_t l2_bytes;
    int64_t hbm_total_bytes;
    int64_t hbm_free_bytes;
    char arch[64];
} ddp_device_info_t;
int ddp_device_info(ddp_ctx *ctx, ddp_device_info_t *info);
/* Waits for all work on the context's device (the reference's
 * cudaDeviceSynchronize after each launch, GA_star/GA_v2.cu:495). */
int ddp_ctx_synchronize(ddp_ctx *ctx);
/* No reference counterpart: a plain device-to-device copy of `bytes` (a
 * kernel named copy_kernel, 16-byte loads and stores), timed `reps` times
 * with HIP events.  Gives the measured copy bandwidth the roofline is also
 * reported against, and a kernel of known traffic to scale HBM counters by.
 * *ms = best time, *gbps = 2*bytes / best time. */
int ddp_copy_bandwidth(ddp_ctx *ctx, int64_t bytes, int32_t reps, float *ms, float *gbps);

/* -------------------------------------------------------------------- graph */

#define DDP_GRAPH_REVERSE 1u  /* also keep the transposed CSR on the device */
#define DDP_GRAPH_REORDER 2u  /* relabel nodes for locality (see ddp_graph_upload_ex) */

/* Replaces GPU_Graph's constructor (LIB/include/graph.cuh:50-72) and the
 * cudaMalloc/cudaMemcpy block of every driver (GA_star/GA_v2.cu:414-447).
 * Copies the host CSR (reference layout above) to HBM. With
 * DDP_GRAPH_REVERSE the transposed CSR is built too, rows ordered as
 * DA_star/genOpCSR.cpp:55-74 writes them. */
int ddp_graph_upload(ddp_ctx *ctx, int32_t N, int32_t E, const int32_t *off,
                     const int32_t *col, const uint32_t *w, uint32_t flags,
                     ddp_graph **out);
/* ddp_graph_upload plus, optionally, the coordinates of ddp_graph_set_coords
 * (xy may be NULL).  With DDP_GRAPH_REORDER the device copy numbers the nodes
 * along a Hilbert curve through xy (breadth-first order without xy) so that
 * neighbours share cache lines; every entry point still takes and returns the
 * caller's node ids (queries, paths, parent arrays, h arrays, updates), the
 * translation happens at the boundary.  Not combinable with
 * ddp_graph_set_grid, whose heuristic reads coordinates from the id.  No
 * reference counterpart (its kernels index the file's ids directly). */
int ddp_graph_upload_ex(ddp_ctx *ctx, int32_t N, int32_t E, const int32_t *off,
                        const int32_t *col, const uint32_t *w, const int32_t *xy,
                        uint32_t flags, ddp_graph **out);
void ddp_graph_destroy(ddp_graph *g);
int ddp_graph_info(const ddp_graph *g, int32_t *N, int32_t *E,
                   uint32_t *min_weight, int32_t *max_out_degree);

/* Replaces A_Star::set_huiristics / GPU_A_Star::set_huiristics
 * (LIB/include/a_star.h:86-90, LIB/include/a_star.cu:96-105) and the Hx.txt
 * load (GA_star/GA_v2.cu:350-353): N int32 values, -1 = cannot reach the
 * target (GA_star/bfs_hx.cu:170); such nodes are never expanded. */
int ddp_graph_set_h(ddp_graph *g, const int32_t *h);
/* Integer plane coordinates (2*N int32, x then y per node) for the closed-form
 * Euclidean heuristic. No reference counterpart (the reference only reads
 * Hx.txt); used by geometric graphs. */
int ddp_graph_set_coords(ddp_graph *g, const int32_t *xy);
/* Row-major grid of the given width (node id = r*width + c) for the
 * closed-form Manhattan heuristic. No reference counterpart (maze/create_maze.cpp:128
 * writes a squared-Euclid Hx.txt instead, which is not admissible). */
int ddp_graph_set_grid(ddp_graph *g, int32_t width);

/* Heuristic selectors */
#define DDP_H_ZERO      0  /* h = 0 (Dijkstra order) */
#define DDP_H_ARRAY     1  /* the array given to ddp_graph_set_h */
#define DDP_H_MANHATTAN 2  /* scale * (|dr| + |dc|) on the grid of ddp_graph_set_grid */
#define DDP_H_EUCLID    3  /* scale * floor(sqrt(dx^2+dy^2)) on ddp_graph_set_coords */
typedef struct {
    int32_t kind;
    int32_t scale;   /* multiplier for closed forms (>= 1); ignored for ARRAY/ZERO */
} ddp_heur_t;

/* ------------------------------------------------------- reverse-BFS (bfs_hx) */

/* Replaces bfs_hx (GA_star/bfs_hx.cu:11-36 kernel, 111-138 host loop,
 * 158-172 output).  BFS from t; h_out[v] = hops(v) * scale, -1 if v is not
 * reached.  reverse = 0 walks g's forward rows (the reference's behaviour:
 * it is given the transposed file); reverse = 1 walks g's transposed rows
 * (needs DDP_GRAPH_REVERSE), i.e. hops from v to t in g.
 * If store_as_h != 0 the result also becomes g's DDP_H_ARRAY heuristic
 * without a host round trip.  h_out may be NULL.  *levels_out (optional) gets
 * the number of BFS levels. */
int ddp_bfs_hops(ddp_ctx *ctx, ddp_graph *g, int32_t t, int32_t reverse,
                 int32_t scale, int32_t store_as_h, int32_t *h_out,
                 int32_t *levels_out);

/* No reference counterpart.  ddp_bfs_hops for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): it calls no hipMalloc/hipFree.  Its
 * distance array, frontier buffers and control words live in a static device
 * array of 12 bytes per node for at most 2,097,152 nodes (24 MiB, held by
 * every GPU the library runs on until the process ends); a larger graph gets
 * DDP_ERR_CAPACITY.  h_out_dev: N int32 in DEVICE memory the caller owns,
 * written in the caller's node numbering (hops x scale, -1 = not reached).
 * No store_as_h.  All calls on one GPU share that array, so calls that
 * overlap on one GPU (several host threads or contexts) race; each call
 * clears what it needs before use and waits for its work before returning. */
int ddp_bfs_hops_noalloc(ddp_ctx *ctx, const ddp_graph *g, int32_t t,
                         int32_t reverse, int32_t scale, int32_t *h_out_dev,
                         int32_t *levels_out);

/* ------------------------------------------------------ single-query GA* */

typedef struct {
    int32_t K;          /* number of priority queues (one per lane); 0 = 1024 */
    int32_t block;      /* lanes per workgroup; grid mode: 64, 256 or 1024; 0 = auto
                           (grid mode: 64 below K = 4096, 256 below 65536, 1024 from there) */
    int32_t heap_cap;   /* entries per queue (LDS part + HBM spill); 0 = auto */
    int32_t lds_cap;    /* heap entries per queue kept in LDS; 0 = auto */
    int32_t max_iters;  /* iteration limit; 0 = auto (bounded) */
    int32_t mode;       /* DDP_ASTAR_AUTO / _GRID / _WORKGROUP / _PILES / _BIDI */
    int32_t reserved[2];
} ddp_astar_opts_t;

/* Execution modes of the single-query search.  GRID spreads the K queues over
 * K/block workgroups that meet at a device-wide barrier every iteration.
 * WORKGROUP keeps all K <= 1024 queues in one workgroup (block = K rounded up
 * to a power of two >= 64): no device-wide barrier, but one CU's worth of
 * memory parallelism.  PILES runs the search on one workgroup of K <= 1024
 * lanes with the near/far piles of the batched engine (DDP_BATCH_PILES)
 * instead of per-lane heaps.  AUTO selects GRID; ddp_tune_astar times all. */
#define DDP_ASTAR_AUTO      0
#define DDP_ASTAR_GRID      1
#define DDP_ASTAR_WORKGROUP 2
#define DDP_ASTAR_PILES     3
/* BIDI: the pile search of PILES run from both ends at once in one workgroup
 * of K (256..1024) lanes -- forward from s over the rows, backward from t
 * over the transposed rows, keys g +- (h_t - h_s) / 2 -- until no open entry
 * on either side can shorten the best meeting.  Needs DDP_GRAPH_REVERSE and a
 * closed-form heuristic (not DDP_H_ARRAY).  parent_out then holds the forward
 * search's parents, with the path's own predecessors from the meeting node
 * on to t. */
#define DDP_ASTAR_BIDI      4

typedef struct {
    int64_t expansions;     /* nodes popped and expanded */
    int64_t edges;          /* edges scanned (tombstones included) */
    int64_t improvements;   /* successful relaxations */
    int64_t stale_pops;     /* popped entries superseded by a better one */
    int32_t iterations;     /* GA* iterations (one pop per queue each) */
    int32_t K;              /* queues used */
    int32_t block;          /* lanes per workgroup used */
    int32_t blocks;         /* workgroups used */
    float kernel_ms;        /* search kernel time (HIP events, ctx stream) */
    float total_ms;         /* entry point wall time */
} ddp_astar_stats_t;

/* Replaces GPU_A_Star<T,U>(graph, start, end, K) + get_path()
 * (LIB/include/a_star.cuh:12-51, LIB/include/a_star.cu:41-69,109-303) and the
 * GA* loop of GA_star/GA_v2.cu:483-575 (extractMin / A_star_expand /
 * keepHeapPQ / setNV / insertPQ / checkMIN).
 * cost: Cx of t (DDP_INF_COST if unreachable).  parent_out (optional, N
 * entries): the parent array (-1 = none), as GA_v2.cu:562 copies it back.
 * path_out (optional, path_cap entries): the path s..t; *path_len gets its
 * length (0 if unreachable, -len if path_cap was too small). */
int ddp_astar(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t s,
              int32_t t, const ddp_astar_opts_t *opts, int32_t *cost,
              int32_t *parent_out, int32_t *path_out, int32_t path_cap,
              int32_t *path_len, ddp_astar_stats_t *stats);

#ifdef __cplusplus
}
#endif

#endif /* DDP_CAPI_H */
