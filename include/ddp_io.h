/*
 * ddp_io.h -- host-side graph files, update batches and seeded generators.
 *
 * Text formats are the reference's (one contract for every driver):
 *   graph.txt / graph_op.txt / graph_cg.txt: "N E", E column ids, N row
 *     offsets (no sentinel), E weights (GA_star/GenerateCSR.cpp:92-132,
 *     readers GA_star/A_star_cpu.cpp:105-140, GA_star/GA_v2.cu:307-348)
 *   Hx.txt: N ints, -1 = cannot reach the target (GA_star/bfs_hx.cu:158-172)
 *   Updates.txt: blocks of "c" then c lines "flag u v w", flag 1 = insert,
 *     0 = delete (DA_star/genDynamicGraph.cpp:136,170-171, readers
 *     DA_star/DA_dyn_insert_del_nocycle.cu:1490-1522)
 * All host graphs use the reference layout (off has N entries).
 */
#ifndef DDP_IO_H
#define DDP_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddp_hgraph ddp_hgraph;

/* Host CSR owned by the library. */
int ddp_hgraph_from_arrays(int32_t N, int32_t E, const int32_t *off, const int32_t *col,
                           const uint32_t *w, ddp_hgraph **out);
void ddp_hgraph_free(ddp_hgraph *g);
int ddp_hgraph_view(const ddp_hgraph *g, int32_t *N, int32_t *E, const int32_t **off,
                    const int32_t **col, const uint32_t **w);
/* Coordinates (2*N ints) of generated geometric graphs, else NULL. */
const int32_t *ddp_hgraph_coords(const ddp_hgraph *g);
/* Attaches (copies) 2*N coordinates to a host graph. */
int ddp_hgraph_set_coords(ddp_hgraph *g, const int32_t *xy);

/* graph.txt reader (replaces the fscanf loops of GA_star/GA_v2.cu:307-348;
 * parallel parse of an mmap'd file).  err receives a message on failure. */
int ddp_read_graph_txt(const char *path, ddp_hgraph **out, char *err, int errlen);
/* ddp_read_graph_txt through a binary cache at cache_path (NULL: path +
 * ".ddpbin").  A cache written for the same text file (its size and
 * modification time are in the cache header) is read instead of parsing;
 * otherwise the text is parsed and the cache rewritten (best effort).  No
 * reference counterpart: the reference parses the text on every run
 * (GA_star/GA_v2.cu:307-349). */
int ddp_read_graph_cached(const char *path, const char *cache_path, ddp_hgraph **out, char *err,
                          int errlen);
/* graph.txt writer with the reference's token layout (GenerateCSR.cpp:92-132). */
int ddp_write_graph_txt(const char *path, const ddp_hgraph *g);
/* Hx.txt: reads exactly n ints (n = N). */
int ddp_read_hx_txt(const char *path, int32_t n, int32_t *h, char *err, int errlen);
/* Hx.txt writer, "%d " per node as GA_star/bfs_hx.cu:158-172. */
int ddp_write_hx_txt(const char *path, int32_t n, const int32_t *h);

/* Transposed CSR in DA_star/genOpCSR.cpp:55-74 order (graph_op.txt). */
int ddp_hgraph_transpose(const ddp_hgraph *g, ddp_hgraph **out);

/* Update batches. */
typedef struct ddp_updates ddp_updates;
typedef struct {
    int32_t flag;  /* 1 = insert, 0 = delete */
    int32_t u, v;
    uint32_t w;    /* ignored for deletes */
} ddp_update_t;
int ddp_read_updates_txt(const char *path, ddp_updates **out, char *err, int errlen);
int ddp_write_updates_txt(const char *path, const ddp_updates *u);
int ddp_updates_count(const ddp_updates *u);  /* number of batches */
/* Pointer to batch b's lines and their count. */
int ddp_updates_batch(const ddp_updates *u, int32_t b, const ddp_update_t **lines, int32_t *n);
void ddp_updates_free(ddp_updates *u);

/* G_{b+1} = G_b minus the first live instance of every deleted (u,v), then
 * inserts appended per source in line order (DA_star/GA_merge.cu:1019-1089).
 * Host implementation used by the drivers' recompute mode and generators. */
int ddp_hgraph_apply_batch(const ddp_hgraph *g, const ddp_update_t *lines, int32_t n,
                           ddp_hgraph **out, int32_t *n_deleted);

/* ---- seeded generators (std::mt19937_64 + explicit integer sampling) ---- */

/* Each node gets `deg` out-edges to targets uniform in [0,N) (self-loops
 * rejected, duplicates kept), weights uniform in [wmin, wmax]. */
int ddp_gen_random(int32_t N, int32_t deg, uint32_t wmin, uint32_t wmax, uint64_t seed,
                   ddp_hgraph **out);
/* width x height 4-connected grid, cells blocked with probability
 * p_block_permille/1000, plus a carved monotone open corridor from (0,0) to
 * (height-1,width-1); edge order up, down, right, left as
 * maze/create_maze.cpp:93-111; unit weights. */
int ddp_gen_grid_maze(int32_t width, int32_t height, int32_t p_block_permille, uint64_t seed,
                      ddp_hgraph **out);
/* N points uniform on the integer lattice [0,scale]^2; edges both ways
 * between points with dx^2+dy^2 <= R^2, R^2 chosen for the mean degree;
 * w = max(1, ceil(sqrt(d2))).  Coordinates available via ddp_hgraph_coords. */
int ddp_gen_geometric(int32_t N, double mean_degree, int32_t scale, uint64_t seed,
                      ddp_hgraph **out);
/* R-MAT: 2^scale nodes, E edges, quadrant probabilities a,b,c (d = 1-a-b-c),
 * duplicates and self-loops kept, weights uniform in [wmin, wmax]. */
int ddp_gen_rmat(int32_t scale, int64_t E, double a, double b, double c, uint32_t wmin,
                 uint32_t wmax, uint64_t seed, ddp_hgraph **out);
/* nb batches of n_ins inserts (u != v uniform, w uniform in [wmin,wmax]) and
 * n_del deletes of distinct live edge instances of the pre-batch graph;
 * within a batch deletes are listed first. */
int ddp_gen_updates(const ddp_hgraph *g, int32_t nb, int32_t n_ins, int32_t n_del, uint32_t wmin,
                    uint32_t wmax, uint64_t seed, ddp_updates **out);
/* Query pairs on a geometric graph: straight-line separation log-uniform in
 * [dmin, dmax] lattice units, target = nearest node to the sampled point,
 * kept only if s and t are connected and distinct. */
int ddp_gen_geo_pairs(const ddp_hgraph *g, int32_t nq, double dmin, double dmax, uint64_t seed,
                      int32_t *src, int32_t *dst);
/* Uniform pairs (s != t) such that t is reachable from s (checked on the host). */
int ddp_gen_reachable_pairs(const ddp_hgraph *g, int32_t nq, uint64_t seed, int32_t *src,
                            int32_t *dst);
/* Connected pairs for large graphs: s uniform among the nodes that reach the
 * hub (largest out-degree), t uniform among the nodes the hub reaches (s != t);
 * two traversals in total.  DDP_ERR_STATE if the hub's reach is too small. */
int ddp_gen_hub_pairs(const ddp_hgraph *g, int32_t nq, uint64_t seed, int32_t *src,
                      int32_t *dst);

#ifdef __cplusplus
}
#endif

#endif /* DDP_IO_H */
