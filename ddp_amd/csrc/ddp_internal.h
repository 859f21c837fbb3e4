// Internal host-side definitions shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdarg>
#include <string>
#include <vector>

#include "../../include/ddp_capi.h"

struct ddp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    std::string err;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Reusable device scratch, keyed by small integer slot ids (grown on demand).
    std::vector<void *> ws_ptr;
    std::vector<size_t> ws_size;
};

struct ddp_graph {
    ddp_ctx *ctx = nullptr;
    int device = 0;  // destroy uses this, never ctx (the context may be gone)
    int32_t N = 0, E = 0;
    uint32_t minw = 0, meanw = 0;  // smallest and mean weight of live edges
    int32_t maxdeg = 0;
    // forward CSR, N+1 offsets (sentinel appended at upload)
    int32_t *off = nullptr, *col = nullptr;
    uint32_t *w = nullptr;
    // transposed CSR (optional); reid[r] = forward edge index of reverse entry r
    int32_t *roff = nullptr, *rcol = nullptr, *reid = nullptr;
    uint32_t *rw = nullptr;
    int32_t RE = 0;
    // heuristics
    int32_t *h = nullptr;
    int32_t h_target = -1;  // device id the h array was computed for (ddp_bfs_hops); -1 = not known
    int32_t *xy = nullptr;
    int32_t grid_w = 0;
    // host copy of the coordinates for work estimates (batched scheduling)
    std::vector<int32_t> host_xy;
    // node relabelling of DDP_GRAPH_REORDER: perm[caller id] = device id,
    // iperm[device id] = caller id; both empty when ids are kept
    std::vector<int32_t> perm, iperm;
    int32_t *d_iperm = nullptr;  // device copy of iperm (relabelled graphs only)
};

// Caller node id -> device node id and back (identity without relabelling).
// Callers range-check ids before mapping them in; out-of-range ids (such as
// the -1 of "no parent") map out unchanged.
inline int32_t ddp_to_dev(const ddp_graph *g, int32_t v) { return g->perm.empty() ? v : g->perm[v]; }
inline int32_t ddp_to_user(const ddp_graph *g, int32_t v)
{
    return (g->iperm.empty() || v < 0 || v >= g->N) ? v : g->iperm[v];
}

// Error helpers: record a message on the context and return the code.
int ddp_set_err(ddp_ctx *ctx, int code, const char *fmt, ...);
int ddp_hip_err(ddp_ctx *ctx, hipError_t e, const char *what, const char *file, int line);

#define DDP_HIP(ctx, expr)                                                   \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) return ddp_hip_err((ctx), e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define DDP_CHECK(cond, ctx, code, ...)                                      \
    do {                                                                     \
        if (!(cond)) return ddp_set_err((ctx), (code), __VA_ARGS__);         \
    } while (0)

// Device scratch owned by the context: returns a buffer of at least `bytes`
// for `slot`, reallocating (contents lost) when it is too small.
int ddp_ws(ddp_ctx *ctx, int slot, size_t bytes, void **out);

// Workspace slot ids (one user per slot at a time).
enum {
    WS_ASTAR_GP = 0,
    WS_ASTAR_HEAP,
    WS_ASTAR_HSIZE,
    WS_ASTAR_PBUF,
    WS_ASTAR_CTRL,
    WS_ASTAR_TOUCH,
    WS_BFS_FRONT,
    WS_BFS_CTRL,
    WS_PATH,
    WS_COUNT_
};

// Device-side graph view, passed to kernels by value.
struct GraphView {
    const int32_t *__restrict__ off;   // N+1
    const int32_t *__restrict__ col;
    const uint32_t *__restrict__ w;
    int32_t N;
    // Versioned graph (serving): a base edge p is visible iff dver[p] > ver;
    // inserted edges hang off lhead[u] as a list ordered oldest to newest whose
    // links pack (entry << 32 | version that inserted it); an entry x is
    // visible iff its link version <= ver and ldel[x] > ver.
    const int32_t *__restrict__ dver;
    int32_t ver;
    // 1: relax with one 64-bit atomicMin on the packed (f, parent) word
    // instead of a strict-decrease CAS loop.  Only for graphs whose weights
    // are all >= 1 and that do not change during the search: a parent then
    // always has a strictly smaller g, so parent chains stay acyclic even
    // though an equal f may swap the parent (to the smaller id).
    int32_t relax_min;
    const uint64_t *lhead;
    const uint64_t *lnext;
    const int32_t *lv;
    const uint32_t *lw;
    const int32_t *ldel;
};

struct HeurView {
    int32_t kind;
    int32_t scale;
    const int32_t *__restrict__ h;   // DDP_H_ARRAY
    const int32_t *__restrict__ xy;  // DDP_H_EUCLID
    int32_t width;                   // DDP_H_MANHATTAN
    int32_t tx, ty;                  // target coordinates (closed forms)
    int32_t sx, sy;                  // source coordinates (bidirectional search)
};

int ddp_make_heur(const ddp_graph *g, ddp_heur_t h, int32_t t, HeurView *out);
GraphView ddp_view(const ddp_graph *g);
// h(s) on the host side (reads one value from the device for DDP_H_ARRAY).
int ddp_heur_at(const ddp_graph *g, const HeurView &hv, int32_t v, int32_t *out);

// Walk a parent array on the device from t to s; writes the s..t path into
// path_out (host).  Returns DDP_OK and sets *len (0 if no path, -len if cap
// too small).
int ddp_extract_path(ddp_ctx *ctx, const uint64_t *gp_dev, int32_t N, int32_t s,
                     int32_t t, int32_t *path_out, int32_t cap, int32_t *len);
