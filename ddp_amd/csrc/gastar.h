// Host-side interface of the single-query GA* grid kernel, shared by the
// entry points that run it (ddp_astar, DA* resume).
#pragma once

#include "ddp_internal.h"

// Control block layout (uint32 words).  Three rotating slots carry, per GA*
// iteration, the push count, the smallest queue top, the smallest pushed f,
// the target's best f and error bits.  Slot 2 plays "iteration -1": whatever
// is pushed there before a launch is drained into the queues at iteration 0.
enum : int {
    GA_SLOT_WORDS = 8,
    GA_S_PUSHCNT = 0, GA_S_TOPMIN = 1, GA_S_PUSHMIN = 2, GA_S_TMIN = 3, GA_S_ERR = 4,
    GA_RES_IT = 32, GA_RES_RR = 33, GA_RES_F = 34, GA_RES_ERR = 35, GA_RES_PLEN = 36,
    GA_BAR = 64,
    GA_CTRL_WORDS = 96,
};

// Error bits of GA_RES_ERR.
enum : uint32_t { GA_ERR_HEAP = 1, GA_ERR_PBUF = 2, GA_ERR_ITERS = 4, GA_ERR_BARRIER = 8 };

struct GaGridPlan {
    int32_t K, B, nblocks, lcap, cap, pcap, maxit;
    bool wg = false;     // one workgroup (ddp_astar only): lane heaps, or piles when `piles`
    bool piles = false;
    bool bidi = false;   // piles from both ends: two (f, parent) arrays, touched logs and pile sets
    uint32_t ncap = 0, fcap = 0, delta0 = 1;  // pile mode: near ring, far buffers, first window
};

// Device buffers of one GA* search state.
struct GaGridState {
    uint64_t *gp = nullptr;      // N packed (f << 32 | parent); bidirectional: 2N, forward then backward
    uint64_t *home = nullptr;    // cap * K heap entries (bidirectional: twice)
    int32_t *hsize = nullptr;    // K
    int32_t *touched = nullptr;  // N, workgroup mode only (bidirectional: 2N)
    uint64_t *pbuf = nullptr;    // 2 * pcap push entries
    uint32_t *ctrl = nullptr;    // GA_CTRL_WORDS
    unsigned long long *stats = nullptr;  // 8 counters after ctrl
    bool owned = false;          // allocated by ddp_ga_alloc_owned (freed by ddp_ga_free)
};

// min_pcap: lower bound on the push buffer (DA* repairs push before a launch).
// allow_wg: honour opts->mode and plan the workgroup mode when it applies.
int ddp_ga_plan(ddp_ctx *ctx, const ddp_graph *g, const ddp_astar_opts_t *o, GaGridPlan *p,
                int64_t min_pcap = 0, bool allow_wg = false);
// State in the context's reusable scratch (valid until the next ddp_astar on ctx).
int ddp_ga_alloc(ddp_ctx *ctx, const ddp_graph *g, const GaGridPlan &P, GaGridState *S);
// State owned by the caller (DA*), freed by ddp_ga_free.
int ddp_ga_alloc_owned(ddp_ctx *ctx, int32_t N, const GaGridPlan &P, GaGridState *S);
void ddp_ga_free(GaGridState *S);
// gp = INF everywhere, queues empty, counters zero.
int ddp_ga_reset(ddp_ctx *ctx, int32_t N, const GaGridPlan &P, const GaGridState &S);
// Fresh search from s: gp[s] = (h(s), -1) and s as the only pending push.
int ddp_ga_seed(ddp_ctx *ctx, const GaGridPlan &P, const GaGridState &S, int32_t s, uint32_t hs);
// Clears the three slots before repair kernels push into slot 2 / buffer 1.
int ddp_ga_prepare_resume(ddp_ctx *ctx, const GaGridState &S);
int ddp_ga_run(ddp_ctx *ctx, const GraphView &gv, const HeurView &hv, const GaGridPlan &P,
               const GaGridState &S, int32_t t, uint32_t F0, int32_t rr0, uint32_t *F_out,
               int32_t *rr_out, int32_t *iters_out, float *kernel_ms);
// Grid mode, asynchronous: resets the state, seeds s and enqueues the grid
// kernel on ctx->stream without waiting.  Stream-ordered work after it reads
// the outcome from S.ctrl (GA_RES_F: f of t, DDP_INF_F if unreachable;
// GA_RES_ERR: GA_ERR_* bits).
int ddp_ga_launch(ddp_ctx *ctx, const GraphView &gv, const HeurView &hv, const GaGridPlan &P,
                  const GaGridState &S, int32_t N, int32_t s, uint32_t hs, int32_t t);
// Workgroup mode: one search from s to completion (no seed needed).  The
// bidirectional plan also needs the transposed rows (gr) and walks the path on
// the device into dpath[0..dpath_cap) (length to *plen: 0 if unreachable,
// -len if dpath_cap is too small) when dpath is given.
int ddp_ga_run_wg(ddp_ctx *ctx, const GraphView &gv, const HeurView &hv, const GaGridPlan &P,
                  const GaGridState &S, int32_t N, int32_t s, int32_t t, uint32_t *F_out, int32_t *iters_out,
                  float *kernel_ms, const GraphView *gr = nullptr, int32_t *dpath = nullptr, int32_t dpath_cap = 0,
                  int32_t *plen = nullptr);
// Reads the four search counters (expanded, edges, improved, stale).
int ddp_ga_stats(ddp_ctx *ctx, const GaGridState &S, unsigned long long out[4]);
