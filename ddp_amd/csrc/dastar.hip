// Dynamic A* (DA*): repair one query's search state after a batch of edge
// inserts and deletes, then resume the GA* grid kernel from its retained
// queues.
//
// Invariant kept between batches (the one A* itself maintains): every finite
// g is the cost of a real path, and every node that is not in a queue has
// relaxed all of its live out-edges with its current g.  A batch restores it:
//   1. deletes (matched on the device by apply_deletes, first live instance
//      of u->v as in DA_star/GA_merge.cu:1019-1037) become tombstones in HBM;
//   2. if a deleted u->v was v's parent edge, v's whole parent subtree is
//      invalidated (found by walking children: y with parent(y) == x), which
//      replaces the ancestor walks and serial cycle fixing of
//      DA_star/DA_dyn_insert_del_nocycle.cu:376-483, 1023-1134;
//   3. each invalidated node re-pulls its best value from live in-edges of
//      nodes outside the invalidated set (a parent chain that avoids every
//      invalidated node, so no cycle can form) and is queued;
//   4. inserted edges u->v are relaxed from u (propogateAdd, NC:486-553);
//   5. GA* resumes: queued nodes are expanded only while they can still beat
//      the target, instead of the unbounded Bellman-Ford waves of
//      insert_propagate / delete_propagate (NC:556-901).
// The graph lives on the device: base rows with tombstones, plus a log of
// inserted edges threaded onto per-source lists (insert order) and per-target
// lists, so the row of u reads exactly as GA_merge.cu's merged row (surviving
// base edges, then inserts in arrival order).  A batch is applied by kernels:
// each delete CASes the first live matching instance (base row first, then the
// source's list), inserts are appended to both lists.  The host does O(batch)
// work per batch (checking and grouping the lines), never O(N).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/ddp_dastar.h"
#include "compact.h"
#include "gastar.h"
#include "slot_search.h"

namespace {

using clk = std::chrono::steady_clock;

// Stable sort of (node, entry) pairs by node id in [0, n): LSD radix with
// 11-bit digits.  An update batch holds thousands of pairs; this is several
// times faster than a comparison sort on them and keeps equal nodes in input
// order, which the insert lists rely on.
void sort_by_node(std::vector<std::pair<int32_t, int32_t>> &v, int32_t n)
{
    if (v.size() < 2) return;
    std::vector<std::pair<int32_t, int32_t>> tmp(v.size());
    int bits = 1;
    while (bits < 31 && (1ll << bits) < (int64_t)n) ++bits;
    for (int sh = 0; sh < bits; sh += 11) {
        uint32_t cnt[2049] = {0};
        for (const auto &p : v) cnt[(((uint32_t)p.first >> sh) & 2047u) + 1]++;
        for (int i = 0; i < 2048; ++i) cnt[i + 1] += cnt[i];
        for (const auto &p : v) tmp[cnt[((uint32_t)p.first >> sh) & 2047u]++] = p;
        v.swap(tmp);
    }
}
float ms_since(clk::time_point t0) { return std::chrono::duration<float, std::milli>(clk::now() - t0).count(); }

enum : int { RC_NEXT = 0, RC_MARKED = 1, RC_ROOTS = 2, RC_REPULLED = 3, RC_PUSHES = 4, RC_WORDS = 16 };

enum : int { RC_DEL_BASE = 5, RC_DEL_LOG = 6, RC_WALKED = 7 };

// The log of inserted edges: entry x is u -> v with weight w, live while
// del[x] == INT32_MAX.  head/next links pack (entry << 32 | 0); the end of a
// list is ~0.  Lists are only appended between searches.
struct InsLog {
    int32_t *u, *v;
    uint32_t *w;
    int32_t *del;
    uint64_t *fhead, *fnext;  // by source
    uint64_t *rhead, *rnext;  // by target
    int32_t *ftail, *rtail;   // last entry of each list (appends)
};

__device__ __forceinline__ bool link_end(uint64_t link) { return (uint32_t)link != 0u; }

// Deletes, one thread per line: the first live instance of u -> v, base row
// first, then u's inserted edges oldest first (DA_star/GA_merge.cu:1019-1037
// on the merged row).  Concurrent deletes of the same pair CAS distinct
// instances, so k deletes remove the first k live ones, as a sequential pass
// would.  A deleted edge that was v's parent edge makes v a root of the
// subtree to invalidate.
__global__ void apply_deletes(const int32_t *pairs, int32_t n, GraphView F, int32_t *col, int32_t *rcol,
                              const int32_t *roff, const int32_t *reid, InsLog L, const uint64_t *gp, int32_t *mark,
                              int32_t epoch, int32_t *front, int32_t *marked, uint32_t *rctrl)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t u = pairs[2 * i], v = pairs[2 * i + 1];
    bool found = false;
    for (int32_t j = F.off[u]; j < F.off[u + 1] && !found; ++j) {
        if (__hip_atomic_load(col + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != v) continue;
        if (atomicCAS(col + j, v, -1) != v) continue;
        found = true;
        for (int32_t r = roff[v]; r < roff[v + 1]; ++r)
            if (reid[r] == j) {
                rcol[r] = -1;
                break;
            }
        atomicAdd(rctrl + RC_DEL_BASE, 1u);
    }
    for (uint64_t link = L.fhead[u]; !found && !link_end(link);) {
        const int32_t x = (int32_t)(link >> 32);
        link = L.fnext[x];
        if (L.v[x] == v && atomicCAS(L.del + x, INT32_MAX, 0) == INT32_MAX) {
            found = true;
            atomicAdd(rctrl + RC_DEL_LOG, 1u);
        }
    }
    if (!found) return;
    const uint64_t word = ld_gp(gp + v);
    if ((uint32_t)(word >> 32) == DDP_INF_F || (uint32_t)word != (uint32_t)u) return;
    if (atomicMax(mark + v, epoch) == epoch) return;
    front[atomicAdd(rctrl + RC_NEXT, 1u)] = v;
    marked[atomicAdd(rctrl + RC_MARKED, 1u)] = v;
    atomicAdd(rctrl + RC_ROOTS, 1u);
}

// New log entries: one thread per group of equal keys (entries arrive sorted
// by key, stably, so a group keeps insert order) appends the group to the
// key's list.  grp holds (key, entry) pairs.
__global__ void log_append(const int32_t *grp, int32_t n, uint64_t *head, uint64_t *next, int32_t *tail)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t key = grp[2 * i];
    if (i > 0 && grp[2 * (i - 1)] == key) return;
    for (int32_t k = i; k < n && grp[2 * k] == key; ++k) {
        const int32_t x = grp[2 * k + 1];
        next[x] = ~0ull;
        const uint64_t link = (uint64_t)(uint32_t)x << 32;
        const int32_t last = tail[key];
        if (last < 0) head[key] = link;
        else next[last] = link;
        tail[key] = x;
    }
}

// New log entries at + k from the uploaded (u, v, w) triples, live (del =
// INT32_MAX) until a later batch deletes them.
__global__ void log_fill(const int32_t *ins, int32_t n, InsLog L, int64_t at)
{
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    L.u[at + k] = ins[3 * k];
    L.v[at + k] = ins[3 * k + 1];
    L.w[at + k] = (uint32_t)ins[3 * k + 2];
    L.del[at + k] = INT32_MAX;
}

__global__ void fill_i32(int32_t *p, int64_t n, int32_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void fill_u64(uint64_t *p, int64_t n, uint64_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// y is a child of a frontier node: marks it once per epoch and lists it;
// nxt[k - base] receives the child that drew k from rctrl[RC_NEXT].
__device__ __forceinline__ void claim(int32_t y, int32_t *mark, int32_t epoch, int32_t *nxt, uint32_t base,
                                      int32_t *marked, uint32_t *rctrl)
{
    if (atomicMax(mark + y, epoch) == epoch) return;
    nxt[atomicAdd(rctrl + RC_NEXT, 1u) - base] = y;
    marked[atomicAdd(rctrl + RC_MARKED, 1u)] = y;
}

__device__ __forceinline__ void claim_child(int32_t y, int32_t x, const uint64_t *gp, int32_t *mark, int32_t epoch,
                                            int32_t *nxt, uint32_t base, int32_t *marked, uint32_t *rctrl)
{
    if ((uint32_t)ld_gp(gp + y) == (uint32_t)x) claim(y, mark, epoch, nxt, base, marked, rctrl);
}

// One node of a level of the parent-subtree walk: the children y of frontier
// node x are its live out-neighbours whose parent is x.
__device__ __forceinline__ void claim_children(const GraphView &G, const InsLog &L, int32_t x, const uint64_t *gp,
                                               int32_t *mark, int32_t epoch, int32_t *nxt, uint32_t base,
                                               int32_t *marked, uint32_t *rctrl)
{
    // the row in groups of 8: all column loads, then all parent-word loads,
    // then the claims (one round trip per stage instead of two per edge)
    constexpr int kG = 8;
    const int32_t rb = G.off[x], re = G.off[x + 1];
    for (int32_t j0 = rb; j0 < re; j0 += kG) {
        int32_t y[kG];
        uint32_t par[kG];
#pragma unroll
        for (int k = 0; k < kG; ++k) y[k] = j0 + k < re ? G.col[j0 + k] : -1;
#pragma unroll
        for (int k = 0; k < kG; ++k) par[k] = y[k] >= 0 ? (uint32_t)ld_gp(gp + y[k]) : 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < kG; ++k)
            if (y[k] >= 0 && par[k] == (uint32_t)x) claim(y[k], mark, epoch, nxt, base, marked, rctrl);
    }
    for (uint64_t link = L.fhead[x]; !link_end(link);) {
        const int32_t e = (int32_t)(link >> 32);
        link = L.fnext[e];
        if (L.del[e] == INT32_MAX) claim_child(L.v[e], x, gp, mark, epoch, nxt, base, marked, rctrl);
    }
}

__global__ void subtree_level(GraphView G, InsLog L, const uint64_t *gp, int32_t *mark, int32_t epoch,
                              const int32_t *cur, int32_t ncur, int32_t *nxt, int32_t *marked, uint32_t *rctrl)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncur) claim_children(G, L, cur[i], gp, mark, epoch, nxt, 0u, marked, rctrl);
}

// The narrow levels of the subtree walk in one workgroup, so that they cost
// no launch and host round trip each.  Starts from the rctrl[RC_NEXT] nodes in
// cur and runs level after level while a level holds at most `wide` nodes.
// rctrl[RC_NEXT] keeps counting across levels (a level's children take the
// draws from its base on), so no reset has to be ordered against the next
// level's draws, and the barriers order everything else: one workgroup, no
// device-wide fence.  On exit rctrl[RC_NEXT] is the size of the level left to
// walk (0 when done), and rctrl[RC_WALKED] the number of levels walked: when it
// is odd, that level sits in the buffer passed as nxt.  Every node is marked at
// most once per epoch, so the loop ends after at most N levels and the count
// stays below 2N.
constexpr int kWalkBlock = 1024;

__global__ void __launch_bounds__(kWalkBlock)
    subtree_walk(GraphView G, InsLog L, const uint64_t *gp, int32_t *mark, int32_t epoch, int32_t *cur,
                 int32_t *nxt, int32_t *marked, uint32_t *rctrl, int32_t wide)
{
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = __hip_atomic_load(rctrl + RC_NEXT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    uint32_t base = s_n;  // draws so far; the current level holds ncur nodes
    int32_t ncur = (int32_t)base;
    uint32_t walked = 0;
    while (ncur > 0 && ncur <= wide) {
        for (int32_t i = threadIdx.x; i < ncur; i += kWalkBlock)
            claim_children(G, L, cur[i], gp, mark, epoch, nxt, base, marked, rctrl);
        // every draw has returned; waiting on the memory counter before the
        // barrier (a bare s_barrier) lands this wave's nxt stores, so the next
        // level reads them
        block_sync_global();
        if (threadIdx.x == 0) s_n = __hip_atomic_load(rctrl + RC_NEXT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t total = s_n;
        ncur = (int32_t)(total - base);
        base = total;
        int32_t *tmp = cur;
        cur = nxt;
        nxt = tmp;
        ++walked;  // thread 0 writes s_n again only after the next level's first barrier
    }
    if (threadIdx.x == 0) {
        rctrl[RC_NEXT] = (uint32_t)ncur;
        rctrl[RC_WALKED] = walked;
    }
}

struct RepairPush {
    uint64_t *buf;     // push buffer of "iteration -1"
    uint32_t *slot;    // control slot 2
    int32_t pcap;
    uint32_t *rctrl;
    int32_t t;

    __device__ void push(bool has, int32_t v, uint32_t f)
    {
        if (has && v == t) {
            atomicMin(slot + GA_S_TMIN, f);  // the target is never queued
            has = false;
        }
        const uint64_t mask = __ballot(has);
        if (!mask) return;
        const int lane = threadIdx.x & 63;
        const int leader = __builtin_ctzll(mask);
        uint32_t base = 0;
        if (lane == leader) {
            base = atomicAdd(slot + GA_S_PUSHCNT, (uint32_t)__popcll(mask));
            atomicAdd(rctrl + RC_PUSHES, (uint32_t)__popcll(mask));
        }
        base = __shfl(base, leader, 64);
        if (has) {
            const uint32_t idx = base + lane_rank(mask);
            if (idx < (uint32_t)pcap) buf[idx] = ((uint64_t)f << 32) | (uint32_t)v;
            else atomicOr(slot + GA_S_ERR, 2u);
            atomicMin(slot + GA_S_PUSHMIN, f);
        }
    }
};

// Invalidated node x: best live in-edge from a node outside the invalidated
// set with a finite value; x's word is overwritten (INF if none).
__global__ void repull(GraphView R, InsLog L, const int32_t *h, uint64_t *gp, const int32_t *mark, int32_t epoch,
                       const int32_t *marked, int32_t n, RepairPush P)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t x = -1;
    uint32_t best = 0xFFFFFFFFu, bp = 0xFFFFFFFFu;
    if (i < n) {
        x = marked[i];
        const uint32_t hx = (uint32_t)h[x];
        auto consider = [&](int32_t p, uint32_t w) {
            if (p < 0 || p == x || mark[p] == epoch) return;
            const uint32_t fp = (uint32_t)(ld_gp(gp + p) >> 32);
            if (fp == DDP_INF_F) return;
            const uint32_t cand = (fp - (uint32_t)h[p]) + w + hx;  // unsigned, as GA_v2.cu:137
            if (cand < DDP_INF_F && (cand < best || (cand == best && (uint32_t)p < bp))) {
                best = cand;
                bp = (uint32_t)p;
            }
        };
        for (int32_t j = R.off[x]; j < R.off[x + 1]; ++j) consider(R.col[j], R.w[j]);
        for (uint64_t link = L.rhead[x]; !link_end(link);) {
            const int32_t e = (int32_t)(link >> 32);
            link = L.rnext[e];
            if (L.del[e] == INT32_MAX) consider(L.u[e], L.w[e]);
        }
        gp[x] = best == 0xFFFFFFFFu ? DDP_GP_INF : (((uint64_t)best << 32) | bp);
        if (best != 0xFFFFFFFFu) atomicAdd(P.rctrl + RC_REPULLED, 1u);
    }
    P.push(x >= 0 && best != 0xFFFFFFFFu, x, best);
}

// Newly inserted u->v (w): relax from u.
__global__ void relax_inserts(const int32_t *ins, int32_t n, const int32_t *h, uint64_t *gp, RepairPush P)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool has = false;
    int32_t v = 0;
    uint32_t cand = 0;
    if (i < n) {
        const int32_t u = ins[3 * i];
        v = ins[3 * i + 1];
        const uint32_t w = (uint32_t)ins[3 * i + 2];
        const uint32_t fu = (uint32_t)(ld_gp(gp + u) >> 32);
        if (fu != DDP_INF_F && u != v) {
            cand = (fu - (uint32_t)h[u]) + w + (uint32_t)h[v];
            uint64_t prev;
            has = cand < DDP_INF_F && relax_gp(gp, v, cand, (uint32_t)u, &prev);
        }
    }
    P.push(has, v, cand);
}

// The search behind DA* in pile mode: one workgroup runs the pile engine of
// the batched slots (slot_search.h) resumed from the retained open set -- the
// dormant pile of the previous search plus this batch's repair pushes -- with
// F(t) read from the target's word.  Its rounds synchronise through LDS
// barriers only, where the grid mode pays a device-wide barrier per
// iteration; a repair's few-hop propagation is bound by that per-round
// latency.  Entries that cannot beat F(t) now are kept (dormant) for later
// batches, which may raise F(t) again.
constexpr int kPileBlock = 1024;

// open_split counters
enum { SP_LIVE = 0, SP_KEEP = 1, SP_OVF = 2, SP_WORDS = 4 };

// Ahead of the one-workgroup search, the whole GPU splits the retained open
// set and this batch's repair pushes around F(t): current entries below it
// (live) seed the search, current entries at or above it stay dormant
// (keep), and stale ones -- the node's word has moved on, so an entry with
// its current f exists elsewhere -- are dropped.  The search then scans only
// the live entries, where it used to scan the whole retained set (1.3M+
// entries on the 1M-node graph of 8 out-edges) on its one CU every batch.
__global__ void open_split(const uint64_t *open, uint32_t n_open, const uint64_t *push, const uint32_t *push_n,
                           uint32_t pcap, const uint64_t *gp, int32_t t, uint64_t *live, uint64_t *keep, uint32_t cap,
                           uint32_t *cnt)
{
    const uint32_t np0 = __hip_atomic_load(push_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t np = np0 < pcap ? np0 : pcap;
    if (np0 > pcap && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(cnt + SP_OVF, 1u);
    const uint64_t n = (uint64_t)n_open + np;
    const uint32_t F0 = (uint32_t)(ld_gp(gp + t) >> 32);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // uniform trips per block: the appends are wave-wide
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const uint64_t i = i0 + threadIdx.x;
        bool sel[2] = {false, false};
        uint64_t key = 0;
        if (i < n) {
            key = i < n_open ? open[i] : push[i - n_open];
            const int32_t v = (int32_t)(uint32_t)key;
            const uint32_t f = (uint32_t)(key >> 32);
            if (v != t && (uint32_t)(ld_gp(gp + v) >> 32) == f) {
                sel[0] = f < F0;
                sel[1] = !sel[0];
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool s1[1] = {sel[k]};
            uint32_t pos[1];
            wave_append<1>(s1, cnt + k, pos);
            if (sel[k]) {
                if (pos[0] < cap) (k ? keep : live)[pos[0]] = key;
                else atomicOr(cnt + SP_OVF, 1u);
            }
        }
    }
}

struct DaPileArgs {
    GraphView g;
    HeurView h;
    uint64_t *gp;
    uint64_t *live, *alt, *dorm;  // live entries in, far-pile partner, dormant pile (kept entries first)
    uint64_t *near;               // ncap entries
    uint32_t cap, ncap, lring, delta0;
    const uint32_t *split;        // open_split counters
    int32_t N, t, maxit;
    uint32_t *res;                // iterations, F(t), error flags, dormant entries
    unsigned long long *stats;
};

__global__ void __launch_bounds__(kPileBlock) dastar_pile_kernel(DaPileArgs a)
{
    extern __shared__ uint64_t s_ring[];
    __shared__ SlotLds<kPileBlock, kPileBlock> S;
    const int lane = threadIdx.x;
    const uint32_t n_live = __hip_atomic_load(a.split + SP_LIVE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t n_keep = __hip_atomic_load(a.split + SP_KEEP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool fits = __hip_atomic_load(a.split + SP_OVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    const uint32_t F0 = (uint32_t)(ld_gp(a.gp + a.t) >> 32);
    const PilePtrs P{a.near, a.live, a.alt, a.ncap, a.cap, s_ring, a.lring};
    const PileResume Z{a.dorm, a.cap, fits ? n_live : 0u, F0, fits ? n_keep : 0u};
    ExpandCounters C;
    uint32_t stale = 0;
    slot_search_piles<kPileBlock, DDP_H_ARRAY, 1, true>(a.g, a.h, a.gp, nullptr, P, a.N, -1, a.t, a.maxit, a.delta0, S,
                                                        C, stale, &Z);
    if (lane == 0) {
        a.res[0] = (uint32_t)S.iters;
        a.res[1] = S.F;
        a.res[2] = S.err | (fits ? 0u : (uint32_t)QERR_HEAP);
        a.res[3] = S.dorm_n;
    }
    slot_flush_stats<kPileBlock>(a.stats, C, stale);
}

__global__ void map_h(const int32_t *in, int32_t *out, int32_t N)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        out[i] = in ? (in[i] < 0 ? 0 : in[i]) : 0;
}

template <class T>
int dev_upload(ddp_ctx *ctx, T **dst, size_t *cap, const T *src, size_t n)
{
    if (*cap < n || !*dst) {
        if (*dst) {
            DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            DDP_HIP(ctx, hipFree(*dst));
            *dst = nullptr;
        }
        const size_t c = std::max<size_t>(std::max<size_t>(n, 16), *cap * 2);
        DDP_HIP(ctx, hipMalloc((void **)dst, c * sizeof(T)));
        *cap = c;
    }
    if (n) DDP_HIP(ctx, hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return DDP_OK;
}

}  // namespace

struct ddp_dastar {
    ddp_ctx *ctx = nullptr;
    int device = 0;
    int32_t N = 0, E = 0, s = 0, t = 0;
    int64_t live_base = 0, live_log = 0;  // live edges (ddp_dastar_num_edges)
    // device graph: base rows with tombstones (d->g, transposed rows too) ...
    ddp_graph *g = nullptr;
    int32_t *d_reid = nullptr;  // forward edge id of each transposed entry
    // ... and the log of inserted edges
    InsLog L{};
    int32_t log_cap = 0, log_used = 0;
    int64_t compact_at = 0;     // log entries that trigger a compaction after a batch (0 = never)
    int32_t compactions = 0;
    int32_t *d_h = nullptr;
    int32_t *d_mark = nullptr, *d_front = nullptr, *d_marked = nullptr;
    uint32_t *d_rctrl = nullptr;
    int32_t *d_work = nullptr;  // per-batch lines: delete pairs, insert triples, list groups
    size_t c_work = 0;
    int32_t epoch = 0;
    // search
    GaGridPlan P{};
    GaGridState S{};
    HeurView hv{};
    int32_t rr = 0;
    uint32_t F = DDP_INF_F;
    // pile mode (DDP_ASTAR_PILES): three rotating buffers (open set in,
    // far-pile partner, dormant pile out), the HBM near ring and result words
    bool piles = false;
    uint64_t *pbuf[3] = {nullptr, nullptr, nullptr};
    uint64_t *pnear = nullptr;
    uint32_t *pres = nullptr;
    uint32_t *d_split = nullptr;  // open_split counters
    uint32_t pcap_open = 0, pncap = 0;
    int32_t open_at = 0;   // which buffer holds the open set
    uint32_t open_n = 0;   // its entries
};

static void log_free(InsLog &L)
{
    void *ptrs[] = {L.u, L.v, L.w, L.del, L.fnext, L.rnext};
    for (void *p : ptrs)
        if (p) hipFree(p);
    L.u = L.v = L.del = nullptr;
    L.w = nullptr;
    L.fnext = L.rnext = nullptr;
}

static void dastar_free(ddp_dastar *d)
{
    ddp_ga_free(&d->S);
    log_free(d->L);
    void *ptrs[] = {d->L.fhead, d->L.rhead, d->L.ftail, d->L.rtail, d->d_reid, d->d_h,
                    d->d_mark, d->d_front, d->d_marked, d->d_rctrl, d->d_work,
                    d->pbuf[0], d->pbuf[1], d->pbuf[2], d->pnear, d->pres, d->d_split};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (d->g) ddp_graph_destroy(d->g);
}

extern "C" void ddp_dastar_close(ddp_dastar *d)
{
    if (!d) return;
    hipSetDevice(d->device);
    dastar_free(d);
    delete d;
}

// Inserted edges are visible to the search through the list fields of the
// view: every link carries version 0 and a live entry's delete word is
// INT32_MAX, so "version <= ver < delete word" with ver = INT32_MAX - 1 reads
// exactly the live entries.
static GraphView fwd_view(const ddp_dastar *d)
{
    GraphView v = ddp_view(d->g);
    v.ver = INT32_MAX - 1;
    v.lhead = d->L.fhead;
    v.lnext = d->L.fnext;
    v.lv = d->L.v;
    v.lw = d->L.w;
    v.ldel = d->L.del;
    return v;
}

static GraphView rev_view(const ddp_dastar *d)
{
    GraphView v{};
    v.off = d->g->roff;
    v.col = d->g->rcol;
    v.w = d->g->rw;
    v.N = d->N;
    return v;
}

// Makes room for `need` log entries, keeping the ones in use (device copy).
static int log_reserve(ddp_dastar *d, int32_t need)
{
    if (need <= d->log_cap) return DDP_OK;
    ddp_ctx *ctx = d->ctx;
    const int32_t cap = (int32_t)std::min<int64_t>(std::max<int64_t>(need, 2LL * d->log_cap + 65536), INT32_MAX);
    DDP_CHECK(need <= cap, ctx, DDP_ERR_NOMEM, "DA*: inserted-edge log would exceed %d entries", INT32_MAX);
    InsLog n = d->L;
    hipError_t e = hipSuccess;
    auto al = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 16));
    };
    n.u = n.v = n.del = nullptr;
    n.w = nullptr;
    n.fnext = n.rnext = nullptr;
    al((void **)&n.u, (size_t)cap * 4);
    al((void **)&n.v, (size_t)cap * 4);
    al((void **)&n.w, (size_t)cap * 4);
    al((void **)&n.del, (size_t)cap * 4);
    al((void **)&n.fnext, (size_t)cap * 8);
    al((void **)&n.rnext, (size_t)cap * 8);
    const size_t used = (size_t)d->log_used;
    if (e == hipSuccess && used) {
        auto cp = [&](void *dst, const void *src, size_t bytes) {
            if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
        };
        cp(n.u, d->L.u, used * 4);
        cp(n.v, d->L.v, used * 4);
        cp(n.w, d->L.w, used * 4);
        cp(n.del, d->L.del, used * 4);
        cp(n.fnext, d->L.fnext, used * 8);
        cp(n.rnext, d->L.rnext, used * 8);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        log_free(n);
        return ddp_hip_err(ctx, e, "DA* log growth", __FILE__, __LINE__);
    }
    log_free(d->L);
    d->L.u = n.u;
    d->L.v = n.v;
    d->L.w = n.w;
    d->L.del = n.del;
    d->L.fnext = n.fnext;
    d->L.rnext = n.rnext;
    d->log_cap = cap;
    return DDP_OK;
}

// Folds the log into the base rows (merge_graph / mergeDiff,
// DA_star/DA_dyn_insert_del_nocycle.cu:940-980, 2236-2284): the base rows
// become the merged rows of the current version (tombstones dropped, live
// inserts appended in insert order), the transposed rows are rebuilt from
// them, and the log is emptied.  Nodes keep their ids, so the search state
// (g values, parents, queues) carries over unchanged.
static int dastar_compact(ddp_dastar *d)
{
    ddp_ctx *ctx = d->ctx;
    const MergeSrc m{d->g->off, d->g->col, d->g->w, d->N, nullptr, INT32_MAX - 1, d->L.fhead, d->L.fnext,
                     d->L.v, d->L.w, d->L.del};
    MergedCsr f, r;
    int rc = ddp_merge_rows(ctx, m, &f);
    if (rc) return rc;
    if ((rc = ddp_transpose_rows(ctx, f, true, &r))) {
        ddp_merged_free(&f);
        return rc;
    }
    void *old[] = {d->g->off, d->g->col, d->g->w, d->g->roff, d->g->rcol, d->g->rw, d->d_reid};
    for (void *p : old)
        if (p) hipFree(p);
    d->g->off = f.off;
    d->g->col = f.col;
    d->g->w = f.w;
    d->g->E = f.E;
    d->g->roff = r.off;
    d->g->rcol = r.col;
    d->g->rw = r.w;
    d->g->RE = r.E;
    d->d_reid = r.eid;
    d->E = f.E;
    const int32_t N = d->N;
    hipLaunchKernelGGL(fill_u64, dim3(1024), dim3(256), 0, ctx->stream, d->L.fhead, (int64_t)N, ~0ull);
    hipLaunchKernelGGL(fill_u64, dim3(1024), dim3(256), 0, ctx->stream, d->L.rhead, (int64_t)N, ~0ull);
    hipLaunchKernelGGL(fill_i32, dim3(1024), dim3(256), 0, ctx->stream, d->L.ftail, (int64_t)N, -1);
    hipLaunchKernelGGL(fill_i32, dim3(1024), dim3(256), 0, ctx->stream, d->L.rtail, (int64_t)N, -1);
    DDP_HIP(ctx, hipGetLastError());
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    d->log_used = 0;
    d->live_base = f.E;
    d->live_log = 0;
    d->compactions += 1;
    return DDP_OK;
}

// One pile-mode search over the open set d->pbuf[open_at] (open_n entries)
// plus the repair pushes counted in control slot 2; the dormant pile it leaves
// becomes the next open set.
static int pile_run(ddp_dastar *d, int32_t *iters, float *kms)
{
    ddp_ctx *ctx = d->ctx;
    DaPileArgs a{};
    a.g = fwd_view(d);
    a.g.relax_min = d->g->minw >= 1;
    a.h = d->hv;
    a.gp = d->S.gp;
    // the open set (pbuf[open_at]) is split into live (the next buffer) and
    // kept entries (the one after, which the search's dormant pile extends);
    // the open set's own buffer is free then and serves as the far-pile partner
    uint64_t *open = d->pbuf[d->open_at];
    a.live = d->pbuf[(d->open_at + 1) % 3];
    a.alt = open;
    a.dorm = d->pbuf[(d->open_at + 2) % 3];
    a.near = d->pnear;
    a.cap = d->pcap_open;
    a.ncap = d->pncap;
    a.lring = 8u * kPileBlock;
    a.delta0 = std::max<uint32_t>(1, d->g->meanw);
    a.split = d->d_split;
    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, ctx->stream, (int32_t *)d->d_split, (int64_t)SP_WORDS, 0);
    hipLaunchKernelGGL(open_split, dim3(1024), dim3(256), 0, ctx->stream, (const uint64_t *)open, d->open_n,
                       (const uint64_t *)(d->S.pbuf + d->P.pcap),
                       (const uint32_t *)(d->S.ctrl + 2 * GA_SLOT_WORDS + GA_S_PUSHCNT), (uint32_t)d->P.pcap,
                       (const uint64_t *)d->S.gp, d->t, a.live, a.dorm, d->pcap_open, d->d_split);
    a.N = d->N;
    a.t = d->t;
    a.maxit = d->P.maxit;
    a.res = d->pres;
    a.stats = d->S.stats;
    const size_t lds = (size_t)a.lring * 8;
    DDP_HIP(ctx, hipFuncSetAttribute((const void *)dastar_pile_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    DDP_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(dastar_pile_kernel, dim3(1), dim3(kPileBlock), lds, ctx->stream, a);
    DDP_HIP(ctx, hipGetLastError());
    DDP_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t res[4];
    DDP_HIP(ctx, hipMemcpyAsync(res, d->pres, 16, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    DDP_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *kms = ms;
    *iters = (int32_t)res[0];
    DDP_CHECK(!(res[2] & QERR_ITERS), ctx, DDP_ERR_TIMEOUT, "DA*: iteration limit %d reached", d->P.maxit);
    DDP_CHECK(!(res[2] & QERR_HEAP), ctx, DDP_ERR_OVERFLOW,
              "DA*: the pile-mode open set outgrew its %u entries", d->pcap_open);
    d->F = res[1];
    d->open_at = (d->open_at + 2) % 3;
    d->open_n = res[3];
    return DDP_OK;
}

static int fresh_search(ddp_dastar *d, int32_t *iters, float *kms)
{
    ddp_ctx *ctx = d->ctx;
    int rc;
    if ((rc = ddp_ga_reset(ctx, d->N, d->P, d->S))) return rc;
    int32_t hs = 0;
    DDP_HIP(ctx, hipMemcpyAsync(&hs, d->d_h + d->s, 4, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    d->rr = 0;
    if (d->s == d->t) {
        d->F = (uint32_t)hs;
        uint64_t word = ((uint64_t)(uint32_t)hs << 32) | 0xFFFFFFFFull;
        DDP_HIP(ctx, hipMemcpyAsync(d->S.gp + d->s, &word, 8, hipMemcpyHostToDevice, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return DDP_OK;
    }
    if (d->piles) {
        // the open set is s alone; no repair pushes
        const uint64_t word = ((uint64_t)(uint32_t)hs << 32) | 0xFFFFFFFFull;
        const uint64_t key = ((uint64_t)(uint32_t)hs << 32) | (uint32_t)d->s;
        DDP_HIP(ctx, hipMemcpyAsync(d->S.gp + d->s, &word, 8, hipMemcpyHostToDevice, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(d->pbuf[d->open_at], &key, 8, hipMemcpyHostToDevice, ctx->stream));
        if ((rc = ddp_ga_prepare_resume(ctx, d->S))) return rc;
        d->open_n = 1;
        return pile_run(d, iters, kms);
    }
    if ((rc = ddp_ga_seed(ctx, d->P, d->S, d->s, (uint32_t)hs))) return rc;
    return ddp_ga_run(ctx, fwd_view(d), d->hv, d->P, d->S, d->t, DDP_INF_F, 0, &d->F, &d->rr, iters, kms);
}

extern "C" int ddp_dastar_open(ddp_ctx *ctx, int32_t N, int32_t E, const int32_t *off, const int32_t *col,
                               const uint32_t *w, const int32_t *h, int32_t s, int32_t t,
                               const ddp_astar_opts_t *opts, ddp_dastar **out, int32_t *cost)
{
    if (!ctx || !out || !cost) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_dastar_open: null argument");
    *out = nullptr;
    DDP_CHECK(N > 0 && s >= 0 && s < N && t >= 0 && t < N, ctx, DDP_ERR_ARG, "ddp_dastar_open: s=%d t=%d N=%d", s,
              t, N);
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    ddp_dastar *d = new ddp_dastar();
    d->ctx = ctx;
    d->device = ctx->device;
    d->N = N;
    d->s = s;
    d->t = t;
    int rc = ddp_graph_upload(ctx, N, E, off, col, w, 0, &d->g);
    if (rc) {
        delete d;
        return rc;
    }
    d->E = E;
    // stable transpose (DA_star/genOpCSR.cpp order) with forward edge ids
    std::vector<int32_t> roff(N + 1, 0), rcol, reid;
    std::vector<uint32_t> rw;
    for (int32_t j = 0; j < E; ++j)
        if (col[j] >= 0) roff[col[j] + 1]++;
    for (int32_t i = 0; i < N; ++i) roff[i + 1] += roff[i];
    const int32_t RE = roff[N];
    d->live_base = RE;
    d->compact_at = std::max<int64_t>(65536, RE / 4);
    rcol.assign(RE, 0);
    rw.assign(RE, 0);
    reid.assign(RE, 0);
    {
        std::vector<int32_t> f(roff.begin(), roff.end() - 1);
        for (int32_t u = 0; u < N; ++u)
            for (int32_t j = off[u]; j < (u + 1 < N ? off[u + 1] : E); ++j) {
                if (col[j] < 0) continue;
                const int32_t p = f[col[j]]++;
                rcol[p] = u;
                rw[p] = w[j];
                reid[p] = j;
            }
    }
    hipError_t e = hipSuccess;
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(dst, std::max<size_t>(bytes, 16));
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    };
    up((void **)&d->g->roff, roff.data(), (size_t)(N + 1) * 4);
    up((void **)&d->g->rcol, rcol.data(), (size_t)RE * 4);
    up((void **)&d->g->rw, rw.data(), (size_t)RE * 4);
    up((void **)&d->d_reid, reid.data(), (size_t)RE * 4);
    d->g->RE = RE;
    auto al = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 16));
    };
    al((void **)&d->L.fhead, (size_t)N * 8);
    al((void **)&d->L.rhead, (size_t)N * 8);
    al((void **)&d->L.ftail, (size_t)N * 4);
    al((void **)&d->L.rtail, (size_t)N * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_u64, dim3(1024), dim3(256), 0, ctx->stream, d->L.fhead, (int64_t)N, ~0ull);
        hipLaunchKernelGGL(fill_u64, dim3(1024), dim3(256), 0, ctx->stream, d->L.rhead, (int64_t)N, ~0ull);
        hipLaunchKernelGGL(fill_i32, dim3(1024), dim3(256), 0, ctx->stream, d->L.ftail, (int64_t)N, -1);
        hipLaunchKernelGGL(fill_i32, dim3(1024), dim3(256), 0, ctx->stream, d->L.rtail, (int64_t)N, -1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMalloc((void **)&d->d_h, (size_t)N * 4);
    int32_t *h_tmp = nullptr;
    if (e == hipSuccess && h) {
        e = hipMalloc((void **)&h_tmp, (size_t)N * 4);
        if (e == hipSuccess) e = hipMemcpyAsync(h_tmp, h, (size_t)N * 4, hipMemcpyHostToDevice, ctx->stream);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(map_h, dim3(1024), dim3(256), 0, ctx->stream, h_tmp, d->d_h, N);
        e = hipStreamSynchronize(ctx->stream);
    }
    if (h_tmp) hipFree(h_tmp);
    if (e == hipSuccess) e = hipMalloc((void **)&d->d_mark, (size_t)N * 4);
    if (e == hipSuccess) e = hipMemsetAsync(d->d_mark, 0, (size_t)N * 4, ctx->stream);
    if (e == hipSuccess) e = hipMalloc((void **)&d->d_front, (size_t)N * 8 + 64);
    if (e == hipSuccess) e = hipMalloc((void **)&d->d_marked, (size_t)N * 4 + 64);
    if (e == hipSuccess) e = hipMalloc((void **)&d->d_rctrl, RC_WORDS * 4);
    if (e != hipSuccess) {
        dastar_free(d);
        delete d;
        return ddp_hip_err(ctx, e, "DA* setup", __FILE__, __LINE__);
    }
    d->hv = HeurView{};
    d->hv.kind = DDP_H_ARRAY;
    d->hv.scale = 1;
    d->hv.h = d->d_h;
    if ((rc = ddp_ga_plan(ctx, d->g, opts, &d->P, 2 * (int64_t)N + 65536)) ||
        (rc = ddp_ga_alloc_owned(ctx, N, d->P, &d->S)) || (rc = log_reserve(d, 1))) {
        dastar_free(d);
        delete d;
        return rc;
    }
    d->piles = opts && opts->mode == DDP_ASTAR_PILES;
    if (d->piles) {
        // room for three open sets of an exhaustive search each; the near ring
        // a power of two of about a quarter of that
        d->pcap_open = (uint32_t)std::min<int64_t>(std::max<int64_t>(1 << 20, 3LL * N), 1LL << 31);
        uint32_t nc = 1;
        while (nc * 2 <= std::max<uint32_t>(d->pcap_open / 4, 1u << 16)) nc *= 2;
        d->pncap = nc;
        hipError_t e2 = hipSuccess;
        for (int k = 0; k < 3 && e2 == hipSuccess; ++k) e2 = hipMalloc((void **)&d->pbuf[k], (size_t)d->pcap_open * 8);
        if (e2 == hipSuccess) e2 = hipMalloc((void **)&d->pnear, (size_t)nc * 8);
        if (e2 == hipSuccess) e2 = hipMalloc((void **)&d->pres, 64);
        if (e2 == hipSuccess) e2 = hipMalloc((void **)&d->d_split, SP_WORDS * 4);
        if (e2 != hipSuccess) {
            dastar_free(d);
            delete d;
            return ddp_hip_err(ctx, e2, "DA* pile state", __FILE__, __LINE__);
        }
    }
    int32_t iters = 0;
    float kms = 0;
    if ((rc = fresh_search(d, &iters, &kms))) {
        dastar_free(d);
        delete d;
        return rc;
    }
    *cost = (int32_t)d->F;
    *out = d;
    return DDP_OK;
}

extern "C" int ddp_dastar_apply(ddp_dastar *d, const ddp_update_t *lines, int32_t n, int32_t mode, int32_t *cost,
                                ddp_dastar_stats_t *st)
{
    if (!d || !cost || (n && !lines)) return DDP_ERR_ARG;
    ddp_ctx *ctx = d->ctx;
    // checked before the batch touches the device: a batch applied without
    // its repair would leave g values over edges that no longer exist
    DDP_CHECK(mode == DDP_DASTAR_REPAIR || mode == DDP_DASTAR_RECOMPUTE, ctx, DDP_ERR_ARG, "unknown DA* mode %d",
              mode);
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    const auto t0 = clk::now();
    const int32_t N = d->N;
    ddp_dastar_stats_t S{};
    for (int32_t i = 0; i < n; ++i) {
        const ddp_update_t &l = lines[i];
        DDP_CHECK(l.flag == 0 || l.flag == 1, ctx, DDP_ERR_ARG, "update line %d: flag %d", i, l.flag);
        DDP_CHECK(l.u >= 0 && l.u < N && l.v >= 0 && l.v < N, ctx, DDP_ERR_ARG, "update line %d: edge %d->%d outside [0,%d)",
                  i, l.u, l.v, N);
    }
    // ---- 1. host, O(batch): split the lines; group new inserts by source and by target
    std::vector<int32_t> pairs, ins;
    std::vector<std::pair<int32_t, int32_t>> fg, rg;  // (key, log entry), insert order kept
    for (int32_t i = 0; i < n; ++i) {
        const ddp_update_t &l = lines[i];
        if (l.flag == 0) {
            ++S.deletes;
            pairs.push_back(l.u);
            pairs.push_back(l.v);
        } else {
            const int32_t x = d->log_used + S.inserts;
            ++S.inserts;
            ins.push_back(l.u);
            ins.push_back(l.v);
            ins.push_back((int32_t)l.w);
            fg.emplace_back(l.u, x);
            rg.emplace_back(l.v, x);
        }
    }
    sort_by_node(fg, N);
    sort_by_node(rg, N);
    int rc = log_reserve(d, d->log_used + S.inserts);
    if (rc) return rc;
    // device work list: [pairs | ins | fg | rg]
    std::vector<int32_t> work;
    work.reserve(pairs.size() + ins.size() + 4 * fg.size());
    work.insert(work.end(), pairs.begin(), pairs.end());
    work.insert(work.end(), ins.begin(), ins.end());
    for (const auto &x : fg) work.insert(work.end(), {x.first, x.second});
    for (const auto &x : rg) work.insert(work.end(), {x.first, x.second});
    if ((rc = dev_upload(ctx, &d->d_work, &d->c_work, work.data(), work.size()))) return rc;
    const int32_t *w_pairs = d->d_work, *w_ins = w_pairs + pairs.size(), *w_fg = w_ins + ins.size(),
                  *w_rg = w_fg + 2 * fg.size();
    const int T = 256;
    auto grid = [&](size_t k) { return dim3((unsigned)std::max<size_t>(1, (k + T - 1) / T)); };
    // ---- 2. device: deletes (they never see this batch's inserts), then the appends
    d->epoch += 1;
    uint32_t rctrl[RC_WORDS] = {0};
    DDP_HIP(ctx, hipMemsetAsync(d->d_rctrl, 0, RC_WORDS * 4, ctx->stream));
    const int32_t npairs = (int32_t)(pairs.size() / 2), nins = S.inserts;
    if (npairs)
        hipLaunchKernelGGL(apply_deletes, grid(npairs), dim3(T), 0, ctx->stream, w_pairs, npairs, fwd_view(d),
                           d->g->col, d->g->rcol, (const int32_t *)d->g->roff, (const int32_t *)d->d_reid, d->L,
                           (const uint64_t *)d->S.gp, d->d_mark, d->epoch, d->d_front, d->d_marked, d->d_rctrl);
    if (nins) {
        hipLaunchKernelGGL(log_fill, grid(nins), dim3(T), 0, ctx->stream, w_ins, nins, d->L, (int64_t)d->log_used);
        hipLaunchKernelGGL(log_append, grid(nins), dim3(T), 0, ctx->stream, w_fg, nins, d->L.fhead, d->L.fnext,
                           d->L.ftail);
        hipLaunchKernelGGL(log_append, grid(nins), dim3(T), 0, ctx->stream, w_rg, nins, d->L.rhead, d->L.rnext,
                           d->L.rtail);
    }
    DDP_HIP(ctx, hipGetLastError());
    DDP_HIP(ctx, hipMemcpyAsync(rctrl, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    d->log_used += nins;
    S.deletes_effective = (int32_t)(rctrl[RC_DEL_BASE] + rctrl[RC_DEL_LOG]);
    d->live_base -= rctrl[RC_DEL_BASE];
    d->live_log += nins - (int64_t)rctrl[RC_DEL_LOG];
    S.host_ms = ms_since(t0);
    int32_t iters = 0;
    float kms = 0;
    if (mode == DDP_DASTAR_RECOMPUTE) {
        if ((rc = fresh_search(d, &iters, &kms))) return rc;
    } else {
        const auto t1 = clk::now();
        // ---- 3. invalidate parent subtrees of deleted tree edges (roots seeded by apply_deletes)
        if ((rc = ddp_ga_prepare_resume(ctx, d->S))) return rc;
        const GraphView F = fwd_view(d), R = rev_view(d);
        int32_t *cur = d->d_front, *nxt = d->d_front + N;
        int32_t ncur = (int32_t)rctrl[RC_NEXT];
        S.roots = (int32_t)rctrl[RC_ROOTS];
        // Narrow levels (up to kWalkWide nodes) run inside one subtree_walk
        // launch; a wider level takes a launch of its own over the whole grid.
        constexpr int32_t kWalkWide = 8 * kWalkBlock;
        while (ncur > 0) {
            if (ncur <= kWalkWide) {
                hipLaunchKernelGGL(subtree_walk, dim3(1), dim3(kWalkBlock), 0, ctx->stream, F, d->L, d->S.gp,
                                   d->d_mark, d->epoch, cur, nxt, d->d_marked, d->d_rctrl, kWalkWide);
            } else {
                DDP_HIP(ctx, hipMemsetAsync(d->d_rctrl + RC_NEXT, 0, 4, ctx->stream));
                hipLaunchKernelGGL(subtree_level, grid(ncur), dim3(T), 0, ctx->stream, F, d->L, d->S.gp, d->d_mark,
                                   d->epoch, cur, ncur, nxt, d->d_marked, d->d_rctrl);
            }
            DDP_HIP(ctx, hipMemcpyAsync(rctrl, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
            DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            const bool walked = ncur <= kWalkWide;
            ncur = (int32_t)rctrl[RC_NEXT];
            if (!walked || (rctrl[RC_WALKED] & 1u)) std::swap(cur, nxt);
        }
        S.invalidated = (int32_t)rctrl[RC_MARKED];
        // ---- 4. re-pull invalidated nodes, 5. relax inserted edges
        RepairPush P{d->S.pbuf + d->P.pcap, d->S.ctrl + 2 * GA_SLOT_WORDS, d->P.pcap, d->d_rctrl, d->t};
        if (S.invalidated)
            hipLaunchKernelGGL(repull, grid(S.invalidated), dim3(T), 0, ctx->stream, R, d->L, d->d_h, d->S.gp,
                               d->d_mark, d->epoch, d->d_marked, S.invalidated, P);
        if (nins)
            hipLaunchKernelGGL(relax_inserts, grid(nins), dim3(T), 0, ctx->stream, w_ins, nins, d->d_h, d->S.gp, P);
        DDP_HIP(ctx, hipGetLastError());
        uint64_t tword = 0;
        DDP_HIP(ctx, hipMemcpyAsync(&tword, d->S.gp + d->t, 8, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(rctrl, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipMemsetAsync(d->S.stats, 0, 64, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        S.repulled = (int32_t)rctrl[RC_REPULLED];
        S.pushes = (int32_t)rctrl[RC_PUSHES];
        S.repair_ms = ms_since(t1);
        // ---- 6. resume GA* from the retained queues plus the repair pushes
        if (d->s == d->t) {
            d->F = (uint32_t)(tword >> 32);
        } else if (d->piles) {
            if ((rc = pile_run(d, &iters, &kms))) return rc;
        } else if ((rc = ddp_ga_run(ctx, F, d->hv, d->P, d->S, d->t, (uint32_t)(tword >> 32), d->rr, &d->F, &d->rr,
                                    &iters, &kms))) {
            return rc;
        }
    }
    S.iterations = iters;
    S.search_ms = kms;
    unsigned long long cnt[4];
    if ((rc = ddp_ga_stats(ctx, d->S, cnt))) return rc;
    S.expansions = (int64_t)cnt[0];
    S.edges = (int64_t)cnt[1];
    S.improvements = (int64_t)cnt[2];
    // list walks grow with the log: fold it into the rows once it is long
    if (d->compact_at > 0 && d->log_used >= d->compact_at) {
        const auto tc = clk::now();
        if ((rc = dastar_compact(d))) return rc;
        S.compacted = 1;
        S.compact_ms = ms_since(tc);
    }
    S.total_ms = ms_since(t0);
    *cost = (int32_t)d->F;
    if (st) *st = S;
    return DDP_OK;
}

extern "C" int ddp_dastar_compact(ddp_dastar *d)
{
    if (!d) return DDP_ERR_ARG;
    DDP_HIP(d->ctx, hipSetDevice(d->ctx->device));
    return dastar_compact(d);
}

extern "C" int ddp_dastar_set_compaction(ddp_dastar *d, int64_t log_entries)
{
    if (!d) return DDP_ERR_ARG;
    DDP_CHECK(log_entries >= 0, d->ctx, DDP_ERR_ARG, "compaction threshold must be >= 0");
    d->compact_at = log_entries;
    return DDP_OK;
}

extern "C" int ddp_dastar_path(ddp_dastar *d, int32_t *path, int32_t cap, int32_t *len)
{
    if (!d || !path || !len) return DDP_ERR_ARG;
    DDP_HIP(d->ctx, hipSetDevice(d->ctx->device));
    if (d->F == DDP_INF_F) {
        *len = 0;
        return DDP_OK;
    }
    return ddp_extract_path(d->ctx, d->S.gp, d->N, d->s, d->t, path, cap, len);
}

extern "C" int ddp_dastar_num_edges(const ddp_dastar *d, int32_t *E)
{
    if (!d || !E) return DDP_ERR_ARG;
    *E = (int32_t)(d->live_base + d->live_log);
    return DDP_OK;
}

// The current graph version as a CSR in the reference layout: per node its
// surviving base edges in row order, then its live inserted edges in insert
// order (GA_merge.cu's merged row).  Reads the device state back.
extern "C" int ddp_dastar_export(const ddp_dastar *d, int32_t *off, int32_t *col, uint32_t *w)
{
    if (!d || !off || !col || !w) return DDP_ERR_ARG;
    ddp_ctx *ctx = d->ctx;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    const int32_t N = d->N, E = d->E, L = d->log_used;
    std::vector<int32_t> boff(N + 1), bcol(E), lu(L), lv(L), ldel(L);
    std::vector<uint32_t> bw(E), lw(L);
    DDP_HIP(ctx, hipMemcpyAsync(boff.data(), d->g->off, (size_t)(N + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (E) {
        DDP_HIP(ctx, hipMemcpyAsync(bcol.data(), d->g->col, (size_t)E * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(bw.data(), d->g->w, (size_t)E * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (L) {
        DDP_HIP(ctx, hipMemcpyAsync(lu.data(), d->L.u, (size_t)L * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(lv.data(), d->L.v, (size_t)L * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(lw.data(), d->L.w, (size_t)L * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(ldel.data(), d->L.del, (size_t)L * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // live inserts grouped by source, entry (= insert) order kept
    std::vector<int32_t> loff(N + 1, 0), lidx;
    for (int32_t x = 0; x < L; ++x)
        if (ldel[x] == INT32_MAX) loff[lu[x] + 1]++;
    for (int32_t i = 0; i < N; ++i) loff[i + 1] += loff[i];
    lidx.resize(loff[N]);
    {
        std::vector<int32_t> f(loff.begin(), loff.end() - 1);
        for (int32_t x = 0; x < L; ++x)
            if (ldel[x] == INT32_MAX) lidx[f[lu[x]]++] = x;
    }
    int32_t k = 0;
    for (int32_t u = 0; u < N; ++u) {
        off[u] = k;
        for (int32_t j = boff[u]; j < boff[u + 1]; ++j)
            if (bcol[j] >= 0) {
                col[k] = bcol[j];
                w[k] = bw[j];
                ++k;
            }
        for (int32_t p = loff[u]; p < loff[u + 1]; ++p) {
            col[k] = lv[lidx[p]];
            w[k] = lw[lidx[p]];
            ++k;
        }
    }
    return DDP_OK;
}
