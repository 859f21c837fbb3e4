// Dynamic A* (DA*): repair one query's search state after a batch of edge
// inserts and deletes, then resume the search from its retained open set --
// the GA* grid kernel's queues, or the one-workgroup pile engine's dormant
// pile (DDP_ASTAR_PILES).
//
// Invariant kept between batches (the one A* itself maintains): every finite
// g is the cost of a real path, and every node that is not in a queue has
// relaxed all of its live out-edges with its current g.  A batch restores it:
//   1. deletes (matched on the device by apply_deletes, first live instance
//      of u->v as in DA_star/GA_merge.cu:1019-1037) become tombstones in HBM;
//   2. if a deleted u->v was v's parent edge, v's whole parent subtree is
//      invalidated (found by walking children: y with parent(y) == x, level
//      by level in one launch over 16 workgroups), which replaces the
//      ancestor walks and serial cycle fixing of
//      DA_star/DA_dyn_insert_del_nocycle.cu:376-483, 1023-1134;
//   3. each invalidated node re-pulls its best value from live in-edges of
//      nodes outside the invalidated set (a parent chain that avoids every
//      invalidated node, so no cycle can form) and is queued;
//   4. inserted edges u->v are relaxed from u (propogateAdd, NC:486-553);
//   5. the search resumes: queued nodes are expanded only while they can still
//      beat the target, instead of the unbounded Bellman-Ford waves of
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
#include "host_sort.h"
#include "slot_search.h"

namespace {

using clk = std::chrono::steady_clock;

// Stable sort of (node, entry) pairs by node id in [0, n): LSD radix with
// 11-bit digits.  An update batch holds thousands of pairs; this is several
// times faster than a comparison sort on them and keeps equal nodes in input
// order, which the insert lists rely on.
float ms_since(clk::time_point t0) { return std::chrono::duration<float, std::milli>(clk::now() - t0).count(); }

enum : int { RC_NEXT = 0, RC_MARKED = 1, RC_ROOTS = 2, RC_REPULLED = 3, RC_PUSHES = 4, RC_WORDS = 16 };

enum : int { RC_DEL_BASE = 5, RC_DEL_LOG = 6, RC_WALKED = 7, RC_WALK_ERR = 8 };

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

// Appends the children claimed in one group to the next level (nxt[k - base]
// for draw k of *next_cnt) and to the marked list: wave-collective (every
// lane of the wave calls it), one counter atomic per wave and list.  Per-node
// atomics on two shared counters serialised in L2 and made the walk of a C3
// batch (12,500 invalidated nodes) take about 340 us.  Both lists hold at
// most cap (N) entries, as each node is claimed once per epoch; a store past
// that is dropped rather than written out of bounds.
template <int K>
__device__ __forceinline__ void append_claims(const bool (&sel)[K], const int32_t (&y)[K], uint32_t *next_cnt,
                                              uint32_t base, int32_t *nxt, uint32_t *marked_cnt, int32_t *marked,
                                              uint32_t cap)
{
    uint32_t pos[K];
    wave_append<K>(sel, next_cnt, pos);
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (sel[k] && pos[k] - base < cap) nxt[pos[k] - base] = y[k];
    wave_append<K>(sel, marked_cnt, pos);
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (sel[k] && pos[k] < cap) marked[pos[k]] = y[k];
}

// Deletes, one thread per line: the first live instance of u -> v, base row
// first, then u's inserted edges oldest first (DA_star/GA_merge.cu:1019-1037
// on the merged row).  Concurrent deletes of the same pair CAS distinct
// instances, so k deletes remove the first k live ones, as a sequential pass
// would.  A deleted edge that was v's parent edge makes v a root of the
// subtree to invalidate.  Every lane reaches the counting at the end, where
// each wave adds its counts with one atomic per counter.
__global__ void apply_deletes(const int32_t *pairs, int32_t n, GraphView F, int32_t *col, int32_t *rcol,
                              const int32_t *roff, const int32_t *reid, InsLog L, const uint64_t *gp, int32_t *mark,
                              int32_t epoch, int32_t *front, int32_t *marked, uint32_t *rctrl)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t u = -1, v = -1;
    bool in_base = false, in_log = false;
    if (i < n) {
        u = pairs[2 * i];
        v = pairs[2 * i + 1];
        // the row in groups of 8 loads issued together; a CAS only where the
        // column matches, in row order (the first live instance wins)
        constexpr int kG = 8;
        const int32_t rb = F.off[u], re = F.off[u + 1];
        int32_t hit = -1;
        for (int32_t j0 = rb; j0 < re && hit < 0; j0 += kG) {
            int32_t c[kG];
#pragma unroll
            for (int k = 0; k < kG; ++k)
                c[k] = j0 + k < re ? __hip_atomic_load(col + j0 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : -1;
#pragma unroll
            for (int k = 0; k < kG; ++k)
                if (hit < 0 && c[k] == v && atomicCAS(col + j0 + k, v, -1) == v) hit = j0 + k;
        }
        if (hit >= 0) {
            in_base = true;
            const int32_t qb = roff[v], qe = roff[v + 1];
            bool done = false;
            for (int32_t r0 = qb; r0 < qe && !done; r0 += kG) {
                int32_t id[kG];
#pragma unroll
                for (int k = 0; k < kG; ++k) id[k] = r0 + k < qe ? reid[r0 + k] : -1;
#pragma unroll
                for (int k = 0; k < kG; ++k)
                    if (!done && id[k] == hit) {
                        rcol[r0 + k] = -1;
                        done = true;
                    }
            }
        }
        for (uint64_t link = L.fhead[u]; !in_base && !in_log && !link_end(link);) {
            const int32_t x = (int32_t)(link >> 32);
            link = L.fnext[x];
            if (L.v[x] == v && atomicCAS(L.del + x, INT32_MAX, 0) == INT32_MAX) in_log = true;
        }
    }
    bool root[1] = {false};
    const int32_t rv[1] = {v};
    if (in_base || in_log) {
        const uint64_t word = ld_gp(gp + v);
        root[0] = (uint32_t)(word >> 32) != DDP_INF_F && (uint32_t)word == (uint32_t)u &&
                  atomicMax(mark + v, epoch) != epoch;
    }
    const uint64_t mb = __ballot(in_base), ml = __ballot(in_log), mr = __ballot(root[0]);
    if ((threadIdx.x & 63) == 0) {
        if (mb) atomicAdd(rctrl + RC_DEL_BASE, (uint32_t)__popcll(mb));
        if (ml) atomicAdd(rctrl + RC_DEL_LOG, (uint32_t)__popcll(ml));
        if (mr) atomicAdd(rctrl + RC_ROOTS, (uint32_t)__popcll(mr));
    }
    append_claims<1>(root, rv, rctrl + RC_NEXT, 0u, front, rctrl + RC_MARKED, marked, (uint32_t)F.N);
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

// Where the children claimed in a level go (wave-collective append<K>):
// GlobalSink lists them in nxt[k - base] for draw k of a device counter (the
// wide levels, one launch over the grid each); WalkSink keeps a level in LDS
// and sends what does not fit to the device buffer `out` (the narrow levels
// of subtree_walk).  Both add each claimed node to the marked list.
struct GlobalSink {
    uint32_t *next_cnt, *marked_cnt;
    uint32_t base;
    int32_t *nxt, *marked;
    uint32_t cap;  // entries of nxt and marked (N)
    template <int K>
    __device__ void append(const bool (&sel)[K], const int32_t (&y)[K])
    {
        append_claims<K>(sel, y, next_cnt, base, nxt, marked_cnt, marked, cap);
    }
};

// One node x of a level of the parent-subtree walk (x < 0: none): its
// children are its live out-neighbours y whose parent is x.  A node has one
// parent and every root is marked before the walk, so a child is reached
// from one lane only (its parent's) once per epoch: reading its mark with
// its word and storing the epoch claims it without an atomic round trip.
// Wave-collective: the loop runs while any lane of the wave has work.
template <class Sink>
__device__ __forceinline__ void claim_children(const GraphView &G, const InsLog &L, int32_t x, const uint64_t *gp,
                                               int32_t *mark, int32_t epoch, Sink &sink)
{
    // Each step takes a group of 8 row entries and one entry of the inserted
    // list together: their columns and list fields in one round trip, then
    // all parent words and marks in another, then the claims.  A node with at
    // most 8 edges and one inserted edge costs one step (the row and list
    // head come in the round trip before).
    constexpr int kG = 8;
    int32_t rb = 0, re = 0;
    uint64_t link = 1ull;  // a non-zero low word ends a list
    if (x >= 0) {
        rb = G.off[x];
        re = G.off[x + 1];
        link = L.fhead[x];
    }
    for (int32_t j0 = rb; __any(j0 < re || !link_end(link)); j0 += kG) {
        int32_t y[kG + 1], mk[kG + 1];
        uint32_t par[kG + 1];
        bool sel[kG + 1];
#pragma unroll
        for (int k = 0; k < kG; ++k) y[k] = j0 + k < re ? G.col[j0 + k] : -1;
        y[kG] = -1;
        if (!link_end(link)) {
            const int32_t e = (int32_t)(link >> 32);
            link = L.fnext[e];
            const int32_t v = L.v[e];
            if (L.del[e] == INT32_MAX) y[kG] = v;
        }
#pragma unroll
        for (int k = 0; k <= kG; ++k) {
            par[k] = 0xFFFFFFFFu;
            mk[k] = epoch;
            if (y[k] >= 0) {
                par[k] = (uint32_t)ld_gp(gp + y[k]);
                mk[k] = __hip_atomic_load(mark + y[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        // Parallel edges (or a base edge and an inserted one) may name the
        // same child twice in one step: it is claimed at its first place
        // only.  Across steps the lane's own mark store precedes its next
        // loads of that address, so a later step sees the mark.
#pragma unroll
        for (int k = 0; k <= kG; ++k) {
            sel[k] = y[k] >= 0 && par[k] == (uint32_t)x && mk[k] != epoch;
#pragma unroll
            for (int q = 0; q < k; ++q) sel[k] = sel[k] && y[q] != y[k];
            if (sel[k]) mark[y[k]] = epoch;
        }
        sink.template append<kG + 1>(sel, y);
    }
}

// Device-wide barrier of the subtree walk over nblocks co-resident
// workgroups (bar[0] arrivals, bar[1] generation, bar[2] published count):
// the last workgroup to arrive reads *cnt -- every workgroup's draws of the
// level are in by then -- and publishes it with the new generation, so each
// workgroup learns the level's total without racing the next level's draws.
// Waits are bounded by limit_ticks of the 100 MHz clock (false on timeout).
__device__ __forceinline__ bool walk_barrier(uint32_t *bar, uint32_t nblocks, uint32_t &gen, const uint32_t *cnt,
                                             uint64_t limit_ticks, int32_t *s_flag, uint32_t *s_total)
{
    block_sync_global();  // every wave's draws and stores reach L2 first
    if (threadIdx.x == 0) {
        bool ok = true;
        __threadfence();
        const uint32_t arrived = atomicAdd(&bar[0], 1u);
        if (arrived == nblocks - 1) {
            __threadfence();
            const uint32_t total = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[2], total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = dev_clock_100mhz();
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (dev_clock_100mhz() - t0 > limit_ticks) {
                    ok = false;
                    break;
                }
            }
        }
        __threadfence();
        *s_total = __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_flag = ok ? 1 : 0;
    }
    gen += 1;
    __syncthreads();
    return *s_flag != 0;
}

// The parent-subtree walk in one launch over kWalkGrid co-resident
// workgroups: level after level, each ended by walk_barrier.  Node i of a
// level goes to workgroup i % kWalkGrid, so even a level of a few hundred
// nodes spreads over kWalkGrid CUs: one workgroup walking alone was bound
// by its CU's gather rate (about 35 ns per node, 42 us for a level of 1,200
// nodes).  Sixteen workgroups (4,096 lanes) still take a C3 level in one
// pass and their barrier is cheaper than 32's or 64's: the repair's device
// time per C3 batch was 0.250, 0.259 and 0.296 ms with 16, 32 and 64.  The
// roots are the rctrl[RC_NEXT] draws apply_deletes made into buf0; later
// levels alternate between buf1 and buf0 (each at most N nodes: every node is
// claimed once per epoch).  On exit rctrl[RC_WALKED] holds the number of
// levels, and RC_WALK_ERR is set if a barrier timed out.
constexpr int kWalkGrid = 16, kWalkGridBlock = 256;

__global__ void __launch_bounds__(kWalkGridBlock)
    subtree_walk(GraphView G, InsLog L, const uint64_t *gp, int32_t *mark, int32_t epoch, int32_t *buf0,
                 int32_t *buf1, int32_t *marked, uint32_t *rctrl, uint32_t *bar, uint64_t limit_ticks)
{
    __shared__ int32_t s_flag;
    __shared__ uint32_t s_total;
    const uint32_t N = (uint32_t)G.N;
    uint32_t gen = 0;
    int32_t *cur = buf0, *nxt = buf1;
    uint32_t base = __hip_atomic_load(rctrl + RC_NEXT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t ncur = base < N ? base : N;
    uint32_t levels = 0;
    while (ncur > 0) {
        GlobalSink sink{rctrl + RC_NEXT, rctrl + RC_MARKED, base, nxt, marked, N};
        const uint32_t per = (ncur + kWalkGrid - 1) / kWalkGrid;  // this workgroup's share, at most
        for (uint32_t s0 = 0; s0 < per; s0 += kWalkGridBlock) {   // uniform trips: the appends are wave-wide
            const uint32_t slot = s0 + threadIdx.x;
            const uint32_t i = slot * kWalkGrid + blockIdx.x;
            claim_children(G, L, (slot < per && i < ncur) ? cur[i] : -1, gp, mark, epoch, sink);
        }
        if (!walk_barrier(bar, gridDim.x, gen, rctrl + RC_NEXT, limit_ticks, &s_flag, &s_total)) {
            if (threadIdx.x == 0) atomicOr(rctrl + RC_WALK_ERR, 1u);
            break;
        }
        const uint32_t total = s_total;
        ncur = total - base < N ? total - base : N;
        base = total;
        int32_t *tmp = cur;
        cur = nxt;
        nxt = tmp;
        ++levels;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) rctrl[RC_WALKED] = levels;
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
        const uint32_t fmin = wave_min_u32(has ? f : 0xFFFFFFFFu);  // one atomicMin per wave
        uint32_t base = 0;
        if (lane == leader) {
            base = atomicAdd(slot + GA_S_PUSHCNT, (uint32_t)__popcll(mask));
            atomicAdd(rctrl + RC_PUSHES, (uint32_t)__popcll(mask));
            atomicMin(slot + GA_S_PUSHMIN, fmin);
        }
        base = __shfl(base, leader, 64);
        if (has) {
            const uint32_t idx = base + lane_rank(mask);
            if (idx < (uint32_t)pcap) buf[idx] = ((uint64_t)f << 32) | (uint32_t)v;
            else atomicOr(slot + GA_S_ERR, 2u);
        }
    }
};

// Invalidated node x: best live in-edge from a node outside the invalidated
// set with a finite value; x's word is overwritten (INF if none).
// Invalidated node marked[i] (i < n) takes its best in-edge from a node left
// valid, as a fresh label; it joins the repair pushes.  Wave-collective.
__device__ __forceinline__ void repull_one(const GraphView &R, const InsLog &L, const int32_t *h, uint64_t *gp,
                                           const int32_t *mark, int32_t epoch, const int32_t *marked, int32_t i,
                                           int32_t n, RepairPush &P)
{
    int32_t x = -1;
    uint32_t best = 0xFFFFFFFFu, bp = 0xFFFFFFFFu;
    if (i < n) {
        x = marked[i];
        const uint32_t hx = (uint32_t)h[x];
        auto consider = [&](int32_t p, uint32_t w, int32_t mp, uint64_t word, int32_t hp) {
            if (p < 0 || p == x || mp == epoch) return;
            const uint32_t fp = (uint32_t)(word >> 32);
            if (fp == DDP_INF_F) return;
            const uint32_t cand = (fp - (uint32_t)hp) + w + hx;  // unsigned, as GA_v2.cu:137
            if (cand < DDP_INF_F && (cand < best || (cand == best && (uint32_t)p < bp))) {
                best = cand;
                bp = (uint32_t)p;
            }
        };
        // in-edges in groups of 8: sources and weights, then their marks,
        // words and heuristics, all issued together
        constexpr int kG = 8;
        const int32_t rb = R.off[x], re = R.off[x + 1];
        for (int32_t j0 = rb; j0 < re; j0 += kG) {
            int32_t p[kG], mp[kG], hp[kG];
            uint32_t w[kG];
            uint64_t word[kG];
#pragma unroll
            for (int k = 0; k < kG; ++k) {
                p[k] = -1;
                w[k] = 0;
                if (j0 + k < re) {
                    p[k] = R.col[j0 + k];
                    w[k] = R.w[j0 + k];
                }
            }
#pragma unroll
            for (int k = 0; k < kG; ++k) {
                mp[k] = epoch;
                word[k] = DDP_GP_INF;
                hp[k] = 0;
                if (p[k] >= 0 && p[k] != x) {
                    mp[k] = mark[p[k]];
                    word[k] = ld_gp(gp + p[k]);
                    hp[k] = h[p[k]];
                }
            }
#pragma unroll
            for (int k = 0; k < kG; ++k) consider(p[k], w[k], mp[k], word[k], hp[k]);
        }
        for (uint64_t link = L.rhead[x]; !link_end(link);) {
            const int32_t e = (int32_t)(link >> 32);
            link = L.rnext[e];
            if (L.del[e] == INT32_MAX) {
                const int32_t q = L.u[e];
                if (q >= 0 && q != x) consider(q, L.w[e], mark[q], ld_gp(gp + q), h[q]);
            }
        }
        gp[x] = best == 0xFFFFFFFFu ? DDP_GP_INF : (((uint64_t)best << 32) | bp);
    }
    const uint64_t got = __ballot(x >= 0 && best != 0xFFFFFFFFu);
    if ((threadIdx.x & 63) == 0 && got) atomicAdd(P.rctrl + RC_REPULLED, (uint32_t)__popcll(got));
    P.push(x >= 0 && best != 0xFFFFFFFFu, x, best);
}

// Grid-stride over the marked list, whose length (rctrl[RC_MARKED], at most
// N) is read on the device: no host round trip between the walk and this.
__global__ void repull(GraphView R, InsLog L, const int32_t *h, uint64_t *gp, const int32_t *mark, int32_t epoch,
                       const int32_t *marked, const uint32_t *rctrl, RepairPush P)
{
    const uint32_t n0 = __hip_atomic_load(rctrl + RC_MARKED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int32_t n = (int32_t)(n0 < (uint32_t)R.N ? n0 : (uint32_t)R.N);
    for (int32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x)  // uniform trips per block
        repull_one(R, L, h, gp, mark, epoch, marked, i0 + (int32_t)threadIdx.x, n, P);
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
// Live entries are always checked; kept ones only with check_all (a random
// word read each: the caller asks for it when the retained set has grown),
// since a stale kept entry costs room, not correctness.
constexpr int kSplitBlock = 256, kSplitPer = 8;  // entries per thread: a workgroup splits 2,048

__global__ void __launch_bounds__(kSplitBlock)
    open_split(const uint64_t *open, uint32_t n_open, const uint64_t *push, const uint32_t *push_n, uint32_t pcap,
               const uint64_t *gp, int32_t t, uint64_t *live, uint64_t *keep, uint32_t cap, uint32_t *cnt,
               int32_t check_all)
{
    __shared__ uint32_t s_wc[2][kSplitBlock / 64], s_base[2];
    const uint32_t np0 = __hip_atomic_load(push_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t np = np0 < pcap ? np0 : pcap;
    if (np0 > pcap && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(cnt + SP_OVF, 1u);
    const uint64_t n = (uint64_t)n_open + np;
    const uint64_t b0 = (uint64_t)blockIdx.x * kSplitBlock * kSplitPer;
    if (b0 >= n) return;  // uniform per workgroup
    const uint32_t F0 = (uint32_t)(ld_gp(gp + t) >> 32);
    // this workgroup's entries, kSplitPer per thread (coalesced: entry b0 + k * block + thread)
    uint64_t key[kSplitPer];
    bool sel[2][kSplitPer];
#pragma unroll
    for (int k = 0; k < kSplitPer; ++k) {
        const uint64_t i = b0 + (uint64_t)k * kSplitBlock + threadIdx.x;
        sel[0][k] = sel[1][k] = false;
        key[k] = 0;
        if (i < n) key[k] = i < n_open ? open[i] : push[i - n_open];
    }
    uint64_t word[kSplitPer];
#pragma unroll
    for (int k = 0; k < kSplitPer; ++k) {
        const uint64_t i = b0 + (uint64_t)k * kSplitBlock + threadIdx.x;
        const int32_t v = (int32_t)(uint32_t)key[k];
        const bool need = i < n && v != t && ((uint32_t)(key[k] >> 32) < F0 || check_all);
        word[k] = need ? ld_gp(gp + v) : 0;
    }
    uint32_t mine[2] = {0, 0};
#pragma unroll
    for (int k = 0; k < kSplitPer; ++k) {
        const uint64_t i = b0 + (uint64_t)k * kSplitBlock + threadIdx.x;
        const int32_t v = (int32_t)(uint32_t)key[k];
        const uint32_t f = (uint32_t)(key[k] >> 32);
        if (i < n && v != t) {
            if (f < F0) sel[0][k] = (uint32_t)(word[k] >> 32) == f;
            else sel[1][k] = !check_all || (uint32_t)(word[k] >> 32) == f;
        }
        mine[0] += sel[0][k];
        mine[1] += sel[1][k];
    }
    // one reservation per workgroup and list: prefix over the threads
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t pre[2];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        const uint32_t x = (uint32_t)wave_incl_scan((int32_t)mine[l]);
        pre[l] = x - mine[l];
        if (lane == 63) s_wc[l][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t tot = 0;
        for (int w = 0; w < kSplitBlock / 64; ++w) tot += s_wc[threadIdx.x][w];
        s_base[threadIdx.x] = tot ? atomicAdd(cnt + threadIdx.x, tot) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        uint32_t pos = s_base[l] + pre[l];
        for (int w = 0; w < wid; ++w) pos += s_wc[l][w];
        uint64_t *dst = l ? keep : live;
#pragma unroll
        for (int k = 0; k < kSplitPer; ++k) {
            if (!sel[l][k]) continue;
            if (pos < cap) dst[pos] = key[k];
            else atomicOr(cnt + SP_OVF, 1u);
            ++pos;
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
    uint32_t *d_walkbar = nullptr;  // subtree_walk's device-wide barrier words
    hipEvent_t ev_r0 = nullptr, ev_r1 = nullptr;  // around the repair's device work
    int32_t *d_work = nullptr;  // per-batch lines: delete pairs, insert triples, list groups
    size_t c_work = 0;
    int32_t *h_work = nullptr;  // pinned host staging of d_work (an asynchronous upload)
    size_t c_hwork = 0;
    std::vector<int32_t> ku, kv, kx, tk, tv;  // host scratch of a batch's insert grouping
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
    uint32_t open_checked = 0;  // open_n after the last split that checked every entry
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
    if (d->h_work) hipHostFree(d->h_work);
    if (d->ev_r0) hipEventDestroy(d->ev_r0);
    if (d->ev_r1) hipEventDestroy(d->ev_r1);
    ddp_ga_free(&d->S);
    log_free(d->L);
    void *ptrs[] = {d->L.fhead, d->L.rhead, d->L.ftail, d->L.rtail, d->d_reid, d->d_h,
                    d->d_mark, d->d_front, d->d_marked, d->d_rctrl, d->d_walkbar, d->d_work,
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
// rctrl_out: also copies the repair counters back, in the same host round trip.
static int pile_run(ddp_dastar *d, int32_t *iters, float *kms, uint32_t *rctrl_out = nullptr,
                    unsigned long long *stats_out = nullptr)
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
    // kept entries are re-checked once the retained set has grown by a quarter
    const bool check_all = d->open_n > d->open_checked + d->open_checked / 4 + 65536;
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
    const uint64_t nsplit = (uint64_t)d->open_n + (uint64_t)d->P.pcap;  // the pushes' count is on the device
    const uint32_t gsplit = (uint32_t)std::max<uint64_t>(1, (nsplit + kSplitBlock * kSplitPer - 1) /
                                                                (kSplitBlock * kSplitPer));
    hipLaunchKernelGGL(open_split, dim3(gsplit), dim3(kSplitBlock), 0, ctx->stream, (const uint64_t *)open, d->open_n,
                       (const uint64_t *)(d->S.pbuf + d->P.pcap),
                       (const uint32_t *)(d->S.ctrl + 2 * GA_SLOT_WORDS + GA_S_PUSHCNT), (uint32_t)d->P.pcap,
                       (const uint64_t *)d->S.gp, d->t, a.live, a.dorm, d->pcap_open, d->d_split, (int32_t)check_all);
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
    if (rctrl_out)
        DDP_HIP(ctx, hipMemcpyAsync(rctrl_out, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (stats_out)  // the search counters in the same round trip
        DDP_HIP(ctx, hipMemcpyAsync(stats_out, d->S.stats, 64, hipMemcpyDeviceToHost, ctx->stream));
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
    if (check_all) d->open_checked = d->open_n;
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
        d->open_checked = 0;
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
    if (e == hipSuccess) e = hipMalloc((void **)&d->d_walkbar, 16);
    if (e == hipSuccess) e = hipEventCreate(&d->ev_r0);
    if (e == hipSuccess) e = hipEventCreate(&d->ev_r1);
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
    int32_t nd = 0;
    for (int32_t i = 0; i < n; ++i) {
        const ddp_update_t &l = lines[i];
        DDP_CHECK(l.flag == 0 || l.flag == 1, ctx, DDP_ERR_ARG, "update line %d: flag %d", i, l.flag);
        DDP_CHECK(l.u >= 0 && l.u < N && l.v >= 0 && l.v < N, ctx, DDP_ERR_ARG, "update line %d: edge %d->%d outside [0,%d)",
                  i, l.u, l.v, N);
        nd += l.flag == 0;
    }
    S.deletes = nd;
    S.inserts = n - nd;
    const int32_t ni = S.inserts;
    int rc = log_reserve(d, d->log_used + ni);
    if (rc) return rc;
    // ---- 1. host, O(batch), written straight into the pinned device work
    // list [pairs | ins | fg | rg]: the delete pairs, the new inserts, and the
    // inserts' log entries grouped by source and by target (insert order kept)
    // (every earlier batch synchronised before returning, so it is free)
    const size_t nwork = 2 * (size_t)nd + 7 * (size_t)ni;
    if (d->c_hwork < nwork || !d->h_work) {
        if (d->h_work) DDP_HIP(ctx, hipHostFree(d->h_work));
        d->h_work = nullptr;
        d->c_hwork = 0;
        const size_t c = std::max<size_t>(nwork, 1 << 16);
        DDP_HIP(ctx, hipHostMalloc((void **)&d->h_work, c * sizeof(int32_t)));
        d->c_hwork = c;
    }
    int32_t *o_pairs = d->h_work, *o_ins = o_pairs + 2 * (size_t)nd, *o_fg = o_ins + 3 * (size_t)ni,
            *o_rg = o_fg + 2 * (size_t)ni;
    d->ku.resize(ni);
    d->kv.resize(ni);
    d->kx.resize(ni);
    for (int32_t i = 0, a = 0, b = 0; i < n; ++i) {
        const ddp_update_t &l = lines[i];
        if (l.flag == 0) {
            o_pairs[2 * a] = l.u;
            o_pairs[2 * a + 1] = l.v;
            ++a;
        } else {
            o_ins[3 * b] = l.u;
            o_ins[3 * b + 1] = l.v;
            o_ins[3 * b + 2] = (int32_t)l.w;
            d->ku[b] = l.u;
            d->kv[b] = l.v;
            d->kx[b] = d->log_used + b;
            ++b;
        }
    }
    radix_pairs(d->ku.data(), d->kx.data(), ni, N, o_fg, d->tk, d->tv);
    radix_pairs(d->kv.data(), d->kx.data(), ni, N, o_rg, d->tk, d->tv);
    if ((rc = dev_upload(ctx, &d->d_work, &d->c_work, (const int32_t *)d->h_work, nwork))) return rc;
    const int32_t *w_pairs = d->d_work, *w_ins = w_pairs + 2 * (size_t)nd, *w_fg = w_ins + 3 * (size_t)ni,
                  *w_rg = w_fg + 2 * (size_t)ni;
    const int T = 256;
    auto grid = [&](size_t k) { return dim3((unsigned)std::max<size_t>(1, (k + T - 1) / T)); };
    // ---- 2. device: deletes (they never see this batch's inserts), then the appends
    d->epoch += 1;
    uint32_t rctrl[RC_WORDS] = {0};
    DDP_HIP(ctx, hipEventRecord(d->ev_r0, ctx->stream));
    DDP_HIP(ctx, hipMemsetAsync(d->d_rctrl, 0, RC_WORDS * 4, ctx->stream));
    const int32_t npairs = nd, nins = ni;
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
    d->log_used += nins;
    // host preparation only: the device work below follows without a host
    // round trip until the repair's counters are read
    S.host_ms = ms_since(t0);
    int32_t iters = 0;
    float kms = 0;
    unsigned long long cnt[8] = {0};  // the search counters (expanded, edges, improved, stale, rebuilds, refills, scanned)
    bool have_cnt = false;
    if (mode == DDP_DASTAR_RECOMPUTE) {
        if ((rc = fresh_search(d, &iters, &kms))) return rc;
        DDP_HIP(ctx, hipMemcpyAsync(rctrl, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        if ((rc = ddp_ga_prepare_resume(ctx, d->S))) return rc;
        const GraphView F = fwd_view(d), R = rev_view(d);
        // ---- 3. invalidate the parent subtrees of deleted tree edges (roots
        // seeded by apply_deletes): every level in one launch
        DDP_HIP(ctx, hipMemsetAsync(d->d_walkbar, 0, 16, ctx->stream));
        hipLaunchKernelGGL(subtree_walk, dim3(kWalkGrid), dim3(kWalkGridBlock), 0, ctx->stream, F, d->L, d->S.gp,
                           d->d_mark, d->epoch, d->d_front, d->d_front + N, d->d_marked, d->d_rctrl, d->d_walkbar,
                           (uint64_t)100 * 1000 * 1000 * 5);  // 5 s of the 100 MHz clock per barrier
        // ---- 4. re-pull invalidated nodes (their count is read on the device), 5. relax inserted edges
        RepairPush P{d->S.pbuf + d->P.pcap, d->S.ctrl + 2 * GA_SLOT_WORDS, d->P.pcap, d->d_rctrl, d->t};
        hipLaunchKernelGGL(repull, dim3(256), dim3(T), 0, ctx->stream, R, d->L, d->d_h, d->S.gp, d->d_mark, d->epoch,
                           (const int32_t *)d->d_marked, (const uint32_t *)d->d_rctrl, P);
        if (nins)
            hipLaunchKernelGGL(relax_inserts, grid(nins), dim3(T), 0, ctx->stream, w_ins, nins, d->d_h, d->S.gp, P);
        DDP_HIP(ctx, hipGetLastError());
        DDP_HIP(ctx, hipEventRecord(d->ev_r1, ctx->stream));
        DDP_HIP(ctx, hipMemsetAsync(d->S.stats, 0, 64, ctx->stream));
        // ---- 6. resume GA* from the retained queues plus the repair pushes
        if (d->piles && d->s != d->t) {
            // the pile search reads F(t) on the device: one host round trip
            // for the repair counters, the search and its counters together
            if ((rc = pile_run(d, &iters, &kms, rctrl, cnt))) return rc;
            have_cnt = true;
        } else {
            uint64_t tword = 0;
            DDP_HIP(ctx, hipMemcpyAsync(&tword, d->S.gp + d->t, 8, hipMemcpyDeviceToHost, ctx->stream));
            DDP_HIP(ctx, hipMemcpyAsync(rctrl, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
            DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if (d->s == d->t) {
                d->F = (uint32_t)(tword >> 32);
            } else if (d->piles) {
                if ((rc = pile_run(d, &iters, &kms))) return rc;
            } else if ((rc = ddp_ga_run(ctx, F, d->hv, d->P, d->S, d->t, (uint32_t)(tword >> 32), d->rr, &d->F,
                                        &d->rr, &iters, &kms))) {
                return rc;
            }
        }
        DDP_CHECK(rctrl[RC_WALK_ERR] == 0, ctx, DDP_ERR_TIMEOUT,
                  "DA*: a device-wide barrier of the subtree walk timed out (are %d workgroups co-resident?)",
                  kWalkGrid);
        S.roots = (int32_t)rctrl[RC_ROOTS];
        S.invalidated = (int32_t)std::min<uint32_t>(rctrl[RC_MARKED], (uint32_t)N);
        S.repulled = (int32_t)rctrl[RC_REPULLED];
        S.pushes = (int32_t)rctrl[RC_PUSHES];
        DDP_HIP(ctx, hipEventElapsedTime(&S.repair_ms, d->ev_r0, d->ev_r1));
    }
    S.deletes_effective = (int32_t)(rctrl[RC_DEL_BASE] + rctrl[RC_DEL_LOG]);
    d->live_base -= rctrl[RC_DEL_BASE];
    d->live_log += nins - (int64_t)rctrl[RC_DEL_LOG];
    S.iterations = iters;
    S.search_ms = kms;
    if (!have_cnt) {
        DDP_HIP(ctx, hipMemcpyAsync(cnt, d->S.stats, 64, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    S.expansions = (int64_t)cnt[0];
    S.edges = (int64_t)cnt[1];
    S.improvements = (int64_t)cnt[2];
    S.pile_refills = (int64_t)cnt[5];
    S.pile_scanned = (int64_t)cnt[6];
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
