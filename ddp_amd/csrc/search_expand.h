// Block-cooperative expansion of the nodes popped in one GA* iteration.
//
// Each lane of the block brings at most one popped node u (its queue's pop).
// Out-degrees are prefix-summed across the block and the union of all rows is
// walked R*B edges per round, so lanes read consecutive CSR entries and a long
// row is split across the whole block instead of serialising one lane (the
// reference gives each thread its own row: GA_star/GA_v2.cu:109-167).
// Within a window every lane owns R edges and issues their loads stage by
// stage (row entries, then heuristic inputs and (f, parent) words, then the
// CAS or atomicMin attempts), so R independent memory round trips overlap
// instead of running back to back.
#pragma once

#include "device_common.h"

template <int V>
struct ExpandLds {
    static_assert(V <= 65536, "owner entries are 16-bit");
    int32_t incl[V];     // inclusive prefix of degrees
    int32_t rowb[V];     // row start
    uint32_t gu[V];      // g(u) = f(u) - h(u), unsigned as GA_v2.cu:137
    int32_t node[V];
    int32_t tmp[16];
    uint16_t own[16 * V];  // edge slot of the round -> virtual lane (a hint)
};

// Owner entries a lane writes per row and window: rows up to this long (every
// row of a road or geometric graph) are located with one LDS read per edge;
// the rest of a longer row falls back to a binary search over the prefix.
constexpr int32_t kOwnRun = 32;

// What expanding u needs besides u itself: its row and g(u).
struct RowSpan {
    int32_t b = 0, deg = 0;
    uint32_t gu = 0;
};

template <int HK>
__device__ __forceinline__ RowSpan row_span(const GraphView &G, const HeurView &H, int32_t u, uint32_t fu)
{
    RowSpan r;
    if (u >= 0) {
        r.gu = fu - (uint32_t)heur_eval<HK>(H, u);
        r.b = G.off[u];
        r.deg = G.off[u + 1] - r.b;
    }
    return r;
}

// Per-lane statistics, 32-bit to spare registers.  Kernels add them to the
// 64-bit device totals at exit, and slot kernels also whenever a wave's
// edge count passes 2^30 (no counter grows faster than the edge count,
// except rebuilds/refills/scanned, which lane 0 counts once per refill).
struct ExpandCounters {
    uint32_t expanded = 0, edges = 0, improved = 0;
    uint32_t rebuilds = 0;  // pile engine: far piles rebuilt after an overflow (lane 0 counts)
    uint32_t refills = 0;   // pile engine: near-ring refills from the far pile (lane 0 counts)
    uint32_t scanned = 0;   // pile engine: far-pile entries read by refills
};

// Edges handled per lane and window (memory requests a lane keeps in flight
// per stage).  Four where registers allow it without lowering occupancy (the
// 1,024-lane pile kernel fits 128 VGPRs at four); two and three measured no
// faster on the C4 batch.
template <int B>
constexpr int expand_unroll()
{
    return B == 512 ? 2 : 4;
}

// Stages 2-4 of one window of R edges per lane: edge r leads to c[r] (< 0:
// none) with weight wt[r] from the node par(r), whose g is gu(r).  Loads are
// issued stage by stage (heuristic inputs and (f, parent) words, then one
// atomic per improving edge), and improvements go to the push policy.  Fb
// as in expand_rows.
//
// DIRECT (atomicMin relaxation, heuristics that need no load): no word is
// read first; every edge that can still beat Fb issues its atomicMin, whose
// returned word says whether it improved.  A small round is a chain of
// dependent round trips, and this takes one out of it; a large round keeps the
// read, which spares the atomics of the many edges that cannot improve.
template <int R, int HK, bool DIRECT = false, class GuOf, class ParOf, class Push>
__device__ __forceinline__ void relax_window(const GraphView &G, const HeurView &H, uint64_t *gp, int32_t t,
                                             int32_t (&c)[R], const uint32_t (&wt)[R], GuOf gu_of, ParOf par_of,
                                             uint32_t Fb, ExpandCounters &C, Push &push)
{
    static_assert(!DIRECT || HK == DDP_H_ZERO || HK == DDP_H_MANHATTAN, "DIRECT: heuristics without loads only");
    // stage 2: heuristic inputs and the current (f, parent) words, all
    // loads issued before the first heuristic is computed.  hin, cur and
    // got below are only written and read by lanes whose edge takes part:
    // giving the others a value would make the compiler merge the two and
    // wait for each load (or CAS) before issuing the next.
    HeurIn hin[R];
    uint64_t cur[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if constexpr (HK == DDP_H_ZERO || HK == DDP_H_MANHATTAN)
            if (c[r] >= 0 && gu_of(r) + wt[r] + (uint32_t)heur_from<HK>(H, c[r], HeurIn{0, 0}) >= Fb)
                c[r] = -1;  // cannot beat the bound: no word read
        if (c[r] >= 0) {  // negative id = deleted edge
            hin[r] = heur_load<HK>(H, c[r]);
            if constexpr (!DIRECT) cur[r] = ld_gp(gp + c[r]);
        }
    }
    int32_t hc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) hc[r] = c[r] >= 0 ? heur_from<HK>(H, c[r], hin[r]) : 0;
    // stage 3: one CAS attempt per improving edge, all issued together
    uint64_t want[R], got[R];
    bool tried[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        tried[r] = false;
        want[r] = 0;
        if (c[r] >= 0 && (HK != DDP_H_ARRAY || hc[r] >= 0)) {  // h = -1: cannot reach t
            const uint32_t cand = gu_of(r) + wt[r] + (uint32_t)hc[r];
            want[r] = ((uint64_t)cand << 32) | (uint32_t)par_of(r);
            if constexpr (DIRECT) tried[r] = true;
            else tried[r] = (uint32_t)(cur[r] >> 32) > cand;
        }
    }
    // Every load above has returned by now (the comparisons used them);
    // saying so keeps the compiler from waiting again before each CAS.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    const bool by_min = DIRECT || G.relax_min != 0;  // uniform
    if (by_min) {
        // one atomicMin per improving edge: the smallest (f, parent) of
        // the round wins with no retry round trips
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tried[r]) got[r] = (uint64_t)atomicMin((unsigned long long *)(gp + c[r]),
                                                       (unsigned long long)want[r]);
    } else if constexpr (!DIRECT) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tried[r])
                got[r] = (uint64_t)atomicCAS((unsigned long long *)(gp + c[r]), (unsigned long long)cur[r],
                                             (unsigned long long)want[r]);
    }
    // stage 4: settle contended CASes, then hand improvements to the push policy
    bool has[R], first[R];
    uint64_t key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        has[r] = false;
        first[r] = false;
        key[r] = 0;
        if (tried[r]) {
            const uint32_t cand = (uint32_t)(want[r] >> 32);
            bool ok;
            uint64_t prev;
            if constexpr (DIRECT) {
                ok = (uint32_t)(got[r] >> 32) > cand;
                prev = got[r];
            } else {
                ok = by_min ? (uint32_t)(got[r] >> 32) > cand : got[r] == cur[r];
                prev = by_min ? got[r] : cur[r];
            }
            if (!DIRECT && !ok && !by_min) {
                uint64_t now = got[r];
                while ((uint32_t)(now >> 32) > cand) {
                    const uint64_t old = atomicCAS((unsigned long long *)(gp + c[r]), (unsigned long long)now,
                                                   (unsigned long long)want[r]);
                    if (old == now) {
                        ok = true;
                        prev = now;
                        break;
                    }
                    now = old;
                }
            }
            if (ok) {
                C.improved++;
                first[r] = (uint32_t)(prev >> 32) == DDP_INF_F;
                if (c[r] == t) {
                    push.target(cand);
                } else {
                    has[r] = true;
                    key[r] = ((uint64_t)cand << 32) | (uint32_t)c[r];
                }
            }
        }
    }
    push.template touch_all<R>(first, c);
    push.template all<R>(has, key);
}

// Push policy interface (block-collective members: every lane calls them):
//   void target(uint32_t f)                    -- t improved to f (t is never queued)
//   void touch(int32_t c, uint64_t prev)       -- c's word was prev before the update
//   void operator()(bool has, uint64_t key)    -- block-collective
//   void touch_all<R>(first[R], c[R])          -- block-collective: touch() of R edges
//                                                 whose previous f was infinite
//   void all<R>(has[R], key[R])                -- block-collective: operator() of R edges
// Each lane brings up to P nodes u[p] (or -1) and their row spans (row_span,
// or the same values loaded earlier by the caller); node p of lane l is
// "virtual lane" l*P + p of a B*P-wide expansion, so a wide round pays the
// scan and barriers once for B*P rows.
//
// Without FINAL_SYNC the function returns without its closing barrier; the
// caller must pass a barrier before the next expand_rows (or before reading
// what the pushes wrote).
//
// Fb: an upper bound the caller no longer needs to beat (the target's f when
// it only wants the target's cost).  Where the heuristic needs no load (zero,
// Manhattan) an edge whose candidate f is already >= Fb is skipped before its
// (f, parent) word is read: it cannot lead to a cheaper path to t, and the
// word's random 8-byte read costs a whole cache line.  Callers that keep
// searching for a later, larger bound (a resumed DA* search) pass UINT32_MAX.
template <int B, int HK, int P, class Push, bool FINAL_SYNC = true, bool DIRECT_OK = false>
__device__ __forceinline__ void expand_rows(const GraphView &G, const HeurView &H, uint64_t *gp, int32_t t,
                                            const int32_t (&u)[P], const RowSpan (&rs)[P], ExpandLds<B * P> &L,
                                            ExpandCounters &C, Push &push, uint32_t Fb = 0xFFFFFFFFu)
{
    constexpr int R = expand_unroll<B>();
    constexpr int V = B * P;
    const int lane = threadIdx.x;
    int32_t deg[P], lsum = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        deg[p] = u[p] >= 0 ? rs[p].deg : 0;
        lsum += deg[p];
        if (u[p] >= 0) C.expanded++;
    }
    int32_t total;
    // no barrier after the scan: the next write to L.tmp comes after the
    // barrier that publishes L below
    const int32_t run0 = block_incl_scan<B, false>(lsum, L.tmp, total) - lsum;
    int32_t ends[P];
    {
        int32_t run = run0;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            run += deg[p];
            ends[p] = run;
            L.incl[lane * P + p] = run;
            L.rowb[lane * P + p] = rs[p].b;
            L.gu[lane * P + p] = rs[p].gu;
            L.node[lane * P + p] = u[p];
        }
    }
    // Owner hints, when the round's edges fit in four windows of B*R slots
    // (16V): each row writes its virtual lane into the slots of its first
    // kOwnRun edges.  A hint is checked against the prefix before use, so a
    // stale one only costs the binary search.  Larger rounds (hub rows of
    // power-law graphs) skip the hints: refilling them per window cost two
    // barriers each and measured slower than searching.
    const bool use_own = total <= 4 * B * R;
    // a round of at most one edge per lane is latency-bound: relax without the
    // word read (atomicMin relaxation, heuristics without loads; relax_window).
    // Only where the caller allows it (DIRECT_OK): the second copy of the
    // relaxation costs registers that the 1,024-lane batch and serve kernels
    // do not have to spare, and their rounds are mostly large.
    const bool direct = DIRECT_OK && G.relax_min != 0 && total <= B;
    if (use_own) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int32_t st = ends[p] - deg[p];
            const int32_t to = ends[p] < st + kOwnRun ? ends[p] : st + kOwnRun;
            for (int32_t x = st; x < to; ++x) L.own[x] = (uint16_t)(lane * P + p);
        }
    }
    __syncthreads();
    int32_t prow = 0;  // row of this lane's previous edge (edges are visited in increasing order)
    for (int32_t base = 0; base < total; base += B * R) {
        int32_t c[R], jj[R];
        uint32_t wt[R];
        // stage 1: locate each edge (LDS only), then issue its column and
        // weight loads for all R edges before any of them is used
        int32_t pos[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int32_t e = base + r * B + lane;
            pos[r] = -1;
            jj[r] = 0;
            if (e < total) {
                // the owner hint, else the row of this lane's previous edge
                // (a lane's edges only move forward, so that row or a later
                // one holds e: hub rows keep it for many windows), else a
                // binary search from that row on
                int32_t lo = use_own ? L.own[e] : prow;
                int32_t ex = 0;
                if (lo < V) ex = lo ? L.incl[lo - 1] : 0;
                if (!(lo < V && ex <= e && e < L.incl[lo])) {
                    int32_t a = prow, hi = V - 1;
                    while (a < hi) {
                        const int32_t mid = (a + hi) >> 1;
                        if (L.incl[mid] > e) hi = mid;
                        else a = mid + 1;
                    }
                    lo = a;
                    ex = lo ? L.incl[lo - 1] : 0;
                }
                prow = lo;
                jj[r] = lo;
                pos[r] = L.rowb[lo] + (e - ex);
                C.edges++;
            }
        }
        int32_t dv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            c[r] = -1;
            wt[r] = 0;
            dv[r] = INT32_MAX;
            if (pos[r] >= 0) {
                c[r] = G.col[pos[r]];
                wt[r] = G.w[pos[r]];
                if (G.dver) dv[r] = G.dver[pos[r]];
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (G.dver && dv[r] <= G.ver) c[r] = -1;  // deleted at or before this version
        auto gu_of = [&](int r) { return L.gu[jj[r]]; };
        auto par_of = [&](int r) { return L.node[jj[r]]; };
        if constexpr (DIRECT_OK && (HK == DDP_H_ZERO || HK == DDP_H_MANHATTAN)) {
            if (direct) relax_window<R, HK, true>(G, H, gp, t, c, wt, gu_of, par_of, Fb, C, push);
            else relax_window<R, HK>(G, H, gp, t, c, wt, gu_of, par_of, Fb, C, push);
        } else {
            relax_window<R, HK>(G, H, gp, t, c, wt, gu_of, par_of, Fb, C, push);
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (!G.lhead) break;
        // Versioned inserts: each lane walks its own node's list, oldest
        // first, stopping at the first entry newer than this version (links
        // carry the inserting version, so newer entries are never read).
        const uint32_t gu = rs[p].gu;
        uint64_t link = u[p] >= 0 ? __hip_atomic_load(G.lhead + u[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : ~0ull;
        for (;;) {
            bool active = false, has = false;
            uint64_t key = 0;
            int32_t x = -1;
            while ((uint32_t)link <= (uint32_t)G.ver) {
                x = (int32_t)(link >> 32);
                link = __hip_atomic_load(G.lnext + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (G.ldel[x] > G.ver) {
                    active = true;
                    break;
                }
            }
            if (active) {
                C.edges++;
                const int32_t cn = G.lv[x];
                const int32_t hcn = heur_eval<HK>(H, cn);
                if (HK != DDP_H_ARRAY || hcn >= 0) {
                    const uint32_t cand = gu + G.lw[x] + (uint32_t)hcn;
                    uint64_t prev;
                    if (relax_gp(gp, cn, cand, (uint32_t)u[p], &prev)) {
                        C.improved++;
                        push.touch(cn, prev);
                        if (cn == t) {
                            push.target(cand);
                        } else {
                            has = true;
                            key = ((uint64_t)cand << 32) | (uint32_t)cn;
                        }
                    }
                }
            }
            push(has, key);
            if (!__syncthreads_or(active)) break;
        }
    }
    if constexpr (FINAL_SYNC) __syncthreads();
}

template <int B, int HK, bool DIRECT_OK = false, class Push>
__device__ __forceinline__ void expand_block(const GraphView &G, const HeurView &H, uint64_t *gp, int32_t t,
                                             int32_t u, uint32_t fu, ExpandLds<B> &L, ExpandCounters &C,
                                             Push &push)
{
    const int32_t uu[1] = {u};
    const RowSpan rs[1] = {row_span<HK>(G, H, u, fu)};
    expand_rows<B, HK, 1, Push, true, DIRECT_OK>(G, H, gp, t, uu, rs, L, C, push);
}
