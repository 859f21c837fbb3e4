// Block-cooperative expansion of the nodes popped in one GA* iteration.
//
// Each lane of the block brings at most one popped node u (its queue's pop).
// Out-degrees are prefix-summed across the block and the union of all rows is
// walked R*B edges per round, so lanes read consecutive CSR entries and a long
// row is split across the whole block instead of serialising one lane (the
// reference gives each thread its own row: GA_star/GA_v2.cu:109-167).
// Within a round every lane owns R edges and issues their loads stage by
// stage (row entries, then heuristic inputs and (f, parent) words, then the
// CAS attempts), so R independent memory round trips overlap instead of
// running back to back.
#pragma once

#include "device_common.h"

template <int B>
struct ExpandLds {
    int32_t incl[B];   // inclusive prefix of degrees
    int32_t rowb[B];   // row start
    uint32_t gu[B];    // g(u) = f(u) - h(u), unsigned as GA_v2.cu:137
    int32_t node[B];
    int32_t tmp[16];
};

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

struct ExpandCounters {
    uint64_t expanded = 0, edges = 0, improved = 0;
    uint64_t rebuilds = 0;  // pile engine: far piles rebuilt after an overflow (lane 0 counts)
};

// Edges handled per lane and round (memory requests a lane keeps in flight
// per stage).  Four where registers allow it without lowering occupancy: the
// 1,024-lane pile kernel needs 110 VGPRs at four (101 at two; eight spills).
template <int B>
constexpr int expand_unroll()
{
    return B == 512 ? 2 : 4;
}

// Push policy interface:
//   void target(uint32_t f)                    -- t improved to f (t is never queued)
//   void touch(int32_t c, uint64_t prev)       -- c's word was prev before the update
//   void operator()(bool has, uint64_t key)    -- block-collective: all lanes call it
// Each lane brings up to P nodes u[p] (or -1) and their row spans (row_span,
// or the same values loaded earlier by the caller); node p of lane l is
// "virtual lane" l*P + p of a B*P-wide expansion, so a wide round pays the
// scan and barriers once for B*P rows.
template <int B, int HK, int P, class Push>
__device__ __forceinline__ void expand_rows(const GraphView &G, const HeurView &H, uint64_t *gp, int32_t t,
                                            const int32_t (&u)[P], const RowSpan (&rs)[P], ExpandLds<B * P> &L,
                                            ExpandCounters &C, Push &push)
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
    int32_t run = block_incl_scan<B>(lsum, L.tmp, total) - lsum;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        run += deg[p];
        L.incl[lane * P + p] = run;
        L.rowb[lane * P + p] = rs[p].b;
        L.gu[lane * P + p] = rs[p].gu;
        L.node[lane * P + p] = u[p];
    }
    __syncthreads();
    for (int32_t base = 0; base < total; base += B * R) {
        int32_t c[R], jj[R];
        uint32_t wt[R];
        // stage 1: locate each edge and load its column and weight
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int32_t e = base + r * B + lane;
            c[r] = -1;
            jj[r] = 0;
            wt[r] = 0;
            if (e < total) {
                int32_t lo = 0, hi = V - 1;
                while (lo < hi) {
                    const int32_t mid = (lo + hi) >> 1;
                    if (L.incl[mid] > e) hi = mid;
                    else lo = mid + 1;
                }
                const int32_t k = e - (lo ? L.incl[lo - 1] : 0);
                jj[r] = lo;
                const int32_t p = L.rowb[lo] + k;
                c[r] = G.col[p];
                wt[r] = G.w[p];
                if (G.dver && G.dver[p] <= G.ver) c[r] = -1;  // deleted at or before this version
                C.edges++;
            }
        }
        // stage 2: heuristic inputs and the current (f, parent) words
        int32_t hc[R];
        uint64_t cur[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hc[r] = 0;
            cur[r] = 0;
            if (c[r] >= 0) {  // negative id = deleted edge
                hc[r] = heur_eval<HK>(H, c[r]);
                cur[r] = ld_gp(gp + c[r]);
            }
        }
        // stage 3: one CAS attempt per improving edge, all issued together
        uint64_t want[R], got[R];
        bool tried[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            tried[r] = false;
            want[r] = 0;
            if (c[r] >= 0 && (HK != DDP_H_ARRAY || hc[r] >= 0)) {  // h = -1: cannot reach t
                const uint32_t cand = L.gu[jj[r]] + wt[r] + (uint32_t)hc[r];
                want[r] = ((uint64_t)cand << 32) | (uint32_t)L.node[jj[r]];
                tried[r] = (uint32_t)(cur[r] >> 32) > cand;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            got[r] = tried[r] ? (uint64_t)atomicCAS((unsigned long long *)(gp + c[r]), (unsigned long long)cur[r],
                                                   (unsigned long long)want[r])
                              : 0;
        // stage 4: settle contended CASes, then hand improvements to the push policy
#pragma unroll
        for (int r = 0; r < R; ++r) {
            bool has = false;
            uint64_t key = 0;
            if (tried[r]) {
                const uint32_t cand = (uint32_t)(want[r] >> 32);
                bool ok = got[r] == cur[r];
                uint64_t prev = cur[r];
                if (!ok) {
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
                    push.touch(c[r], prev);
                    if (c[r] == t) {
                        push.target(cand);
                    } else {
                        has = true;
                        key = ((uint64_t)cand << 32) | (uint32_t)c[r];
                    }
                }
            }
            push(has, key);
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
    __syncthreads();
}

template <int B, int HK, class Push>
__device__ __forceinline__ void expand_block(const GraphView &G, const HeurView &H, uint64_t *gp, int32_t t,
                                             int32_t u, uint32_t fu, ExpandLds<B> &L, ExpandCounters &C,
                                             Push &push)
{
    const int32_t uu[1] = {u};
    const RowSpan rs[1] = {row_span<HK>(G, H, u, fu)};
    expand_rows<B, HK, 1>(G, H, gp, t, uu, rs, L, C, push);
}
