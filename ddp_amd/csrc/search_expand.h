// Block-cooperative expansion of the nodes popped in one GA* iteration.
//
// Each lane of the block brings at most one popped node u (its queue's pop).
// Out-degrees are prefix-summed across the block and the union of all rows is
// walked B edges at a time, so lanes read consecutive CSR entries and a long
// row is split across the whole block instead of serialising one lane (the
// reference gives each thread its own row: GA_star/GA_v2.cu:109-167).
#pragma once

#include "device_common.h"

template <int B>
struct ExpandLds {
    int32_t incl[B];   // inclusive prefix of degrees
    int32_t rowb[B];   // base row start
    int32_t based[B];  // base row length
    int32_t drow[B];   // delta row start
    uint32_t gu[B];    // g(u) = f(u) - h(u), unsigned as GA_v2.cu:137
    int32_t node[B];
    int32_t tmp[16];
};

struct ExpandCounters {
    uint64_t expanded = 0, edges = 0, improved = 0;
};

// Push policy interface:
//   void target(uint32_t f)                    -- t improved to f (t is never queued)
//   void touch(int32_t c, uint64_t prev)       -- c's word was prev before the update
//   void operator()(bool has, uint64_t key)    -- block-collective: all lanes call it
template <int B, int HK, class Push>
__device__ __forceinline__ void expand_block(const GraphView &G, const HeurView &H, uint64_t *gp,
                                             int32_t t, int32_t u, uint32_t fu, ExpandLds<B> &L,
                                             ExpandCounters &C, Push &push)
{
    const int lane = threadIdx.x;
    int32_t deg = 0, b = 0, bd = 0, db = 0;
    uint32_t gu = 0;
    if (u >= 0) {
        const int32_t hu = heur_eval<HK>(H, u);
        gu = fu - (uint32_t)hu;
        b = G.off[u];
        bd = G.off[u + 1] - b;
        deg = bd;
        if (G.doff) {
            db = G.doff[u];
            deg += G.doff[u + 1] - db;
        }
        C.expanded++;
    }
    int32_t total;
    const int32_t incl = block_incl_scan<B>(deg, L.tmp, total);
    L.incl[lane] = incl;
    L.rowb[lane] = b;
    L.based[lane] = bd;
    L.drow[lane] = db;
    L.gu[lane] = gu;
    L.node[lane] = u;
    __syncthreads();
    for (int32_t base = 0; base < total; base += B) {
        const int32_t e = base + lane;
        bool has = false;
        uint64_t key = 0;
        if (e < total) {
            int32_t lo = 0, hi = B - 1;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (L.incl[mid] > e) hi = mid;
                else lo = mid + 1;
            }
            const int32_t j = lo;
            const int32_t k = e - (j ? L.incl[j - 1] : 0);
            int32_t c;
            uint32_t wt;
            if (k < L.based[j]) {
                const int32_t p = L.rowb[j] + k;
                c = G.col[p];
                wt = G.w[p];
            } else {
                const int32_t p = L.drow[j] + (k - L.based[j]);
                c = G.dcol[p];
                wt = G.dw[p];
            }
            C.edges++;
            if (c >= 0) {  // negative id = deleted edge
                const int32_t hc = heur_eval<HK>(H, c);
                if (HK != DDP_H_ARRAY || hc >= 0) {  // h = -1: cannot reach t
                    // unsigned relax expression of GA_star/GA_v2.cu:137
                    const uint32_t cand = L.gu[j] + wt + (uint32_t)hc;
                    uint64_t prev;
                    if (relax_gp(gp, c, cand, (uint32_t)L.node[j], &prev)) {
                        C.improved++;
                        push.touch(c, prev);
                        if (c == t) {
                            push.target(cand);
                        } else {
                            has = true;
                            key = ((uint64_t)cand << 32) | (uint32_t)c;
                        }
                    }
                }
            }
        }
        push(has, key);
    }
    __syncthreads();
}
