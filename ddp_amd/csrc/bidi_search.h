// Bidirectional pile search of one query in one workgroup ("slot").
//
// A forward search from s over the CSR rows and a backward search from t
// over the transposed rows run together: each round the workgroup takes up to
// B open entries, the first n0 lanes from the forward piles and the rest from
// the backward piles, and expands them in one cooperative pass (the edges of
// both directions share the prefix scan and the load stages).  Each direction
// keeps the near/far piles of slot_search_piles, its own (key, parent) array
// and touched log.  The reference has no counterpart: GA* searches from s
// only (GA_star/GA_v2.cu:483-575); the answer is the same exact cost.
//
// Keys use the average of the two closed-form heuristics as the potential:
//   p(v) = floor((h_t(v) - h_s(v)) / 2),
//   forward key  k_f(v) = g_f(v) + p(v),   backward key  k_b(v) = g_b(v) - p(v),
// where h_t estimates the distance to t and h_s the distance from s.  Both
// heuristics are consistent (w(u,v) >= h_t(u) - h_t(v) and
// w(u,v) >= h_s(v) - h_s(u)), so w(u,v) - p(u) + p(v) >= 0: every edge has a
// non-negative reduced cost in either direction, and k_f + k_b = g_f + g_b.
// Keys are never negative: g_f(v) >= h_s(v) makes k_f(v) >= (g_f + h_t)/2.
// With h = 0 this is plain bidirectional Dijkstra.
//
// Meeting.  When a direction lowers a node's key it then reads the other
// direction's word for that node (after its own atomic has completed, so of
// two concurrent improvements of one node at least one sees the other) and
// offers mu = k_f + k_b, the length of a real s-t path, to an LDS minimum.
//
// Bounds and termination.  Each direction's open keys never fall: an
// expansion only creates keys at least its own (non-negative reduced costs).
// So the smallest live key found by a full refill (near ring empty) bounds
// every key that direction will ever hold; it is published as L_d.  An entry
// with k_d + L_other >= mu cannot lead to a shorter path (any s-t path through
// it is at least that long, or it was already offered to mu), so it is
// dropped, and the search ends when L_f + L_b >= mu (no open entry on either
// side can improve mu) or when a side runs out (its bound is infinite).  This
// is the stopping rule of bidirectional Dijkstra on the reduced costs, and it
// stays exact for the label-correcting pile order because it only uses lower
// bounds of what is still open.
//
// Paths: forward parents lead from the meeting node back to s, backward
// parents from it on to t.
#pragma once

#include "slot_search.h"

// The forward potential p(v) from the inputs heur_load read for v.
template <int HK>
__device__ __forceinline__ int32_t pot_from(const HeurView &H, int32_t v, HeurIn in)
{
    if constexpr (HK == DDP_H_MANHATTAN) {
        const int32_t r = v / H.width, c = v - r * H.width;
        int32_t a = r - H.ty, b = c - H.tx, x = r - H.sy, y = c - H.sx;
        a = a < 0 ? -a : a;
        b = b < 0 ? -b : b;
        x = x < 0 ? -x : x;
        y = y < 0 ? -y : y;
        return ((a + b) * H.scale - (x + y) * H.scale) >> 1;  // arithmetic shift: floor
    } else if constexpr (HK == DDP_H_EUCLID) {
        const int32_t ht = (int32_t)isqrt_sum_sq((int64_t)in.a - H.tx, (int64_t)in.b - H.ty) * H.scale;
        const int32_t hs = (int32_t)isqrt_sum_sq((int64_t)in.a - H.sx, (int64_t)in.b - H.sy) * H.scale;
        return (ht - hs) >> 1;
    } else {
        return 0;
    }
}

template <int HK>
__device__ __forceinline__ void set_source(HeurView &hv, int32_t s)
{
    if constexpr (HK == DDP_H_MANHATTAN) {
        hv.sy = s / hv.width;
        hv.sx = s - hv.sy * hv.width;
    } else if constexpr (HK == DDP_H_EUCLID) {
        hv.sx = hv.xy[2 * s];
        hv.sy = hv.xy[2 * s + 1];
    }
}

// Pile state of one direction (the SlotLds pile fields).
struct BidiDir {
    uint32_t lring_head, lring_tail;  // LDS near ring positions (monotonic)
    uint32_t near_head, near_tail;    // HBM near ring positions (monotonic)
    uint32_t far_n, far_n2;           // entries in the current / next far pile
    uint32_t thr, delta;              // near window [.., thr) and its width
    uint32_t phase_n;                 // near entries taken since the last refill
    uint32_t ovf;                     // a push found both piles full and was dropped
    uint32_t no_early;                // nothing in the far pile adjoins the window
    int32_t ntouch;
};

template <int B>
struct BidiLds {
    ExpandLds<B> L;
    uint32_t red[16];
    unsigned long long meet;  // mu << 32 | meeting node
    uint32_t bound[2];        // L_d: lower bound of direction d's open keys (UINT32_MAX: none left)
    uint32_t err;
    int32_t plen;
    int32_t iters;
    BidiDir d[2];
};

constexpr uint64_t kNoMeet = ((uint64_t)DDP_INF_F << 32) | 0xFFFFFFFFull;

// a[one] for a direction that differs between lanes: both elements are read
// with constant indices (uniform values) and one is selected per lane.
// Indexing the pair with a per-lane value instead puts the whole structure in
// scratch memory, and every such read becomes a memory round trip ahead of
// the load it feeds.
template <class T>
__device__ __forceinline__ T pick2(const T (&a)[2], bool one)
{
    const T x0 = a[0], x1 = a[1];
    return one ? x1 : x0;
}

// Everything a round needs about the two directions.  With a versioned view
// (the server), base entry j of direction d is visible iff dver[d][j] > ver,
// and inserted edges hang off per-node lists (direction 0: by source, 1: by
// target) whose links pack entry << 32 | inserting version; entry x is
// visible iff its link version <= ver < ldel[x], and its neighbour in
// direction d is lnb[d][x].
struct BidiSide {
    const int32_t *off[2];
    const int32_t *col[2];
    const uint32_t *w[2];
    uint64_t *gp[2];
    int32_t *touched[2];
    PilePtrs P[2];  // by value: kept in registers (a pointer would pin them in memory)
    const int32_t *dver[2];
    int32_t ver;
    const uint64_t *lhead[2];
    const uint64_t *lnext[2];
    const int32_t *lnb[2];
    const uint32_t *lw;
    const int32_t *ldel;
};

// After a round's relaxations: R edges per lane, ok[r] if the lane lowered
// node c[r] of direction dr[r] to key want[r] >> 32 (parent in the low word),
// first[r] if that node's word left INF.  Reads the other direction's word of
// every improved node (after this lane's atomics have completed, so of two
// concurrent improvements of one node at least one sees the other), logs first
// touches, queues the entries that can still matter and offers the meetings.
// Wave-collective.
template <int B, int R>
__device__ __forceinline__ void bidi_settle(const BidiSide &D, int32_t s, int32_t t, const bool (&ok)[R],
                                            const bool (&first)[R], const bool (&dr)[R], const int32_t (&c)[R],
                                            const uint64_t (&want)[R], const uint32_t (&F)[2],
                                            uint64_t *const (&far)[2], BidiLds<B> &S)
{
    wait_vmem();
    uint64_t oth[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (ok[r]) oth[r] = ld_gp(pick2(D.gp, !dr[r]) + c[r]);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        bool fd[R], put[R];
        uint64_t key[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool mine = ok[r] && (dr[r] ? 1 : 0) == d;
            fd[r] = mine && first[r];
            key[r] = (want[r] & 0xFFFFFFFF00000000ull) | (uint32_t)c[r];
            // the other search's root is never queued: paths end there
            put[r] = mine && c[r] != (d ? s : t) && (uint32_t)(want[r] >> 32) < F[d];
        }
        uint32_t tp[R];
        wave_append<R>(fd, (uint32_t *)&S.d[d].ntouch, tp);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (fd[r]) D.touched[d][tp[r]] = c[r];
        BidiDir &Q = S.d[d];
        const bool lds_ok = D.P[d].lcap && Q.near_tail == Q.near_head;
        if (pile_place<R>(D.P[d], Q.thr, lds_ok, Q.lring_head, Q.near_head, &Q.lring_tail, &Q.near_tail, far[d],
                          &Q.far_n, put, key))
            Q.ovf = 1;
    }
    // meetings: k_f + k_b of every improved node the other side has reached
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (ok[r] && (uint32_t)(oth[r] >> 32) != DDP_INF_F) {
            const uint64_t mu = (want[r] >> 32) + (oth[r] >> 32);
            if (mu < DDP_INF_F) atomicMin(&S.meet, (unsigned long long)((mu << 32) | (uint32_t)c[r]));
        }
    }
}

// One cooperative pass over the rows of the entries taken this round: lanes
// below n0 hold forward nodes, the others backward nodes.  Loads are issued
// stage by stage as in expand_rows.  With VERS, deleted base entries are
// skipped and each lane then walks its node's inserted-edge list.  No closing
// barrier.
template <int B, int HK, bool VERS>
__device__ __forceinline__ void expand_bidi(const BidiSide &D, const HeurView &H, bool relax_min, int32_t s,
                                            int32_t t, int32_t u, RowSpan rs, int32_t n0, const uint32_t (&F)[2],
                                            uint64_t *const (&far)[2], BidiLds<B> &S, ExpandCounters &C)
{
    constexpr int R = expand_unroll<B>();
    const int lane = threadIdx.x;
    ExpandLds<B> &L = S.L;
    const int32_t deg = u >= 0 ? rs.deg : 0;
    if (u >= 0) C.expanded++;
    int32_t total;
    const int32_t run0 = block_incl_scan<B, false>(deg, L.tmp, total) - deg;
    L.incl[lane] = run0 + deg;
    L.rowb[lane] = rs.b;
    L.gu[lane] = rs.gu;
    L.node[lane] = u;
    const bool use_own = total <= 4 * B * R;
    // a round of at most one edge per lane is latency-bound: relax without the
    // word read, the atomicMin's returned word telling whether it improved
    // (heuristics without loads; as relax_window's DIRECT)
    const bool direct = (HK == DDP_H_ZERO || HK == DDP_H_MANHATTAN) && relax_min && total <= B;
    if (use_own) {
        const int32_t to = deg < kOwnRun ? run0 + deg : run0 + kOwnRun;
        for (int32_t x = run0; x < to; ++x) L.own[x] = (uint16_t)lane;
    }
    __syncthreads();
    int32_t prow = 0;
    for (int32_t base = 0; base < total; base += B * R) {
        int32_t pos[R], jj[R], c[R];
        bool dr[R];
        uint32_t wt[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int32_t e = base + r * B + lane;
            pos[r] = -1;
            jj[r] = 0;
            if (e < total) {
                int32_t lo = use_own ? L.own[e] : prow;
                int32_t ex = 0;
                if (lo < B) ex = lo ? L.incl[lo - 1] : 0;
                if (!(lo < B && ex <= e && e < L.incl[lo])) {
                    int32_t a = prow, hi = B - 1;
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
            dr[r] = jj[r] >= n0;
            if (pos[r] >= 0) {
                const int k = dr[r] ? 1 : 0;
                c[r] = pick2(D.col, k != 0)[pos[r]];
                wt[r] = pick2(D.w, k != 0)[pos[r]];
                if constexpr (VERS) dv[r] = pick2(D.dver, k != 0)[pos[r]];
            }
        }
        // stage 2: heuristic inputs and current words (written only by lanes that take part)
        HeurIn hin[R];
        uint64_t cur[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (VERS && dv[r] <= D.ver) c[r] = -1;  // deleted at or before this version
            if (c[r] >= 0) {
                hin[r] = heur_load<HK>(H, c[r]);
                if (!direct) cur[r] = ld_gp(pick2(D.gp, dr[r]) + c[r]);
            }
        }
        // stage 3: candidates and one atomic per improving edge
        uint64_t want[R], got[R];
        bool tried[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            tried[r] = false;
            want[r] = 0;
            if (c[r] >= 0) {
                const int32_t pc = pot_from<HK>(H, c[r], hin[r]);
                const uint32_t cand = L.gu[jj[r]] + wt[r] + (uint32_t)(dr[r] ? -pc : pc);
                want[r] = ((uint64_t)cand << 32) | (uint32_t)L.node[jj[r]];
                tried[r] = direct || (uint32_t)(cur[r] >> 32) > cand;
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): every load above was consumed
        if (relax_min) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tried[r])
                    got[r] = (uint64_t)atomicMin((unsigned long long *)(pick2(D.gp, dr[r]) + c[r]),
                                                 (unsigned long long)want[r]);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tried[r])
                    got[r] = (uint64_t)atomicCAS((unsigned long long *)(pick2(D.gp, dr[r]) + c[r]),
                                                 (unsigned long long)cur[r], (unsigned long long)want[r]);
        }
        // stage 4: settle contended CASes
        bool ok[R], first[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ok[r] = false;
            first[r] = false;
            if (tried[r]) {
                const uint32_t cand = (uint32_t)(want[r] >> 32);
                bool good = relax_min ? (uint32_t)(got[r] >> 32) > cand : got[r] == cur[r];
                uint64_t prev = relax_min ? got[r] : cur[r];
                if (!good && !relax_min) {
                    uint64_t now = got[r];
                    while ((uint32_t)(now >> 32) > cand) {
                        const uint64_t old =
                            atomicCAS((unsigned long long *)(pick2(D.gp, dr[r]) + c[r]), (unsigned long long)now,
                                      (unsigned long long)want[r]);
                        if (old == now) {
                            good = true;
                            prev = now;
                            break;
                        }
                        now = old;
                    }
                }
                ok[r] = good;
                if (good) {
                    C.improved++;
                    first[r] = (uint32_t)(prev >> 32) == DDP_INF_F;
                }
            }
        }
        bidi_settle<B, R>(D, s, t, ok, first, dr, c, want, F, far, S);
    }
    if constexpr (VERS) {
        // Inserted edges: each lane walks its own node's list, oldest first,
        // stopping at the first entry newer than this version.
        const int d = lane >= n0 ? 1 : 0;
        uint64_t link =
            u >= 0 ? __hip_atomic_load(pick2(D.lhead, d != 0) + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0ull;
        for (;;) {
            bool active = false;
            bool ok[1] = {false}, first[1] = {false}, dr[1] = {d != 0};
            int32_t cn[1] = {-1};
            uint64_t want[1] = {0};
            int32_t x = -1;
            while ((uint32_t)link <= (uint32_t)D.ver) {
                x = (int32_t)(link >> 32);
                link = __hip_atomic_load(pick2(D.lnext, d != 0) + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (D.ldel[x] > D.ver) {
                    active = true;
                    break;
                }
            }
            if (active) {
                C.edges++;
                cn[0] = pick2(D.lnb, d != 0)[x];
                const int32_t pc = pot_from<HK>(H, cn[0], heur_load<HK>(H, cn[0]));
                const uint32_t cand = rs.gu + D.lw[x] + (uint32_t)(d ? -pc : pc);
                uint64_t prev;
                if (relax_gp(pick2(D.gp, d != 0), cn[0], cand, (uint32_t)u, &prev)) {
                    C.improved++;
                    ok[0] = true;
                    first[0] = (uint32_t)(prev >> 32) == DDP_INF_F;
                    want[0] = ((uint64_t)cand << 32) | (uint32_t)u;
                }
            }
            bidi_settle<B, 1>(D, s, t, ok, first, dr, cn, want, F, far, S);
            if (!__syncthreads_or(active)) break;
        }
    }
}

// Refill of direction d's near rings from its far pile (the refill of
// slot_search_piles); a full refill (near rings empty) publishes the
// direction's bound.  Block-collective; ends with a barrier.  The direction
// is a template argument so that every pair is indexed by a constant.
template <int B, int HK, int d>
__device__ __forceinline__ void bidi_refill(const BidiSide &D, int32_t oroot, bool early, uint32_t F, uint32_t lh,
                                            uint32_t hh, uint64_t *(&far)[2], uint64_t *(&far2)[2], BidiLds<B> &S,
                                            ExpandCounters &C)
{
    constexpr uint32_t Vd = B / 2;  // lanes a direction gets when both have work
    const int lane = threadIdx.x;
    BidiDir &Q = S.d[d];
    const PilePtrs &P = D.P[d];
    uint64_t *gp = D.gp[d];
    if (Q.ovf) {
        // rebuild the far pile from the touched log (entries were dropped)
        const int32_t nt = Q.ntouch;
        if (lane == 0) Q.far_n2 = 0;
        __syncthreads();
        for (int32_t i0 = 0; i0 < nt; i0 += B) {
            const int32_t i = i0 + lane;
            int32_t v = -1;
            uint32_t f = 0xFFFFFFFFu;
            if (i < nt) {
                v = __hip_atomic_load(D.touched[d] + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                f = (uint32_t)(ld_gp(gp + v) >> 32);
            }
            const uint32_t idx = far_reserve(f < F && v != oroot, &Q.far_n2);  // the other root is never queued
            if (idx == 0xFFFFFFFFu) continue;
            if (idx < P.fcap) far2[d][idx] = ((uint64_t)f << 32) | (uint32_t)v;
            else atomicOr(&S.err, (uint32_t)QERR_HEAP);
        }
        block_sync_global();
        if (lane == 0) {
            Q.far_n = Q.far_n2;
            Q.ovf = 0;
            C.rebuilds++;
        }
        uint64_t *tmp = far[d];
        far[d] = far2[d];
        far2[d] = tmp;
        __syncthreads();
        if (S.err) return;
    }
    const uint32_t nf = Q.far_n < P.fcap ? Q.far_n : P.fcap;
    if (lane == 0) C.refills++;
    C.scanned += 2 * ((nf > (uint32_t)lane) ? (nf - 1 - lane) / B + 1 : 0);
    constexpr int KC = 2;
    uint64_t *fp = far[d];
    auto load_chunk = [&](uint32_t base, uint64_t (&k)[KC], bool (&live)[KC]) {
        uint64_t wv[KC];
#pragma unroll
        for (int j = 0; j < KC; ++j)
            if (base + (uint32_t)j * B + lane < nf) k[j] = fp[base + (uint32_t)j * B + lane];
#pragma unroll
        for (int j = 0; j < KC; ++j)
            if (base + (uint32_t)j * B + lane < nf && (uint32_t)(k[j] >> 32) < F) wv[j] = ld_gp(gp + (uint32_t)k[j]);
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            live[j] = false;
            if (base + (uint32_t)j * B + lane < nf) {
                const uint32_t f = (uint32_t)(k[j] >> 32);
                live[j] = f < F && (uint32_t)(wv[j] >> 32) == f;
            } else {
                k[j] = 0;
            }
        }
    };
    uint32_t mymin = 0xFFFFFFFFu;
    for (uint32_t base = 0; base < nf; base += (uint32_t)KC * B) {
        uint64_t k2[KC];
        bool l2[KC];
        load_chunk(base, k2, l2);
#pragma unroll
        for (int j = 0; j < KC; ++j)
            if (l2[j]) mymin = (uint32_t)(k2[j] >> 32) < mymin ? (uint32_t)(k2[j] >> 32) : mymin;
    }
    const uint32_t fmin = block_min_u32<B>(mymin, S.red);
    if (fmin >= F) {
        // nothing live in the far pile: bounds only rise and mu only falls, so
        // those entries are hopeless for good
        if (lane == 0) {
            Q.far_n = 0;
            if (!early) S.bound[d] = 0xFFFFFFFFu;  // near rings empty too: this side is done
        }
        __syncthreads();
        return;
    }
    if (early && fmin >= sat_add_u32(Q.thr, Q.delta)) {
        if (lane == 0) Q.no_early = 1;
        __syncthreads();
        return;
    }
    if (lane == 0) {
        Q.no_early = 0;
        if (!early) {
            if (Q.phase_n < Vd) Q.delta = sat_add_u32(Q.delta, Q.delta);
            else if (Q.phase_n > 8 * Vd && Q.delta > 1) Q.delta >>= 1;
            Q.phase_n = 0;
            S.bound[d] = fmin;  // every open key of this side is in the far pile now
        }
        const uint32_t nt = sat_add_u32(fmin, Q.delta);
        Q.thr = nt > Q.thr ? nt : Q.thr;
        Q.far_n2 = 0;
    }
    __syncthreads();
    const uint32_t thr = Q.thr;
    const bool lds_ok = P.lcap != 0 && Q.near_tail == hh;
    for (uint32_t base = 0; base < nf; base += (uint32_t)KC * B) {
        uint64_t k2[KC];
        bool l2[KC];
        load_chunk(base, k2, l2);
        (void)pile_place<KC>(P, thr, lds_ok, lh, hh, &Q.lring_tail, &Q.near_tail, far2[d], &Q.far_n2, l2, k2);
    }
    block_sync_global();
    if (lane == 0) Q.far_n = Q.far_n2;
    uint64_t *tmp = far[d];
    far[d] = far2[d];
    far2[d] = tmp;
    __syncthreads();
}

// Runs one query to completion.  On return (block-uniform) S.meet holds
// mu << 32 | meeting node (kNoMeet if t is unreachable; for s == t, h(s) << 32
// | s) and S.err the error flags; both (key, parent) arrays and touched logs
// still describe the search for slot_path_bidi and slot_reset_bidi.
template <int B, int HK, bool VERS = false>
__device__ __forceinline__ void slot_search_bidi(const BidiSide &D, HeurView hv, bool relax_min, int32_t N, int32_t s, int32_t t,
                                 int32_t maxit, uint32_t delta0, BidiLds<B> &S, ExpandCounters &C, uint32_t &stale)
{
    constexpr uint32_t Vd = B / 2;
    const int lane = threadIdx.x;
    if (lane == 0) {
        S.meet = kNoMeet;
        S.bound[0] = S.bound[1] = 0;
        S.err = 0;
        S.iters = 0;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            S.d[d] = BidiDir{};
            S.d[d].delta = delta0 > 0 ? delta0 : 1;
        }
    }
    __syncthreads();
    bool run = s >= 0 && s < N && t >= 0 && t < N;
    if (run) {
        set_target<HK>(hv, t);
        set_source<HK>(hv, s);
        if (s == t) {
            run = false;
            if (lane == 0) S.meet = ((uint64_t)(uint32_t)heur_eval<HK>(hv, s) << 32) | (uint32_t)s;
        } else if (lane == 0) {
            const int32_t ps = pot_from<HK>(hv, s, heur_load<HK>(hv, s));
            const int32_t pt = pot_from<HK>(hv, t, heur_load<HK>(hv, t));
            const uint32_t k0 = (uint32_t)ps, k1 = (uint32_t)(-pt);  // g = 0 at the roots
            const int32_t root[2] = {s, t};
            const uint32_t k[2] = {k0, k1};
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const uint64_t key = ((uint64_t)k[d] << 32) | (uint32_t)root[d];
                D.gp[d][root[d]] = ((uint64_t)k[d] << 32) | 0xFFFFFFFFull;
                D.touched[d][0] = root[d];
                S.d[d].ntouch = 1;
                if (D.P[d].lcap) {
                    D.P[d].lring[0] = key;
                    S.d[d].lring_tail = 1;
                } else {
                    D.P[d].near[0] = key;
                    S.d[d].near_tail = 1;
                }
                S.d[d].thr = sat_add_u32(k[d], S.d[d].delta);
                S.bound[d] = k[d];
            }
        }
    } else if (lane == 0) {
        S.err = QERR_BADQ;
    }
    block_sync_global();
    uint64_t *far[2] = {D.P[0].far0, D.P[1].far0}, *far2[2] = {D.P[0].far1, D.P[1].far1};
    uint32_t lh[2] = {0, 0}, hh[2] = {0, 0};  // ring heads (uniform)
    int32_t it = 0;
    while (run) {
        // Every LDS word the round's decisions need, read together: they are
        // independent, so one wait covers them (read one by one between the
        // branches, they cost about 1,000 cycles of a 14,000-cycle C2 round).
        // The counters were last changed by the previous round's LDS atomics,
        // which its closing barrier ordered; only HBM ring contents need the
        // store drain below.
        const uint64_t meet = S.meet;
        const uint32_t b0 = S.bound[0], b1 = S.bound[1];
        uint32_t lt[2], nt[2], fn[2], ov[2], ne[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            lt[d] = S.d[d].lring_tail;
            nt[d] = S.d[d].near_tail;
            fn[d] = S.d[d].far_n;
            ov[d] = S.d[d].ovf;
            ne[d] = S.d[d].no_early;
        }
        const uint32_t mu = (uint32_t)(meet >> 32);
        if ((uint64_t)b0 + b1 >= mu) break;  // no open entry on either side can shorten mu
        const uint32_t F[2] = {b1 >= mu ? 0u : mu - b1, b0 >= mu ? 0u : mu - b0};
        bool from_lds[2];
        uint32_t avail[2];
        bool any_hbm = false;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            from_lds[d] = lt[d] != lh[d];
            any_hbm |= !from_lds[d];
        }
        if (any_hbm) block_sync_global();  // HBM entries ahead: every wave's stores have landed
        int rd = -1;
        bool early = false;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            avail[d] = from_lds[d] ? lt[d] - lh[d] : nt[d] - hh[d];
            const bool e = avail[d] > 0 && (lt[d] - lh[d]) + (nt[d] - hh[d]) < Vd / kThinRing && fn[d] > 0 &&
                           !ov[d] && !ne[d];
            if (rd < 0 && (avail[d] == 0 || e)) {
                rd = d;
                early = avail[d] > 0;
            }
        }
        if (rd >= 0) {
            if (early && from_lds[rd]) block_sync_global();  // the far pile is read below
            if (rd) bidi_refill<B, HK, 1>(D, s, early, F[1], lh[1], hh[1], far, far2, S, C);
            else bidi_refill<B, HK, 0>(D, t, early, F[0], lh[0], hh[0], far, far2, S, C);
            if (S.err) break;
            continue;
        }
        if (it >= maxit) {
            if (lane == 0) atomicOr(&S.err, (uint32_t)QERR_ITERS);
            break;
        }
        ++it;
        // lanes: each side gets at least half when it has that many
        const uint32_t n0 = min(avail[0], (uint32_t)B - min(avail[1], Vd));
        const uint32_t n1 = min(avail[1], (uint32_t)B - n0);
        int32_t u = -1;
        RowSpan rs;
        if ((uint32_t)lane < n0 + n1) {
            const int d = (uint32_t)lane >= n0 ? 1 : 0;
            const uint32_t idx = d ? lane - n0 : lane;
            // values, not lvalues, are selected: a conditional over two fields
            // selects their addresses, and a load at a per-lane offset pins D
            // in scratch memory
            const bool one = d != 0;
            uint64_t *const lring0 = D.P[0].lring, *const lring1 = D.P[1].lring;
            uint64_t *const near0 = D.P[0].near, *const near1 = D.P[1].near;
            const uint32_t lcap0 = D.P[0].lcap, lcap1 = D.P[1].lcap, ncap0 = D.P[0].ncap, ncap1 = D.P[1].ncap;
            const uint64_t key = pick2(from_lds, one)
                                     ? (one ? lring1 : lring0)[(pick2(lh, one) + idx) & ((one ? lcap1 : lcap0) - 1)]
                                     : (one ? near1 : near0)[(pick2(hh, one) + idx) & ((one ? ncap1 : ncap0) - 1)];
            const uint32_t f = (uint32_t)(key >> 32);
            const int32_t node = (int32_t)(uint32_t)key;
            // the node's word, row and potential input in one round trip
            const uint64_t word = ld_gp(pick2(D.gp, one) + node);
            const int32_t *off = pick2(D.off, one);
            const int32_t rb = off[node], re = off[node + 1];
            const int32_t pn = pot_from<HK>(hv, node, heur_load<HK>(hv, node));
            if (f < pick2(F, one) && (uint32_t)(word >> 32) == f) {
                u = node;
                rs.b = rb;
                rs.deg = re - rb;
                rs.gu = f - (uint32_t)(d ? -pn : pn);
            } else {
                ++stale;
            }
        }
        // Pushes of this round start after expand_bidi's first barrier, when
        // every lane has read its entry, so the freed slots may be reused.
        const uint32_t nd[2] = {n0, n1};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            if (from_lds[d]) lh[d] += nd[d];
            else hh[d] += nd[d];
        }
        if (lane == 0) {
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                S.d[d].lring_head = lh[d];
                S.d[d].near_head = hh[d];
                S.d[d].phase_n += nd[d];
            }
        }
        expand_bidi<B, HK, VERS>(D, hv, relax_min, s, t, u, rs, (int32_t)n0, F, far, S, C);
        __syncthreads();  // the round's pushes, meetings and errors are visible to the next round
        if (S.err) break;
    }
    if (lane == 0) S.iters = it;
    block_sync_global();
}

// Lane 0 writes the path s..t through the meeting node into out[0..cap) and
// its length into S.plen (0 if unreachable, -len if longer than cap).
template <int B>
__device__ void slot_path_bidi(const uint64_t *gpf, const uint64_t *gpb, int32_t N, int32_t s, int32_t t,
                               int32_t *out, int32_t cap, BidiLds<B> &S)
{
    if (threadIdx.x != 0) return;
    const uint64_t meet = S.meet;
    int32_t n = 0;
    if ((uint32_t)(meet >> 32) != DDP_INF_F) {
        const int32_t m = (int32_t)(uint32_t)meet;
        // m back to s along forward parents
        int32_t v = m;
        for (int32_t step = 0; step <= N; ++step) {
            if (n < cap) out[n] = v;
            ++n;
            if (v == s) break;
            const uint32_t p = (uint32_t)ld_gp(gpf + v);
            if (p >= (uint32_t)N) {
                n = -1;
                break;
            }
            v = (int32_t)p;
        }
        if (n > 0 && n <= cap)
            for (int32_t i = 0; i < n / 2; ++i) {  // m..s -> s..m
                const int32_t x = out[i];
                out[i] = out[n - 1 - i];
                out[n - 1 - i] = x;
            }
        // m on to t along backward parents
        v = m;
        for (int32_t step = 0; n > 0 && v != t && step <= N; ++step) {
            const uint32_t p = (uint32_t)ld_gp(gpb + v);
            if (p >= (uint32_t)N) {
                n = -1;
                break;
            }
            v = (int32_t)p;
            if (n < cap) out[n] = v;
            ++n;
        }
        if (n > 0 && v != t) n = -1;
        if (n < 0) {
            n = 0;
            atomicOr(&S.err, (uint32_t)QERR_PATH);
        }
    }
    S.plen = (n > cap) ? -n : n;
}

// Clears every word either direction touched (barriers on both ends, as slot_reset).
template <int B>
__device__ void slot_reset_bidi(uint64_t *gpf, uint64_t *gpb, const int32_t *tf, const int32_t *tb, BidiLds<B> &S)
{
    __syncthreads();
    const int32_t n0 = S.d[0].ntouch, n1 = S.d[1].ntouch;
    for (int32_t i = (int32_t)threadIdx.x; i < n0; i += B)
        gpf[__hip_atomic_load(tf + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)] = DDP_GP_INF;
    for (int32_t i = (int32_t)threadIdx.x; i < n1; i += B)
        gpb[__hip_atomic_load(tb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)] = DDP_GP_INF;
    block_sync_global();
}
