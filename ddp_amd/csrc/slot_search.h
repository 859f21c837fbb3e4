// One query searched by one workgroup ("slot"): used by the batched kernel
// and the serving kernel.  The slot owns an N-entry (f, parent) array, a log
// of the nodes it touched (reset when the query ends, so the array never
// needs a full clear) and B lane-owned heaps; an improved node is pushed into
// the heap of the lane that relaxed it, so a search never leaves its
// workgroup.
#pragma once

#include "search_expand.h"

enum : uint32_t { QERR_HEAP = 1, QERR_ITERS = 2, QERR_BADQ = 4, QERR_PATH = 8 };

// V: expansion width (B lanes x entries per lane; the pile engine can take
// more than one entry per lane and round).
template <int B, int V = B>
struct SlotLds {
    ExpandLds<V> L;
    uint32_t red[16];
    uint32_t F;        // best f of the target so far
    uint32_t err;
    int32_t ntouch;
    int32_t plen;
    int32_t iters;     // iterations of the last search (pop rounds)
    // pile engine (slot_search_piles)
    uint32_t lring_head, lring_tail;  // LDS near ring positions (monotonic)
    uint32_t near_head, near_tail;  // HBM near ring positions (monotonic)
    uint32_t far_n, far_n2;         // entries in the current / next far pile
    uint32_t thr, delta;            // near window [.., thr) and its width
    uint32_t phase_n;               // near entries taken since the last refill
    uint32_t ovf;                   // a push found both piles full and was dropped
    uint32_t no_early;              // the far pile holds nothing near the window: no early refill
    uint32_t dorm_n;                // resumed searches: entries set aside in the dormant pile
};

template <int B>
struct SlotPush {
    const LaneHeap *H;
    int32_t *size;
    uint32_t *s_F;
    int32_t *touched;
    int32_t *s_ntouch;
    uint32_t err = 0;

    __device__ void target(uint32_t f) { atomicMin(s_F, f); }
    __device__ void touch(int32_t c, uint64_t prev)
    {
        if ((uint32_t)(prev >> 32) == DDP_INF_F) touched[atomicAdd(s_ntouch, 1)] = c;
    }
    __device__ void operator()(bool has, uint64_t key)
    {
        if (has && !heap_push(*H, *size, key)) err |= QERR_HEAP;
    }
    template <int R>
    __device__ void touch_all(const bool (&first)[R], const int32_t (&c)[R])
    {
        uint32_t pos[R];
        wave_append<R>(first, (uint32_t *)s_ntouch, pos);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (first[r]) touched[pos[r]] = c[r];
    }
    template <int R>
    __device__ void all(const bool (&has)[R], const uint64_t (&key)[R])
    {
#pragma unroll
        for (int r = 0; r < R; ++r) (*this)(has[r], key[r]);
    }
};

template <int HK>
__device__ __forceinline__ void set_target(HeurView &hv, int32_t t)
{
    if constexpr (HK == DDP_H_MANHATTAN) {
        hv.ty = t / hv.width;
        hv.tx = t - hv.ty * hv.width;
    } else if constexpr (HK == DDP_H_EUCLID) {
        hv.tx = hv.xy[2 * t];
        hv.ty = hv.xy[2 * t + 1];
    }
}

// Runs one query to completion.  On return (block-uniform) S.F holds the
// target's cost (DDP_INF_F if unreachable) and S.err the error flags; gp and
// touched still describe the search so the caller can walk the path.
template <int B, int HK>
__device__ void slot_search(const GraphView &G, HeurView hv, uint64_t *gp, int32_t *touched, const LaneHeap &H,
                            uint64_t *s_heap, int32_t N, int32_t s, int32_t t, int32_t maxit, SlotLds<B> &S,
                            ExpandCounters &C, uint32_t &stale)
{
    const int lane = threadIdx.x;
    if (lane == 0) {
        S.F = DDP_INF_F;
        S.ntouch = 0;
        S.err = 0;
        S.iters = 0;
    }
    __syncthreads();
    bool run = s >= 0 && s < N && t >= 0 && t < N;
    int32_t size = 0;
    if (run) {
        set_target<HK>(hv, t);
        const int32_t hs = heur_eval<HK>(hv, s);
        if (HK == DDP_H_ARRAY && hs < 0) {
            run = false;  // h(s) = -1: s cannot reach t
        } else if (s == t) {
            run = false;
            if (lane == 0) S.F = (uint32_t)hs;
        } else if (lane == 0) {
            gp[s] = ((uint64_t)(uint32_t)hs << 32) | 0xFFFFFFFFull;
            touched[0] = s;
            S.ntouch = 1;
            heap_push(H, size, ((uint64_t)(uint32_t)hs << 32) | (uint32_t)s);
        }
    } else if (lane == 0) {
        S.err = QERR_BADQ;
    }
    block_sync_global();
    SlotPush<B> push{&H, &size, &S.F, touched, &S.ntouch};
    if (run) {
        for (int32_t it = 0;; ++it) {
            const uint32_t mytop = size > 0 ? (uint32_t)(s_heap[lane] >> 32) : 0xFFFFFFFFu;
            const uint32_t topmin = block_min_u32<B>(mytop, S.red);
            if constexpr (B == 64) __syncthreads();
            if (S.F <= topmin || it >= maxit) {  // no queued node can improve t, or out of iterations
                if (lane == 0) {
                    S.iters = it;
                    if (S.F > topmin) atomicOr(&S.err, (uint32_t)QERR_ITERS);
                }
                break;
            }
            int32_t u = -1;
            uint32_t fu = 0;
            while (size > 0) {
                uint64_t key;
                heap_pop(H, size, key);
                const int32_t node = (int32_t)(uint32_t)key;
                const uint32_t f = (uint32_t)(key >> 32);
                if ((uint32_t)(ld_gp(gp + node) >> 32) == f) {
                    u = node;
                    fu = f;
                    break;
                }
                ++stale;
            }
            expand_block<B, HK>(G, hv, gp, t, u, fu, S.L, C, push);
            if (push.err) {
                atomicOr(&S.err, push.err);
                push.err = 0;
            }
        }
    }
    block_sync_global();
}

// ---------------------------------------------------------------- piles
//
// A second queue discipline for one-slot searches: no heaps.  Open entries
// (f << 32 | node) sit in two piles: a "near" ring holding the entries with f
// below a threshold (its front in LDS, the rest in HBM) and a "far" pile in
// HBM with the rest.  Each round the
// workgroup takes up to B near entries at once (one per lane, a coalesced
// read), drops stale ones (f no longer the node's word) and those with
// f >= F(t), and expands the rest cooperatively; an improvement goes to the
// near ring when its f is below the threshold, else to the far pile.  When the
// near ring runs dry (or thin, kThinRing, while the next window adjoins), one
// pass over the far pile finds the smallest live f, the threshold moves to it
// plus a window width, and a second pass moves the entries under it to the
// near ring and compacts the rest into the other far buffer.  The search ends when both piles are empty or every live far entry
// has f >= F(t).  This is a label-correcting search (a node can be expanded
// again after a later improvement, as GA*'s K queues also allow); it returns
// the optimal cost for an admissible heuristic because an entry is only
// dropped when it is stale or cannot lead to anything below F(t).  The window
// width adapts: doubled when a phase fed the lanes fewer than B entries,
// halved when it ran past 8B (a sweep on the C4 batch: narrower windows
// expand fewer nodes out of order, down to where rounds run short of work).
// Overflow: a push that finds the near ring full goes to the far pile; one
// that finds the far pile full too is dropped and flagged, and the next refill
// rebuilds the far pile from the slot's touched log (every node whose word
// left INF, at most N) with its current f.  That re-queues some nodes that were
// already expanded at their final f, which is harmless in a label-correcting
// search, so results stay exact; only a far pile smaller than the touched set
// is an error (QERR_HEAP).

// Ring reservation by one lane for its wave: up to `want` slots of a ring
// whose tail may never run more than cap ahead of head (it may grant fewer).
__device__ __forceinline__ void ring_reserve(uint32_t *tail, uint32_t head, uint32_t cap, uint32_t want,
                                             uint32_t &base, uint32_t &got)
{
    uint32_t cur = __hip_atomic_load(tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
        const uint32_t used = cur - head;
        got = used < cap ? (cap - used < want ? cap - used : want) : 0u;
        if (got == 0) break;
        if (__hip_atomic_compare_exchange_weak(tail, &cur, cur + got, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP))
            break;
    }
    base = cur;
}

// Wave-aggregated append (one LDS atomic per wave); returns the lane's index,
// or UINT32_MAX if it wanted none.  Every lane of the wave must call it.
__device__ __forceinline__ uint32_t far_reserve(bool want, uint32_t *count)
{
    const uint64_t m = __ballot(want);
    if (m == 0) return 0xFFFFFFFFu;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    return want ? base + lane_rank(m) : 0xFFFFFFFFu;
}

struct PilePtrs {
    uint64_t *near;  // ncap entries (power of two), ring in HBM
    uint64_t *far0;  // fcap entries
    uint64_t *far1;  // fcap entries
    uint32_t ncap, fcap;
    uint64_t *lring = nullptr;  // lcap entries (power of two, or 0), ring in LDS
    uint32_t lcap = 0;
};

// Places up to R entries per lane with one reservation per level and wave:
// entries with f < thr fill the LDS ring (if lds_ok), then the HBM ring;
// what does not fit, and entries with f >= thr, are appended to fdst at
// *fcount.  Returns true if some entry found fdst full and was dropped.
// Every lane of the wave must call it.
template <int R>
__device__ __forceinline__ bool pile_place(const PilePtrs &P, uint32_t thr, bool lds_ok, uint32_t lhead,
                                           uint32_t hhead, uint32_t *ltail, uint32_t *htail, uint64_t *fdst,
                                           uint32_t *fcount, const bool (&put)[R], const uint64_t (&key)[R])
{
    bool nr[R], fr[R];
    uint64_t m[R];
    uint32_t off[R], tot = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        nr[r] = put[r] && (uint32_t)(key[r] >> 32) < thr;
        fr[r] = put[r] && !nr[r];
        m[r] = __ballot(nr[r]);
        off[r] = tot;
        tot += (uint32_t)__popcll(m[r]);
    }
    if (tot) {
        uint32_t lb = 0, lg = 0, hb = 0, hg = 0;
        if ((threadIdx.x & 63) == 0) {
            if (lds_ok) ring_reserve(ltail, lhead, P.lcap, tot, lb, lg);
            if (lg < tot) ring_reserve(htail, hhead, P.ncap, tot - lg, hb, hg);
        }
        lb = __builtin_amdgcn_readfirstlane(lb);  // lane 0's values (every lane is active)
        lg = __builtin_amdgcn_readfirstlane(lg);
        hb = __builtin_amdgcn_readfirstlane(hb);
        hg = __builtin_amdgcn_readfirstlane(hg);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!nr[r]) continue;
            const uint32_t q = off[r] + lane_rank(m[r]);
            if (q < lg) P.lring[(lb + q) & (P.lcap - 1)] = key[r];
            else if (q < lg + hg) P.near[(hb + q - lg) & (P.ncap - 1)] = key[r];
            else fr[r] = true;  // both rings full
        }
    }
    uint32_t pos[R];
    wave_append<R>(fr, fcount, pos);
    bool dropped = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!fr[r]) continue;
        if (pos[r] < P.fcap) fdst[pos[r]] = key[r];
        else dropped = true;
    }
    return dropped;
}

// A resumed search (DA*): the open set arrives as the first n_in entries of
// the far pile, the target's current f is F0, and entries that cannot beat
// F(t) now are kept in a dormant pile instead of being dropped -- a later
// batch may raise F(t) again.  The dormant pile may already hold dorm0
// entries (kept from before); new ones go behind them.  (f, parent) words
// persist between searches, so nothing is logged for a reset.
struct PileResume {
    uint64_t *dorm;  // dcap entries
    uint32_t dcap;
    uint32_t n_in;
    uint32_t F0;
    uint32_t dorm0;
};

// Wave-collective: appends the selected keys to the dormant pile.
template <int R, class SL>
__device__ __forceinline__ void dorm_append(const PileResume &Z, SL &S, const bool (&sel)[R], const uint64_t (&key)[R])
{
    uint32_t pos[R];
    wave_append<R>(sel, &S.dorm_n, pos);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!sel[r]) continue;
        if (pos[r] < Z.dcap) Z.dorm[pos[r]] = key[r];
        else atomicOr(&S.err, (uint32_t)QERR_HEAP);
    }
}

template <int B, int V, bool RESUME = false>
struct PilePush {
    const PilePtrs *P;
    uint64_t *far;
    SlotLds<B, V> *S;
    int32_t *touched;
    const PileResume *Z = nullptr;
    uint32_t err = 0;

    __device__ void target(uint32_t f) { atomicMin(&S->F, f); }
    __device__ void touch(int32_t c, uint64_t prev)
    {
        if constexpr (!RESUME)
            if ((uint32_t)(prev >> 32) == DDP_INF_F) touched[atomicAdd(&S->ntouch, 1)] = c;
    }
    // near candidates: the LDS ring, then the HBM ring, then the far pile
    __device__ void operator()(bool has, uint64_t key)
    {
        const bool h[1] = {has};
        const uint64_t k[1] = {key};
        all<1>(h, k);
    }
    template <int R>
    __device__ void touch_all(const bool (&first)[R], const int32_t (&c)[R])
    {
        if constexpr (RESUME) return;
        uint32_t pos[R];
        wave_append<R>(first, (uint32_t *)&S->ntouch, pos);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (first[r]) touched[pos[r]] = c[r];
    }
    // R candidates per lane; see pile_place.  The two rings act as one FIFO:
    // while the HBM ring holds entries (pushed when the LDS ring was full),
    // new ones queue behind them there, since rounds drain the LDS ring
    // first.  (Letting later pushes overtake through the LDS ring made a
    // Dijkstra-order search on R-MAT expand 65% more nodes.)
    template <int R>
    __device__ void all(const bool (&has)[R], const uint64_t (&key)[R])
    {
        const uint32_t F = __hip_atomic_load(&S->F, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        bool put[R];
#pragma unroll
        for (int r = 0; r < R; ++r) put[r] = has[r] && (uint32_t)(key[r] >> 32) < F;
        if constexpr (RESUME) {
            bool dz[R];
#pragma unroll
            for (int r = 0; r < R; ++r) dz[r] = has[r] && !put[r];
            dorm_append<R>(*Z, *S, dz, key);
        }
        const bool lds_ok = P->lcap && S->near_tail == S->near_head;
        if (pile_place<R>(*P, S->thr, lds_ok, S->lring_head, S->near_head, &S->lring_tail, &S->near_tail, far,
                          &S->far_n, put, key))
            S->ovf = 1;  // dropped: the next refill rebuilds the far pile
    }
};

__device__ __forceinline__ uint32_t sat_add_u32(uint32_t a, uint32_t b)
{
    return a > 0xFFFFFFFFu - b ? 0xFFFFFFFFu : a + b;
}

// Entries each lane takes from the near ring per round.  One: two per lane
// (2,048 per round at 1,024 lanes) measured slower on the C4 batch (139k
// against 148k queries/s) -- wider rounds expand more nodes out of f order
// (80M against 63M) and do not shorten the longest searches, whose rounds are
// bound by the memory requests one CU keeps in flight, not by per-round costs.
template <int B>
constexpr int pile_entries_per_lane()
{
    return 1;
}

// Same contract as slot_search (S.F, S.err, gp and touched describe the search
// on return).  delta0 is the first window width.  W entries per lane per round.
//
// The near ring has two levels: P.lring in LDS (P.lcap entries) and P.near in
// HBM for what does not fit, kept in FIFO order together (see PilePush).
// Rounds take from the LDS ring while it holds entries; an ordinary round then
// ends with a plain barrier, since everything
// the next round reads is in LDS or was written by atomics that completed.
// Reading HBM-resident entries (the HBM ring, the far pile, the touched log)
// is always preceded by a barrier at which every wave has drained its stores.
// The ring heads are uniform registers, mirrored in LDS for the pushes'
// capacity checks.
// A near ring holding fewer than V / kThinRing entries is refilled early
// (when the next window adjoins the current one).  Swept on the C4 batch:
// 32 and 64 were fastest (64 expands 3% more nodes than no early refills,
// 32 19%); 16 and 4 ran the window ahead of f order; 128 helped less and
// 256 (three or fewer entries) refilled so often that it was slower.
constexpr uint32_t kThinRing = 64;

//
// RESUME (a DA* search resumed after a batch): the search starts from the
// open set of PileResume Z -- its first Z.n_in entries already in P.far0 --
// with F(t) = Z.F0 instead of from s, and an entry that cannot beat F(t) goes
// to Z's dormant pile instead of being dropped (see PileResume).  On return
// the dormant pile holds every open entry left: the search's open set for the
// next resume.  A far-pile rebuild re-queues every node with a finite word.
template <int B, int HK, int W = pile_entries_per_lane<B>(), bool RESUME = false, bool DIRECT_OK = false>
__device__ void slot_search_piles(const GraphView &G, HeurView hv, uint64_t *gp, int32_t *touched,
                                  const PilePtrs &P, int32_t N, int32_t s, int32_t t, int32_t maxit,
                                  uint32_t delta0, SlotLds<B, B * W> &S, ExpandCounters &C, uint32_t &stale,
                                  const PileResume *Z = nullptr)
{
    constexpr uint32_t V = (uint32_t)B * W;
    const int lane = threadIdx.x;
    const uint32_t nmask = P.ncap - 1, lmask = P.lcap ? P.lcap - 1 : 0;
    if (lane == 0) {
        S.F = DDP_INF_F;
        S.dorm_n = 0;
        S.ntouch = 0;
        S.err = 0;
        S.iters = 0;
        S.lring_head = S.lring_tail = 0;
        S.near_head = S.near_tail = 0;
        S.far_n = S.far_n2 = 0;
        S.delta = delta0 > 0 ? delta0 : 1;
        S.thr = 0;
        S.phase_n = 0;
        S.ovf = 0;
        S.no_early = 0;
    }
    __syncthreads();
    bool run = s >= 0 && s < N && t >= 0 && t < N;
    if constexpr (RESUME) {
        run = t >= 0 && t < N && Z->n_in > 0;
        set_target<HK>(hv, t);
        if (lane == 0) {
            S.F = Z->F0;
            S.far_n = Z->n_in;  // the open set; the first refill sets the window
            S.dorm_n = Z->dorm0;
        }
    } else if (run) {
        set_target<HK>(hv, t);
        const int32_t hs = heur_eval<HK>(hv, s);
        if (HK == DDP_H_ARRAY && hs < 0) {
            run = false;  // h(s) = -1: s cannot reach t
        } else if (s == t) {
            run = false;
            if (lane == 0) S.F = (uint32_t)hs;
        } else if (lane == 0) {
            const uint64_t key = ((uint64_t)(uint32_t)hs << 32) | (uint32_t)s;
            gp[s] = ((uint64_t)(uint32_t)hs << 32) | 0xFFFFFFFFull;
            touched[0] = s;
            S.ntouch = 1;
            if (P.lcap) {
                P.lring[0] = key;
                S.lring_tail = 1;
            } else {
                P.near[0] = key;
                S.near_tail = 1;
            }
            S.thr = sat_add_u32((uint32_t)hs, S.delta);
        }
    } else if (lane == 0) {
        S.err = QERR_BADQ;
    }
    block_sync_global();
    uint64_t *far = P.far0, *far2 = P.far1;
    uint32_t lh = 0, hh = 0;  // ring heads (uniform)
    int32_t it = 0;
    while (run) {
        const uint32_t lt = S.lring_tail, F = S.F;
        bool from_lds = lt != lh;
        uint32_t avail = lt - lh;
        if (!from_lds) {
            block_sync_global();  // HBM entries ahead: every wave's stores have landed
            avail = S.near_tail - hh;
        }
        // Early refill: a near ring this thin holds stragglers of the window
        // while the far pile waits; move the next window in beside them.
        // (Not while a rebuild is pending: it re-queues every touched node
        // under F(t), and with the near ring drained F(t) prunes far more.)
        const bool early = avail > 0 && (lt - lh) + (S.near_tail - hh) < V / kThinRing && S.far_n > 0 && !S.ovf &&
                           !S.no_early;
        if (early && from_lds) block_sync_global();  // the far pile is read below
        if (avail == 0 || early) {
            if (S.ovf) {
                // ---- rebuild the far pile from the touched log (entries were dropped);
                // a resumed search, which logs nothing, scans every node
                const int32_t nt = RESUME ? N : S.ntouch;
                if (lane == 0) S.far_n2 = 0;
                __syncthreads();
                for (int32_t i0 = 0; i0 < nt; i0 += B) {  // uniform trip count: the reservations are wave-wide
                    const int32_t i = i0 + lane;
                    int32_t v = -1;
                    uint32_t f = 0xFFFFFFFFu;
                    if (i < nt) {
                        v = RESUME ? i : __hip_atomic_load(touched + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        f = (uint32_t)(ld_gp(gp + v) >> 32);
                    }
                    if constexpr (RESUME) {  // finite but hopeless now: dormant
                        const bool dz[1] = {f != DDP_INF_F && f >= F && v != t};
                        const uint64_t dk[1] = {((uint64_t)f << 32) | (uint32_t)v};
                        dorm_append<1>(*Z, S, dz, dk);
                    }
                    const uint32_t idx = far_reserve(f < F, &S.far_n2);  // f < F also excludes t
                    if (idx == 0xFFFFFFFFu) continue;
                    if (idx < P.fcap) far2[idx] = ((uint64_t)f << 32) | (uint32_t)v;
                    else atomicOr(&S.err, (uint32_t)QERR_HEAP);
                }
                block_sync_global();
                if (lane == 0) {
                    S.far_n = S.far_n2;
                    S.ovf = 0;
                    C.rebuilds++;
                }
                uint64_t *tmp = far;
                far = far2;
                far2 = tmp;
                __syncthreads();
                if (S.err) break;
            }
            // ---- refill the near rings from the far pile
            const uint32_t nf = S.far_n < P.fcap ? S.far_n : P.fcap;
            if (lane == 0) C.refills++;
            C.scanned += 2 * ((nf > (uint32_t)lane) ? (nf - 1 - lane) / B + 1 : 0);
            // The far pile is read in chunks of KC*B entries, each chunk's
            // loads issued together (entries, then their (f, parent) words),
            // once for the minimum and again for the split.  (Keeping the
            // first chunk in registers between the passes pushed the
            // 1,024-lane kernel into spills.)
            constexpr int KC = 2;
            auto load_chunk = [&](uint32_t base, uint64_t (&k)[KC], bool (&live)[KC]) {
                uint64_t wv[KC];
#pragma unroll
                for (int j = 0; j < KC; ++j)
                    if (base + (uint32_t)j * B + lane < nf) k[j] = far[base + (uint32_t)j * B + lane];
#pragma unroll
                for (int j = 0; j < KC; ++j)
                    if (base + (uint32_t)j * B + lane < nf && (uint32_t)(k[j] >> 32) < F)
                        wv[j] = ld_gp(gp + (uint32_t)k[j]);
#pragma unroll
                for (int j = 0; j < KC; ++j) {
                    live[j] = false;
                    if (base + (uint32_t)j * B + lane < nf) {
                        const uint32_t f = (uint32_t)(k[j] >> 32);
                        live[j] = f < F && (uint32_t)(wv[j] >> 32) == f;  // else hopeless or stale
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
            if (fmin >= F) {  // nothing in the far pile can improve t (also: far pile empty)
                if constexpr (RESUME) {  // kept for later batches
                    for (uint32_t i0 = 0; i0 < nf; i0 += B) {
                        const uint32_t i = i0 + lane;
                        uint64_t k[1] = {i < nf ? far[i] : 0};
                        const bool dz[1] = {i < nf && (uint32_t)(k[0] >> 32) >= F && (int32_t)(uint32_t)k[0] != t};
                        dorm_append<1>(*Z, S, dz, k);
                    }
                }
                if (!early) break;
                if (lane == 0) S.far_n = 0;  // F only falls: those entries are hopeless for good
                __syncthreads();
                continue;
            }
            // An early refill only takes the next window when it adjoins the
            // current one; entries further out would run ahead of f order (a
            // maze's detours).  Otherwise wait for the ring to drain.
            if (early && fmin >= sat_add_u32(S.thr, S.delta)) {
                if (lane == 0) S.no_early = 1;
                __syncthreads();
                continue;
            }
            if (lane == 0) {
                S.no_early = 0;
                if (!early) {
                    if (S.phase_n < V) S.delta = sat_add_u32(S.delta, S.delta);
                    else if (S.phase_n > 8 * V && S.delta > 1) S.delta >>= 1;
                    S.phase_n = 0;
                }
                const uint32_t nt = sat_add_u32(fmin, S.delta);
                S.thr = nt > S.thr ? nt : S.thr;
                S.far_n2 = 0;
            }
            __syncthreads();
            const uint32_t thr = S.thr;
            // FIFO with what the rings still hold (early refills): the LDS ring
            // takes entries only while the HBM ring is empty
            const bool lds_ok = P.lcap != 0 && S.near_tail == hh;
            // second pass: near entries to the rings (both empty here), the rest
            // compacted into far2 (fewer than nf <= fcap entries: nothing drops)
            for (uint32_t base = 0; base < nf; base += (uint32_t)KC * B) {  // uniform trips
                uint64_t k2[KC];
                bool l2[KC];
                load_chunk(base, k2, l2);
                (void)pile_place<KC>(P, thr, lds_ok, lh, hh, &S.lring_tail, &S.near_tail, far2, &S.far_n2, l2,
                                     k2);
                if constexpr (RESUME) {  // hopeless now, kept for later batches
                    bool dz[KC];
#pragma unroll
                    for (int j = 0; j < KC; ++j)
                        dz[j] = base + (uint32_t)j * B + lane < nf && (uint32_t)(k2[j] >> 32) >= F &&
                                (int32_t)(uint32_t)k2[j] != t;
                    dorm_append<KC>(*Z, S, dz, k2);
                }
            }
            block_sync_global();
            if (lane == 0) S.far_n = S.far_n2;
            uint64_t *tmp = far;
            far = far2;
            far2 = tmp;
            __syncthreads();
            if (S.err) break;
            continue;
        }
        if (it >= maxit) {
            if (lane == 0) atomicOr(&S.err, (uint32_t)QERR_ITERS);
            break;
        }
        ++it;
        // ---- take up to V entries from one ring, W per lane (coalesced: entry w*B + lane)
        const uint32_t n = avail < V ? avail : V;
        int32_t u[W];
        RowSpan rs[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            u[w] = -1;
            const uint32_t idx = (uint32_t)(w * B + lane);
            if (idx < n) {
                const uint64_t key = from_lds ? P.lring[(lh + idx) & lmask] : P.near[(hh + idx) & nmask];
                const uint32_t f = (uint32_t)(key >> 32);
                const int32_t node = (int32_t)(uint32_t)key;
                // the node's word, row and heuristic input in one round trip:
                // the row is loaded before the entry is known to be current
                const uint64_t word = ld_gp(gp + node);
                const int32_t rb = G.off[node], re = G.off[node + 1];
                const int32_t hn = heur_eval<HK>(hv, node);
                if (f < F && (uint32_t)(word >> 32) == f) {
                    u[w] = node;
                    rs[w].b = rb;
                    rs[w].deg = re - rb;
                    rs[w].gu = f - (uint32_t)hn;
                } else {
                    ++stale;
                }
            }
        }
        // Pushes of this round start after expand_rows' first barrier, when
        // every lane has read its entry, so the freed slots may be reused.
        if (from_lds) lh += n;
        else hh += n;
        if (lane == 0) {
            S.lring_head = lh;
            S.near_head = hh;
            S.phase_n += n;
        }
        PilePush<B, B * W, RESUME> push{&P, far, &S, touched, Z};
        // a resumed search keeps every relaxation: a later batch may raise F(t)
        expand_rows<B, HK, W, PilePush<B, B * W, RESUME>, false, DIRECT_OK>(G, hv, gp, t, u, rs, S.L, C, push,
                                                                  RESUME ? 0xFFFFFFFFu : F);
        if (push.err) atomicOr(&S.err, push.err);
        __syncthreads();  // the round's pushes and errors are visible to the next round
        if (S.err) break;
    }
    if (lane == 0) S.iters = it;
    block_sync_global();
}

// Lane 0 writes the path s..t into out[0..cap) and returns its length in
// S.plen (0 if unreachable, -len if longer than cap).
template <int B, int V>
__device__ void slot_path(const uint64_t *gp, int32_t N, int32_t s, int32_t t, int32_t *out, int32_t cap,
                          SlotLds<B, V> &S)
{
    if (threadIdx.x == 0) {
        int32_t n = 0;
        if (S.F != DDP_INF_F) {
            int32_t v = t;
            for (int32_t step = 0; step <= N; ++step) {
                if (n < cap) out[n] = v;
                ++n;
                if (v == s) break;
                const uint32_t p = (uint32_t)ld_gp(gp + v);
                if (p >= (uint32_t)N) {
                    n = 0;
                    atomicOr(&S.err, (uint32_t)QERR_PATH);
                    break;
                }
                v = (int32_t)p;
            }
            if (n > 0 && n <= cap)
                for (int32_t i = 0; i < n / 2; ++i) {  // t..s -> s..t (this lane wrote every entry)
                    const int32_t x = out[i];
                    out[i] = out[n - 1 - i];
                    out[n - 1 - i] = x;
                }
        }
        S.plen = (n > cap) ? -n : n;
    }
}

// Clears every (f, parent) word the query touched.  Starts with a barrier so
// lane 0's path walk has finished reading parents; ends with one so the next
// query's atomics see the cleared words.
template <int B, int V>
__device__ void slot_reset(uint64_t *gp, const int32_t *touched, SlotLds<B, V> &S)
{
    __syncthreads();
    const int32_t nt = S.ntouch;
    for (int32_t i = (int32_t)threadIdx.x; i < nt; i += B)
        gp[__hip_atomic_load(touched + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)] = DDP_GP_INF;
    block_sync_global();
}

template <int B>
__device__ void slot_flush_stats(unsigned long long *stats, ExpandCounters &C, uint32_t &stale)
{
    const uint64_t v0 = wave_sum_u64(C.expanded), v1 = wave_sum_u64(C.edges), v2 = wave_sum_u64(C.improved),
                   v3 = wave_sum_u64(stale), v4 = wave_sum_u64(C.rebuilds), v5 = wave_sum_u64(C.refills),
                   v6 = wave_sum_u64(C.scanned);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(stats + 0, (unsigned long long)v0);
        atomicAdd(stats + 1, (unsigned long long)v1);
        atomicAdd(stats + 2, (unsigned long long)v2);
        atomicAdd(stats + 3, (unsigned long long)v3);
        if (v4) atomicAdd(stats + 4, (unsigned long long)v4);
        if (v5) atomicAdd(stats + 5, (unsigned long long)v5);
        if (v6) atomicAdd(stats + 6, (unsigned long long)v6);
    }
    C = ExpandCounters{};
    stale = 0;
}
