// Single-query GA*: K priority queues over one persistent grid (grid mode),
// or all of them inside one workgroup (workgroup mode, further down).
//
// Reference loop (GA_star/GA_v2.cu:488-559): per iteration the host launches
// extractMin, A_star_expand, keepHeapPQ, setNV (an O(N) flag scan), insertPQ,
// copies flags and all K queue sizes back, and runs checkMIN.  Here the whole
// loop is one kernel:
//   * queue q is a binary heap owned by one lane (lane q % B of workgroup
//     q / B); its first lcap entries sit in LDS, the rest spill to HBM;
//   * each iteration every queue pops its best valid entry, the workgroup
//     expands the popped nodes cooperatively (search_expand.h), and every
//     improvement CASes a packed (f << 32 | parent) word -- no per-node locks;
//   * improved nodes are appended to a push buffer with one atomic per wave
//     (ballot + mbcnt), replacing setNV's O(N) scan; after the grid barrier
//     queue q takes entries p with (rr + p) % K == q, the round-robin
//     distribution of insertPQ (GA_v2.cu:241-273) made continuous across
//     iterations so queue sizes stay balanced;
//   * termination is decided on the device: once the target's f is <= the
//     smallest queue top at the previous iteration's start and every f pushed
//     during it, no open node can improve it (GA_v2.cu:277-289, 547-557).
//     The decision reads only values frozen by the barrier, so all workgroups
//     agree without another barrier.
// Stale heap entries (a node improved after it was queued) are skipped when
// popped; this replaces keepHeapPQ's in-place decrease-key repair.
#include <algorithm>
#include <chrono>
#include <cstring>

#include <type_traits>

#include "bidi_search.h"
#include "gastar.h"

namespace {

enum : uint32_t {
    ERR_HEAP = GA_ERR_HEAP, ERR_PBUF = GA_ERR_PBUF, ERR_ITERS = GA_ERR_ITERS, ERR_BARRIER = GA_ERR_BARRIER,
    ERR_PATH = 16
};

struct GridArgs {
    GraphView g;
    HeurView h;
    uint64_t *gp;
    uint64_t *home;      // heap spill / home: entry i of queue q at home[i*K + q]
    int32_t *hsize;      // K
    uint64_t *pbuf;      // 2 * pcap
    uint32_t *ctrl;      // GA_CTRL_WORDS
    unsigned long long *stats;  // expanded, edges, improved, stale
    int32_t K, lcap, cap, pcap, t, maxit;
    uint32_t F0;
    int32_t rr0;
    uint64_t barrier_ticks;
};

template <int B>
struct GridPush {
    uint64_t *buf;
    uint32_t *pushcnt;
    int32_t pcap;
    uint32_t tmin = 0xFFFFFFFFu, pushmin = 0xFFFFFFFFu;
    uint32_t err = 0;

    __device__ void target(uint32_t f) { tmin = f < tmin ? f : tmin; }
    __device__ void touch(int32_t, uint64_t) {}
    __device__ void operator()(bool has, uint64_t key)
    {
        const uint64_t mask = __ballot(has);
        if (mask == 0) return;
        const int lane = threadIdx.x & 63;
        const int leader = __builtin_ctzll(mask);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(pushcnt, (uint32_t)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (has) {
            const uint32_t idx = base + lane_rank(mask);
            if (idx < (uint32_t)pcap) buf[idx] = key;
            else err |= ERR_PBUF;
            const uint32_t f = (uint32_t)(key >> 32);
            pushmin = f < pushmin ? f : pushmin;
        }
    }
    template <int R>
    __device__ void touch_all(const bool (&)[R], const int32_t (&)[R]) {}
    template <int R>
    __device__ void all(const bool (&has)[R], const uint64_t (&key)[R])
    {
        uint32_t pos[R];
        wave_append<R>(has, pushcnt, pos);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!has[r]) continue;
            if (pos[r] < (uint32_t)pcap) buf[pos[r]] = key[r];
            else err |= ERR_PBUF;
            const uint32_t f = (uint32_t)(key[r] >> 32);
            pushmin = f < pushmin ? f : pushmin;
        }
    }
};

template <int B, int HK>
__global__ void __launch_bounds__(B) ga_grid_kernel(GridArgs a)
{
    extern __shared__ uint64_t s_heap[];  // lcap * B
    __shared__ ExpandLds<B> L;
    __shared__ uint32_t s_red[16];
    __shared__ int32_t s_flag;

    const int lane = threadIdx.x;
    const int32_t K = a.K;
    const int32_t q = blockIdx.x * B + lane;
    uint32_t *ctrl = a.ctrl;
    const LaneHeap H{s_heap + lane, a.home + q, B, a.lcap, (int64_t)K, a.cap};
    int32_t size = a.hsize[q];
    for (int32_t i = 0; i < size && i < a.lcap; ++i) s_heap[i * B + lane] = a.home[(int64_t)i * K + q];

    uint32_t gen = ld_u32(ctrl + GA_BAR + 1);
    uint32_t F = a.F0;
    int32_t rr = a.rr0;
    uint32_t err_local = 0, err_seen = 0;
    ExpandCounters C;
    uint32_t stale = 0;
    int32_t it = 0;
    for (;; ++it) {
        // -- drain the previous iteration's pushes into this lane's queue
        const int ps = (it + 2) % 3;
        const uint32_t n = ld_u32(ctrl + ps * GA_SLOT_WORDS + GA_S_PUSHCNT);
        const uint32_t topm = ld_u32(ctrl + ps * GA_SLOT_WORDS + GA_S_TOPMIN);
        const uint32_t pm = ld_u32(ctrl + ps * GA_SLOT_WORDS + GA_S_PUSHMIN);
        const uint32_t tm = ld_u32(ctrl + ps * GA_SLOT_WORDS + GA_S_TMIN);
        err_seen |= ld_u32(ctrl + ps * GA_SLOT_WORDS + GA_S_ERR);
        F = tm < F ? tm : F;
        const uint32_t navail = n < (uint32_t)a.pcap ? n : (uint32_t)a.pcap;
        const uint64_t *buf = a.pbuf + (int64_t)((it + 1) & 1) * a.pcap;
        for (int64_t p = ((int64_t)q - rr) % K + (q < rr ? K : 0); p < (int64_t)navail; p += K) {
            const uint64_t e = __hip_atomic_load(buf + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!heap_push(H, size, e)) err_local |= ERR_HEAP;
        }
        rr = (int32_t)(((int64_t)rr + n) % K);
        // -- termination (uniform: every input was frozen by the last barrier)
        const uint32_t open_min = topm < pm ? topm : pm;
        if (it > 0 && F <= open_min) break;
        if (err_seen) break;
        if (it >= a.maxit) {
            err_local |= ERR_ITERS;
            break;
        }
        const int cs = it % 3;
        if (blockIdx.x == 0 && lane == 0) {
            uint32_t *ns = ctrl + ((it + 1) % 3) * GA_SLOT_WORDS;
            st_u32(ns + GA_S_PUSHCNT, 0);
            st_u32(ns + GA_S_TOPMIN, 0xFFFFFFFFu);
            st_u32(ns + GA_S_PUSHMIN, 0xFFFFFFFFu);
            st_u32(ns + GA_S_TMIN, 0xFFFFFFFFu);
            st_u32(ns + GA_S_ERR, 0);
        }
        // -- publish the smallest queue top (stale tops only make this smaller)
        const uint32_t mytop = size > 0 ? (uint32_t)(s_heap[lane] >> 32) : 0xFFFFFFFFu;
        const uint32_t bmin = block_min_u32<B>(mytop, s_red);
        if (lane == 0 && bmin != 0xFFFFFFFFu) atomicMin(ctrl + cs * GA_SLOT_WORDS + GA_S_TOPMIN, bmin);
        // -- pop the best entry whose f is still current
        int32_t u = -1;
        uint32_t fu = 0;
        while (size > 0) {
            uint64_t key;
            heap_pop(H, size, key);
            const int32_t node = (int32_t)(uint32_t)key;
            const uint32_t f = (uint32_t)(key >> 32);
            if ((uint32_t)(ld_gp(a.gp + node) >> 32) == f) {
                u = node;
                fu = f;
                break;
            }
            ++stale;
        }
        // -- expand and push
        GridPush<B> push{a.pbuf + (int64_t)(it & 1) * a.pcap, ctrl + cs * GA_SLOT_WORDS + GA_S_PUSHCNT, a.pcap};
        expand_block<B, HK>(a.g, a.h, a.gp, a.t, u, fu, L, C, push);
        err_local |= push.err;
        const uint32_t wt = wave_min_u32(push.tmin);
        const uint32_t wp = wave_min_u32(push.pushmin);
        if ((lane & 63) == 0) {
            if (wt != 0xFFFFFFFFu) atomicMin(ctrl + cs * GA_SLOT_WORDS + GA_S_TMIN, wt);
            if (wp != 0xFFFFFFFFu) atomicMin(ctrl + cs * GA_SLOT_WORDS + GA_S_PUSHMIN, wp);
        }
        if (err_local) atomicOr(ctrl + cs * GA_SLOT_WORDS + GA_S_ERR, err_local);
        if (!grid_barrier(ctrl + GA_BAR, gridDim.x, gen, a.barrier_ticks, &s_flag)) {
            atomicOr(ctrl + GA_RES_ERR, (uint32_t)ERR_BARRIER);
            break;
        }
    }
    // -- persist queues for a later resume (DA*) and report
    for (int32_t i = 0; i < size && i < a.lcap; ++i) a.home[(int64_t)i * K + q] = s_heap[i * B + lane];
    a.hsize[q] = size;
    if (err_local) atomicOr(ctrl + GA_RES_ERR, err_local);
    if (err_seen) atomicOr(ctrl + GA_RES_ERR, err_seen);
    if (blockIdx.x == 0 && lane == 0) {
        ctrl[GA_RES_IT] = (uint32_t)it;
        ctrl[GA_RES_RR] = (uint32_t)rr;
        ctrl[GA_RES_F] = F;
    }
    uint64_t v0 = wave_sum_u64(C.expanded), v1 = wave_sum_u64(C.edges), v2 = wave_sum_u64(C.improved),
             v3 = wave_sum_u64(stale);
    if ((lane & 63) == 0) {
        atomicAdd(a.stats + 0, (unsigned long long)v0);
        atomicAdd(a.stats + 1, (unsigned long long)v1);
        atomicAdd(a.stats + 2, (unsigned long long)v2);
        atomicAdd(a.stats + 3, (unsigned long long)v3);
    }
}

__global__ void fill_u64_kernel(uint64_t *p, int64_t n, uint64_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Fresh search: gp[s] = (h(s), parent -1); s is the only push of "iteration -1".
__global__ void seed_kernel(uint64_t *gp, uint64_t *pbuf1, uint32_t *ctrl, int32_t s, uint32_t hs)
{
    if (threadIdx.x != 0) return;
    gp[s] = ((uint64_t)hs << 32) | 0xFFFFFFFFull;
    pbuf1[0] = ((uint64_t)hs << 32) | (uint32_t)s;
    for (int k = 0; k < 3; ++k) {
        uint32_t *sl = ctrl + k * GA_SLOT_WORDS;
        sl[GA_S_PUSHCNT] = 0;
        sl[GA_S_TOPMIN] = 0xFFFFFFFFu;
        sl[GA_S_PUSHMIN] = 0xFFFFFFFFu;
        sl[GA_S_TMIN] = 0xFFFFFFFFu;
        sl[GA_S_ERR] = 0;
    }
    ctrl[2 * GA_SLOT_WORDS + GA_S_PUSHCNT] = 1;
    ctrl[2 * GA_SLOT_WORDS + GA_S_PUSHMIN] = hs;
    ctrl[GA_RES_IT] = 0;
    ctrl[GA_RES_RR] = 0;
    ctrl[GA_RES_F] = DDP_INF_F;
    ctrl[GA_RES_ERR] = 0;
    ctrl[GA_BAR] = 0;
    ctrl[GA_BAR + 1] = 0;
}

__global__ void unpack_parent_kernel(const uint64_t *gp, int32_t *par, int32_t N)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        par[i] = (int32_t)(uint32_t)gp[i];
}

using KernelFn = void (*)(GridArgs);

template <int B>
KernelFn pick_kernel(int hk)
{
    switch (hk) {
    case DDP_H_ZERO: return ga_grid_kernel<B, DDP_H_ZERO>;
    case DDP_H_ARRAY: return ga_grid_kernel<B, DDP_H_ARRAY>;
    case DDP_H_MANHATTAN: return ga_grid_kernel<B, DDP_H_MANHATTAN>;
    default: return ga_grid_kernel<B, DDP_H_EUCLID>;
    }
}

KernelFn pick(int B, int hk)
{
    switch (B) {
    case 64: return pick_kernel<64>(hk);
    case 256: return pick_kernel<256>(hk);
    default: return pick_kernel<1024>(hk);
    }
}

// Workgroup mode: all K <= 1024 queues live in one workgroup, one per lane,
// and an improved node goes to the queue of the lane that relaxed it.  An
// iteration then synchronises through LDS barriers only, where the grid
// mode pays a device-wide barrier and its atomics; searches whose frontier
// holds a few thousand nodes (or whose iteration count is set by a long
// path) are bound by that per-iteration latency, not by bandwidth.
struct WgArgs {
    GraphView g;
    GraphView gr;        // transposed rows (bidirectional mode)
    HeurView h;
    uint64_t *gp;        // N (bidirectional: 2N, forward then backward)
    uint64_t *home;      // heap spill (entry i of queue q at home[i*B + q]), or the piles (two sets: bidirectional)
    int32_t *touched;    // N: nodes whose word left INF (bidirectional: 2N)
    int32_t *path;       // bidirectional mode: the path s..t is walked on the device into path[0..path_cap)
    int32_t path_cap;
    uint32_t *ctrl;
    unsigned long long *stats;
    int32_t N, s, t, lcap, cap, maxit;
    uint32_t ncap, fcap, delta0;  // pile mode: near ring, far buffers, first window
    uint32_t lring;               // pile mode: entries of the LDS near ring (dynamic LDS)
};

enum { WM_HEAPS = 0, WM_PILES = 1, WM_BIDI = 2 };

// WM_HEAPS: lane heaps (slot_search); WM_PILES: near/far piles
// (slot_search_piles); WM_BIDI: piles from both ends (slot_search_bidi)
template <int B, int HK, int WM>
__global__ void __launch_bounds__(B) ga_wg_kernel(WgArgs a)
{
    extern __shared__ uint64_t s_heap[];  // lcap * B (heaps), near rings (piles)
    constexpr int SV = (WM == WM_PILES) ? B * pile_entries_per_lane<B>() : B;  // expansion width
    __shared__ std::conditional_t<WM == WM_BIDI, BidiLds<B>, SlotLds<B, SV>> S;
    const int lane = threadIdx.x;
    const LaneHeap H{s_heap + lane, a.home + lane, B, a.lcap, (int64_t)B, a.cap};
    ExpandCounters C;
    uint32_t stale = 0;
    if constexpr (WM == WM_BIDI) {
        const int32_t N = a.N;
        uint64_t *bh = a.home + (int64_t)a.cap * B;
        const PilePtrs PF{a.home, a.home + a.ncap, a.home + a.ncap + a.fcap, a.ncap, a.fcap, s_heap, a.lring / 2};
        const PilePtrs PB{bh, bh + a.ncap, bh + a.ncap + a.fcap, a.ncap, a.fcap, s_heap + a.lring / 2, a.lring / 2};
        const BidiSide D{{a.g.off, a.gr.off}, {a.g.col, a.gr.col}, {a.g.w, a.gr.w}, {a.gp, a.gp + N},
                         {a.touched, a.touched + N}, {PF, PB}};
        slot_search_bidi<B, HK>(D, a.h, a.g.relax_min != 0, N, a.s, a.t, a.maxit, a.delta0, S, C, stale);
        if (a.path) slot_path_bidi(a.gp, a.gp + N, N, a.s, a.t, a.path, a.path_cap, S);
        if (lane == 0) {
            const uint32_t e = S.err;
            a.ctrl[GA_RES_IT] = (uint32_t)S.iters;
            a.ctrl[GA_RES_F] = (uint32_t)(S.meet >> 32);
            a.ctrl[GA_RES_ERR] = ((e & QERR_HEAP) ? ERR_HEAP : 0u) | ((e & QERR_ITERS) ? ERR_ITERS : 0u) |
                                 ((e & QERR_PATH) ? ERR_PATH : 0u);
            a.ctrl[GA_RES_PLEN] = (uint32_t)(a.path ? S.plen : 0);
        }
    } else {
        if constexpr (WM == WM_PILES) {
            const PilePtrs P{a.home, a.home + a.ncap, a.home + a.ncap + a.fcap, a.ncap, a.fcap, s_heap, a.lring};
            slot_search_piles<B, HK, pile_entries_per_lane<B>(), false, true>(a.g, a.h, a.gp, a.touched, P, a.N, a.s,
                                                                              a.t, a.maxit, a.delta0, S, C, stale);
        } else {
            slot_search<B, HK>(a.g, a.h, a.gp, a.touched, H, s_heap, a.N, a.s, a.t, a.maxit, S, C, stale);
        }
        if (lane == 0) {
            const uint32_t e = S.err;
            a.ctrl[GA_RES_IT] = (uint32_t)S.iters;
            a.ctrl[GA_RES_F] = S.F;
            a.ctrl[GA_RES_ERR] = ((e & QERR_HEAP) ? ERR_HEAP : 0u) | ((e & QERR_ITERS) ? ERR_ITERS : 0u);
        }
    }
    slot_flush_stats<B>(a.stats, C, stale);
}

using WgFn = void (*)(WgArgs);

template <int B, int WM>
WgFn pick_wg_b(int hk)
{
    switch (hk) {
    case DDP_H_ZERO: return ga_wg_kernel<B, DDP_H_ZERO, WM>;
    case DDP_H_ARRAY:
        if constexpr (WM == WM_BIDI) return nullptr;  // the potential needs h toward s as well as t
        else return ga_wg_kernel<B, DDP_H_ARRAY, WM>;
    case DDP_H_MANHATTAN: return ga_wg_kernel<B, DDP_H_MANHATTAN, WM>;
    default: return ga_wg_kernel<B, DDP_H_EUCLID, WM>;
    }
}

template <int WM>
WgFn pick_wg_p(int B, int hk)
{
    if constexpr (WM == WM_BIDI) {
        switch (B) {
        case 256: return pick_wg_b<256, WM>(hk);
        case 512: return pick_wg_b<512, WM>(hk);
        case 1024: return pick_wg_b<1024, WM>(hk);
        default: return nullptr;
        }
    }
    switch (B) {
    case 64: return pick_wg_b<64, WM>(hk);
    case 128: return pick_wg_b<128, WM>(hk);
    case 256: return pick_wg_b<256, WM>(hk);
    case 512: return pick_wg_b<512, WM>(hk);
    default: return pick_wg_b<1024, WM>(hk);
    }
}

WgFn pick_wg(int B, int hk, int wm)
{
    return wm == WM_BIDI ? pick_wg_p<WM_BIDI>(B, hk) : wm == WM_PILES ? pick_wg_p<WM_PILES>(B, hk)
                                                                       : pick_wg_p<WM_HEAPS>(B, hk);
}

}  // namespace

int ddp_ga_plan(ddp_ctx *ctx, const ddp_graph *g, const ddp_astar_opts_t *o, GaGridPlan *p, int64_t min_pcap,
                bool allow_wg)
{
    int32_t B = (o && o->block) ? o->block : 0;
    int32_t K = (o && o->K) ? o->K : 1024;
    DDP_CHECK(K >= 1, ctx, DDP_ERR_ARG, "K must be >= 1 (got %d)", K);
    const int32_t mode = (o && allow_wg) ? o->mode : DDP_ASTAR_GRID;
    DDP_CHECK(mode >= DDP_ASTAR_AUTO && mode <= DDP_ASTAR_BIDI, ctx, DDP_ERR_ARG, "unknown mode %d", mode);
    const bool one_wg = mode == DDP_ASTAR_WORKGROUP || mode == DDP_ASTAR_PILES || mode == DDP_ASTAR_BIDI;
    DDP_CHECK(!(one_wg && K > 1024), ctx, DDP_ERR_ARG, "one-workgroup modes hold at most 1024 lanes (K=%d)", K);
    if (one_wg) {
        const bool piles = mode == DDP_ASTAR_PILES || mode == DDP_ASTAR_BIDI;
        int32_t WB = mode == DDP_ASTAR_BIDI ? 256 : 64;  // two directions: half a block each
        while (WB < K) WB *= 2;
        DDP_CHECK(B == 0 || B == WB, ctx, DDP_ERR_ARG,
                  "one-workgroup mode runs K=%d on a block of %d lanes (got block %d)", K, WB, B);
        int32_t lcap = (o && o->lds_cap) ? o->lds_cap : std::max(1, std::min(32, 8192 / WB));
        if (piles) lcap = 0;
        DDP_CHECK((int64_t)lcap * WB * 8 <= 96 * 1024, ctx, DDP_ERR_ARG,
                  "lds_cap %d x block %d exceeds the LDS budget", lcap, WB);
        int64_t cap = (o && o->heap_cap) ? o->heap_cap : (2 * (int64_t)g->N + WB - 1) / WB + 256;
        if (piles && !(o && o->heap_cap))  // room for an exhaustive search: 3N entries
            cap = std::max<int64_t>(2048, (3 * (int64_t)g->N + WB - 1) / WB);
        cap = std::max<int64_t>(cap, std::max(lcap, 16));
        DDP_CHECK(cap < (1 << 30) && cap * WB < (1LL << 32), ctx, DDP_ERR_ARG, "heap_cap too large");
        const int32_t maxit =
            (o && o->max_iters) ? o->max_iters : (int32_t)std::min<int64_t>(4LL * g->N + 1024, 1 << 30);
        *p = GaGridPlan{WB, WB, 1, lcap, (int32_t)cap, 1024, maxit, true};
        if (piles) {
            const uint64_t total = (uint64_t)cap * WB;
            uint64_t nc = 1;
            while (nc * 2 <= total / 4) nc *= 2;
            p->piles = true;
            p->bidi = mode == DDP_ASTAR_BIDI;
            p->ncap = (uint32_t)nc;
            p->fcap = (uint32_t)std::min<uint64_t>((total - nc) / 2, 0xFFFFFFFFull);
            p->delta0 = std::max<uint32_t>(1, g->meanw);
        }
        return DDP_OK;
    }
    // the widest grids on wide workgroups: fewer workgroups make each
    // device-wide barrier cheaper.  One R-MAT scale-23 search: K = 65,536 took
    // 33 ms on 64 workgroups of 1,024 lanes and 42 ms on 256 of 256; K = 32,768
    // took 37 ms on 128 of 256 lanes and 41 ms on 32 of 1,024.
    if (B == 0) B = K >= 65536 ? 1024 : K >= 4096 ? 256 : 64;
    DDP_CHECK(B == 64 || B == 256 || B == 1024, ctx, DDP_ERR_ARG,
              "block must be 64, 256 or 1024 (got %d)", B);
    K = ((K + B - 1) / B) * B;
    const int32_t nblocks = K / B;
    int32_t lcap = (o && o->lds_cap) ? o->lds_cap : std::max(1, std::min(32, 8192 / B));
    DDP_CHECK((int64_t)lcap * B * 8 <= 96 * 1024, ctx, DDP_ERR_ARG,
              "lds_cap %d x block %d exceeds the LDS budget", lcap, B);
    int64_t cap = (o && o->heap_cap) ? o->heap_cap : (2 * (int64_t)g->N + K - 1) / K + 256;
    cap = std::max<int64_t>(cap, lcap);
    DDP_CHECK(cap < (1 << 30), ctx, DDP_ERR_ARG, "heap_cap too large");
    int64_t pcap = (int64_t)K * std::max(1, g->maxdeg) + K;
    pcap = std::min<int64_t>(pcap, (int64_t)g->E + 2 * (int64_t)K + 64);
    pcap = std::max<int64_t>(pcap, std::max<int64_t>(1024, min_pcap));
    DDP_CHECK(pcap < (1LL << 31), ctx, DDP_ERR_ARG, "push buffer too large");
    int32_t maxit = (o && o->max_iters) ? o->max_iters : (int32_t)std::min<int64_t>(4LL * g->N + 1024, 1 << 30);
    *p = GaGridPlan{K, B, nblocks, lcap, (int32_t)cap, (int32_t)pcap, maxit};
    // co-residency check for the cooperative launch
    int per_cu = 0;
    DDP_HIP(ctx, hipFuncSetAttribute((const void *)pick(B, DDP_H_ZERO), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)((size_t)lcap * B * 8)));
    DDP_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)pick(B, DDP_H_ZERO), B,
                                                             (size_t)lcap * B * 8));
    DDP_CHECK((int64_t)per_cu * ctx->prop.multiProcessorCount >= nblocks, ctx, DDP_ERR_ARG,
              "K=%d needs %d co-resident workgroups of %d lanes; the device holds %d", K, nblocks, B,
              per_cu * ctx->prop.multiProcessorCount);
    return DDP_OK;
}

int ddp_ga_alloc(ddp_ctx *ctx, const ddp_graph *g, const GaGridPlan &P, GaGridState *S)
{
    void *p;
    int rc;
    const size_t dirs = P.bidi ? 2 : 1;
    if ((rc = ddp_ws(ctx, WS_ASTAR_GP, dirs * g->N * 8, &p))) return rc;
    S->gp = (uint64_t *)p;
    if ((rc = ddp_ws(ctx, WS_ASTAR_HEAP, dirs * P.cap * P.K * 8, &p))) return rc;
    S->home = (uint64_t *)p;
    if ((rc = ddp_ws(ctx, WS_ASTAR_HSIZE, (size_t)P.K * 4, &p))) return rc;
    S->hsize = (int32_t *)p;
    if ((rc = ddp_ws(ctx, WS_ASTAR_PBUF, (size_t)P.pcap * 16, &p))) return rc;
    S->pbuf = (uint64_t *)p;
    if ((rc = ddp_ws(ctx, WS_ASTAR_CTRL, GA_CTRL_WORDS * 4 + 64, &p))) return rc;
    S->ctrl = (uint32_t *)p;
    S->stats = (unsigned long long *)((char *)p + GA_CTRL_WORDS * 4);
    S->touched = nullptr;
    if (P.wg) {
        if ((rc = ddp_ws(ctx, WS_ASTAR_TOUCH, dirs * g->N * 4, &p))) return rc;
        S->touched = (int32_t *)p;
    }
    S->owned = false;
    return DDP_OK;
}

int ddp_ga_alloc_owned(ddp_ctx *ctx, int32_t N, const GaGridPlan &P, GaGridState *S)
{
    *S = GaGridState{};
    S->owned = true;
    hipError_t e = hipSuccess;
    auto al = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 64));
    };
    al((void **)&S->gp, (size_t)N * 8);
    al((void **)&S->home, (size_t)P.cap * P.K * 8);
    al((void **)&S->hsize, (size_t)P.K * 4);
    al((void **)&S->pbuf, (size_t)P.pcap * 16);
    al((void **)&S->ctrl, GA_CTRL_WORDS * 4 + 64);
    if (e != hipSuccess) {
        ddp_ga_free(S);
        return ddp_hip_err(ctx, e, "GA* state allocation", __FILE__, __LINE__);
    }
    S->stats = (unsigned long long *)((char *)S->ctrl + GA_CTRL_WORDS * 4);
    return DDP_OK;
}

void ddp_ga_free(GaGridState *S)
{
    if (!S->owned) return;
    void *ptrs[] = {S->gp, S->home, S->hsize, S->pbuf, S->ctrl};
    for (void *p : ptrs)
        if (p) hipFree(p);
    *S = GaGridState{};
}

int ddp_ga_reset(ddp_ctx *ctx, int32_t N, const GaGridPlan &P, const GaGridState &S)
{
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1024), dim3(256), 0, ctx->stream, S.gp, (int64_t)N * (P.bidi ? 2 : 1),
                       (uint64_t)DDP_GP_INF);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(std::max(1, std::min(1024, P.K / 256 + 1))), dim3(256), 0,
                       ctx->stream, S.hsize, (int64_t)P.K, 0);
    DDP_HIP(ctx, hipMemsetAsync(S.stats, 0, 64, ctx->stream));
    DDP_HIP(ctx, hipGetLastError());
    return DDP_OK;
}

int ddp_ga_seed(ddp_ctx *ctx, const GaGridPlan &P, const GaGridState &S, int32_t s, uint32_t hs)
{
    hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(64), 0, ctx->stream, S.gp, S.pbuf + P.pcap, S.ctrl, s, hs);
    DDP_HIP(ctx, hipGetLastError());
    return DDP_OK;
}

__global__ void prepare_resume_kernel(uint32_t *ctrl)
{
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; ++k) {
        uint32_t *sl = ctrl + k * GA_SLOT_WORDS;
        sl[GA_S_PUSHCNT] = 0;
        sl[GA_S_TOPMIN] = 0xFFFFFFFFu;
        sl[GA_S_PUSHMIN] = 0xFFFFFFFFu;
        sl[GA_S_TMIN] = 0xFFFFFFFFu;
        sl[GA_S_ERR] = 0;
    }
}

int ddp_ga_prepare_resume(ddp_ctx *ctx, const GaGridState &S)
{
    hipLaunchKernelGGL(prepare_resume_kernel, dim3(1), dim3(64), 0, ctx->stream, S.ctrl);
    DDP_HIP(ctx, hipGetLastError());
    return DDP_OK;
}

int ddp_ga_stats(ddp_ctx *ctx, const GaGridState &S, unsigned long long out[4])
{
    DDP_HIP(ctx, hipMemcpyAsync(out, S.stats, 32, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DDP_OK;
}

namespace {

GridArgs grid_args(const GraphView &gv, const HeurView &hv, const GaGridPlan &P, const GaGridState &S, int32_t t,
                   uint32_t F0, int32_t rr0)
{
    GridArgs a{};
    a.g = gv;
    a.h = hv;
    a.gp = S.gp;
    a.home = S.home;
    a.hsize = S.hsize;
    a.pbuf = S.pbuf;
    a.ctrl = S.ctrl;
    a.stats = S.stats;
    a.K = P.K;
    a.lcap = P.lcap;
    a.cap = P.cap;
    a.pcap = P.pcap;
    a.t = t;
    a.maxit = P.maxit;
    a.F0 = F0;
    a.rr0 = rr0;
    a.barrier_ticks = 100ull * 1000 * 1000 * 20;  // 20 s of the 100 MHz clock
    return a;
}

// Ordinary launch: ddp_ga_plan sized the grid to what the device holds at
// once and the grid barrier is software with bounded waits (a cooperative
// launch would also reserve the global wave-sync unit, one process only).
int grid_enqueue(ddp_ctx *ctx, GridArgs &a, const GaGridPlan &P, int hk)
{
    void *args[] = {&a};
    KernelFn fn = pick(P.B, hk);
    DDP_HIP(ctx, hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)((size_t)P.lcap * P.B * 8)));
    DDP_HIP(ctx, hipLaunchKernel((const void *)fn, dim3(P.nblocks), dim3(P.B), args, (size_t)P.lcap * P.B * 8,
                                 ctx->stream));
    return DDP_OK;
}

}  // namespace

int ddp_ga_launch(ddp_ctx *ctx, const GraphView &gv, const HeurView &hv, const GaGridPlan &P,
                  const GaGridState &S, int32_t N, int32_t s, uint32_t hs, int32_t t)
{
    DDP_CHECK(!P.wg, ctx, DDP_ERR_STATE, "GA*: not a grid plan");
    int rc;
    if ((rc = ddp_ga_reset(ctx, N, P, S))) return rc;
    if ((rc = ddp_ga_seed(ctx, P, S, s, hs))) return rc;  // also clears the barrier and the result words
    GridArgs a = grid_args(gv, hv, P, S, t, DDP_INF_F, 0);
    return grid_enqueue(ctx, a, P, hv.kind);
}

// Runs the GA* loop from the current state (fresh or resumed).
int ddp_ga_run(ddp_ctx *ctx, const GraphView &gv, const HeurView &hv, const GaGridPlan &P,
               const GaGridState &S, int32_t t, uint32_t F0, int32_t rr0, uint32_t *F_out,
               int32_t *rr_out, int32_t *iters_out, float *kernel_ms)
{
    GridArgs a = grid_args(gv, hv, P, S, t, F0, rr0);
    // a previous launch that timed out may have left the barrier mid-count
    DDP_HIP(ctx, hipMemsetAsync(S.ctrl + GA_BAR, 0, 8, ctx->stream));
    DDP_HIP(ctx, hipMemsetAsync(S.ctrl + GA_RES_ERR, 0, 4, ctx->stream));
    DDP_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    int rc = grid_enqueue(ctx, a, P, hv.kind);
    if (rc) return rc;
    DDP_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t res[4];
    DDP_HIP(ctx, hipMemcpyAsync(res, S.ctrl + GA_RES_IT, 16, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    DDP_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    if (iters_out) *iters_out = (int32_t)res[0];
    if (rr_out) *rr_out = (int32_t)res[1];
    if (F_out) *F_out = res[2];
    const uint32_t err = res[3];
    DDP_CHECK(!(err & ERR_BARRIER), ctx, DDP_ERR_TIMEOUT, "GA*: grid barrier wait timed out");
    DDP_CHECK(!(err & ERR_ITERS), ctx, DDP_ERR_TIMEOUT, "GA*: iteration limit %d reached", P.maxit);
    DDP_CHECK(!(err & ERR_HEAP), ctx, DDP_ERR_OVERFLOW,
              "GA*: a queue exceeded heap_cap=%d entries; raise heap_cap", P.cap);
    DDP_CHECK(!(err & ERR_PBUF), ctx, DDP_ERR_OVERFLOW, "GA*: push buffer (%d entries) overflowed", P.pcap);
    return DDP_OK;
}

int ddp_ga_run_wg(ddp_ctx *ctx, const GraphView &gv, const HeurView &hv, const GaGridPlan &P,
                  const GaGridState &S, int32_t N, int32_t s, int32_t t, uint32_t *F_out, int32_t *iters_out,
                  float *kernel_ms, const GraphView *gr, int32_t *dpath, int32_t dpath_cap, int32_t *plen)
{
    DDP_CHECK(P.wg && S.touched && P.B == P.K && P.nblocks == 1, ctx, DDP_ERR_STATE, "GA*: not a workgroup plan");
    DDP_CHECK(!P.bidi || gr, ctx, DDP_ERR_STATE, "GA*: the bidirectional plan needs transposed rows");
    WgArgs a{};
    a.g = gv;
    if (gr) a.gr = *gr;
    a.path = dpath;
    a.path_cap = dpath_cap;
    a.h = hv;
    a.gp = S.gp;
    a.home = S.home;  // cap * K entries
    a.touched = S.touched;
    a.ctrl = S.ctrl;
    a.stats = S.stats;
    a.N = N;
    a.s = s;
    a.t = t;
    a.lcap = P.lcap;
    a.cap = P.cap;
    a.maxit = P.maxit;
    a.ncap = P.ncap;
    a.fcap = P.fcap;
    a.delta0 = P.delta0;
    a.lring = P.piles ? 8u * (uint32_t)P.B : 0u;  // LDS near ring
    const size_t lds = P.piles ? (size_t)a.lring * 8 : (size_t)P.lcap * P.B * 8;
    WgFn fn = pick_wg(P.B, hv.kind, P.bidi ? WM_BIDI : P.piles ? WM_PILES : WM_HEAPS);
    DDP_CHECK(fn, ctx, DDP_ERR_ARG, "no workgroup kernel for block %d with this heuristic and mode", P.B);
    DDP_HIP(ctx, hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DDP_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(fn, dim3(1), dim3(P.B), lds, ctx->stream, a);
    DDP_HIP(ctx, hipGetLastError());
    DDP_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t res[5];
    DDP_HIP(ctx, hipMemcpyAsync(res, S.ctrl + GA_RES_IT, 20, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    DDP_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    if (iters_out) *iters_out = (int32_t)res[0];
    if (F_out) *F_out = res[2];
    if (plen) *plen = (int32_t)res[4];
    const uint32_t err = res[3];
    DDP_CHECK(!(err & ERR_ITERS), ctx, DDP_ERR_TIMEOUT, "GA*: iteration limit %d reached", P.maxit);
    DDP_CHECK(!(err & ERR_HEAP), ctx, DDP_ERR_OVERFLOW,
              "GA*: a queue exceeded heap_cap=%d entries; raise heap_cap", P.cap);
    DDP_CHECK(!(err & ERR_PATH), ctx, DDP_ERR_STATE, "GA*: the parent chains through the meeting node are broken");
    return DDP_OK;
}

extern "C" int ddp_astar(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t s, int32_t t,
                         const ddp_astar_opts_t *opts, int32_t *cost, int32_t *parent_out,
                         int32_t *path_out, int32_t path_cap, int32_t *path_len,
                         ddp_astar_stats_t *stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !g || !cost) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_astar: null argument");
    DDP_CHECK(s >= 0 && s < g->N && t >= 0 && t < g->N, ctx, DDP_ERR_ARG,
              "ddp_astar: s=%d t=%d outside [0,%d)", s, t, g->N);
    DDP_CHECK(!path_out || path_len, ctx, DDP_ERR_ARG, "ddp_astar: path_out needs path_len");
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    s = ddp_to_dev(g, s);
    t = ddp_to_dev(g, t);
    HeurView hv;
    int rc = ddp_make_heur(g, h, t, &hv);
    if (rc) return rc;
    GaGridPlan P;
    if ((rc = ddp_ga_plan(ctx, g, opts, &P, 0, true))) return rc;
    if (P.bidi) {
        DDP_CHECK(g->roff, ctx, DDP_ERR_STATE, "DDP_ASTAR_BIDI needs a graph uploaded with DDP_GRAPH_REVERSE");
        DDP_CHECK(hv.kind != DDP_H_ARRAY, ctx, DDP_ERR_ARG,
                  "DDP_ASTAR_BIDI needs a closed-form heuristic (DDP_H_ZERO, _MANHATTAN or _EUCLID)");
    }
    GaGridState S;
    if ((rc = ddp_ga_alloc(ctx, g, P, &S))) return rc;
    std::vector<int32_t> bpath;  // bidirectional mode: the path walked on the device (device ids)
    int32_t hs = 0;
    if ((rc = ddp_heur_at(g, hv, s, &hs))) return rc;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->K = P.K;
        stats->block = P.B;
        stats->blocks = P.nblocks;
    }
    if ((rc = ddp_ga_reset(ctx, g->N, P, S))) return rc;
    if (hv.kind == DDP_H_ARRAY && hs < 0) {
        // h(s) = -1: the heuristic says s cannot reach t.
        *cost = DDP_INF_COST;
        if (path_len) *path_len = 0;
    } else if (s == t) {
        *cost = hs;
        if (path_out && path_cap >= 1) {
            path_out[0] = s;
            *path_len = 1;
        } else if (path_len) {
            *path_len = -1;
        }
    } else {
        uint32_t F = 0;
        int32_t iters = 0;
        float kms = 0;
        if (P.wg && P.bidi) {
            GraphView gv = ddp_view(g);
            gv.relax_min = g->minw >= 1;
            GraphView gr = gv;
            gr.off = g->roff;
            gr.col = g->rcol;
            gr.w = g->rw;
            // the path is walked on the device; a parent array needs all of it
            const int32_t dcap = parent_out ? g->N : (path_out ? std::max(path_cap, 1) : 0);
            int32_t *dpath = nullptr, plen = 0;
            void *p;
            if (dcap > 0) {
                if ((rc = ddp_ws(ctx, WS_PATH, ((size_t)dcap + 64) * 4, &p))) return rc;
                dpath = (int32_t *)p;
            }
            rc = ddp_ga_run_wg(ctx, gv, hv, P, S, g->N, s, t, &F, &iters, &kms, &gr, dpath, dcap, &plen);
            if (!rc && plen != 0) {
                bpath.resize(plen > 0 ? plen : 0);
                if (plen > 0)
                    DDP_HIP(ctx, hipMemcpy(bpath.data(), dpath, (size_t)plen * 4, hipMemcpyDeviceToHost));
                else
                    bpath.assign(1, plen);  // too long for dcap: only the (negative) length is known
            }
        } else if (P.wg) {
            GraphView gv = ddp_view(g);
            gv.relax_min = g->minw >= 1;
            rc = ddp_ga_run_wg(ctx, gv, hv, P, S, g->N, s, t, &F, &iters, &kms);
        } else {
            if ((rc = ddp_ga_seed(ctx, P, S, s, (uint32_t)hs))) return rc;
            rc = ddp_ga_run(ctx, ddp_view(g), hv, P, S, t, DDP_INF_F, 0, &F, nullptr, &iters, &kms);
        }
        if (rc) return rc;
        *cost = (int32_t)F;
        if (stats) {
            unsigned long long st[4];
            if ((rc = ddp_ga_stats(ctx, S, st))) return rc;
            stats->expansions = (int64_t)st[0];
            stats->edges = (int64_t)st[1];
            stats->improvements = (int64_t)st[2];
            stats->stale_pops = (int64_t)st[3];
            stats->iterations = iters;
            stats->kernel_ms = kms;
        }
        if (path_out && P.bidi) {
            if (F == DDP_INF_F || bpath.empty()) {
                *path_len = 0;
            } else if (bpath.size() == 1 && bpath[0] < 0) {
                *path_len = bpath[0];
            } else if ((int64_t)bpath.size() > path_cap) {
                *path_len = -(int32_t)bpath.size();
            } else {
                std::memcpy(path_out, bpath.data(), bpath.size() * 4);
                *path_len = (int32_t)bpath.size();
            }
        } else if (path_out) {
            if (F == DDP_INF_F) *path_len = 0;
            else if ((rc = ddp_extract_path(ctx, S.gp, g->N, s, t, path_out, path_cap, path_len))) return rc;
        } else if (path_len) {
            *path_len = 0;
        }
    }
    if (parent_out) {
        int32_t *dpar = nullptr;
        void *p;
        // reuse the push buffer region as a staging area when large enough
        if ((size_t)P.pcap * 16 >= (size_t)g->N * 4) dpar = (int32_t *)S.pbuf;
        else {
            if ((rc = ddp_ws(ctx, WS_PATH, (size_t)g->N * 4 + 256, &p))) return rc;
            dpar = (int32_t *)p;
        }
        hipLaunchKernelGGL(unpack_parent_kernel, dim3(1024), dim3(256), 0, ctx->stream, S.gp, dpar, g->N);
        DDP_HIP(ctx, hipMemcpyAsync(parent_out, dpar, (size_t)g->N * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        // bidirectional mode: the forward search's parents, and along the path
        // from the meeting node to t the path's own predecessors, so that the
        // parent chain from t leads back to s
        if (P.bidi && !(bpath.size() == 1 && bpath[0] < 0))
            for (size_t i = 1; i < bpath.size(); ++i) parent_out[bpath[i]] = bpath[i - 1];
        if (!g->iperm.empty()) {  // device ids and order -> caller ids and order
            std::vector<int32_t> tmp(parent_out, parent_out + g->N);
            for (int32_t i = 0; i < g->N; ++i) parent_out[g->iperm[i]] = ddp_to_user(g, tmp[i]);
        }
    }
    if (path_out && path_len && !g->iperm.empty())  // written only when the length is positive
        for (int32_t i = 0; i < *path_len; ++i) path_out[i] = ddp_to_user(g, path_out[i]);
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (stats)
        stats->total_ms =
            std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return DDP_OK;
}
