// Batched A*: many (s, t) queries in one persistent launch.
//
// The reference answers one query per process run (GA_star/GA_v2.cu:298-577)
// and its applications loop over queries on the host (WSN/wsn_2011.cu:355-456,
// k-short/GA_merge.cu:702-731); its global PQ symbol (GA_v2.cu:26) makes GA*
// non-reentrant.  Here each workgroup owns one query slot -- a private
// (f, parent) array and B lane-owned heaps -- and pulls query indices from a
// device work counter until the batch is exhausted.  Within a workgroup one
// GA* iteration pops one node per lane-queue and expands them cooperatively;
// an improved node is pushed into the heap of the lane that relaxed it, so the
// search never leaves the workgroup (no grid barrier, no cross-CU traffic).
// The slot's touched nodes are logged on first touch and reset when the query
// ends, so slots are reused without clearing N entries per query.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include <type_traits>

#include "../../include/ddp_batch.h"
#include "bidi_search.h"

namespace {

struct BatchArgs {
    GraphView g;
    GraphView gr;        // transposed rows (bidirectional engine)
    HeurView h;
    const int32_t *qs, *qt, *order;
    int32_t nq;
    uint32_t *next;
    uint64_t *gp;        // slots * N (bidirectional: slots * 2N, forward then backward)
    int32_t *touched;    // as gp
    uint64_t *home;      // slots * cap * B (bidirectional: twice that, one half per direction)
    int32_t *cost, *plen, *path;
    int32_t path_cap;
    uint32_t *qerr;
    uint64_t *qtime;     // 2 per query: start/end, 100 MHz device clock
    unsigned long long *stats;
    int32_t N, lcap, cap, maxit;
    uint32_t ncap, fcap, delta0;  // pile engine: near ring, far buffers, first window
    uint32_t lring;               // pile engine: entries of the LDS near ring (dynamic LDS)
};

enum { ENG_HEAPS = 0, ENG_PILES = 1, ENG_BIDI = 2 };

template <int B, int HK, int ENG>
__global__ void __launch_bounds__(B) batch_kernel(BatchArgs a)
{
    extern __shared__ uint64_t s_heap[];  // lcap * B (heap engine), near rings (pile engines)
    constexpr int SV = (ENG == ENG_PILES) ? B * pile_entries_per_lane<B>() : B;  // expansion width
    __shared__ std::conditional_t<ENG == ENG_BIDI, BidiLds<B>, SlotLds<B, SV>> S;
    __shared__ int32_t s_q;

    const int lane = threadIdx.x;
    const int64_t slot = blockIdx.x;
    const int32_t N = a.N;
    constexpr int DIRS = ENG == ENG_BIDI ? 2 : 1;
    uint64_t *gp = a.gp + slot * DIRS * N;
    int32_t *touched = a.touched + slot * DIRS * N;
    uint64_t *home = a.home + slot * DIRS * (int64_t)a.cap * B;  // this slot's queue memory
    const LaneHeap H{s_heap + lane, home + lane, B, a.lcap, (int64_t)B, a.cap};
    const PilePtrs P{home, home + a.ncap, home + a.ncap + a.fcap, a.ncap, a.fcap, s_heap, a.lring};
    ExpandCounters C;
    uint32_t stale = 0;
    for (;;) {
        if (lane == 0) s_q = (int32_t)atomicAdd(a.next, 1u);
        __syncthreads();
        const int32_t qi = s_q;
        if (qi >= a.nq) break;
        const int32_t q = a.order ? a.order[qi] : qi;
        const int32_t s = a.qs[q], t = a.qt[q];
        if (lane == 0) a.qtime[2 * (int64_t)q] = dev_clock_100mhz();
        if constexpr (ENG == ENG_BIDI) {
            // one half of the queue memory and of the LDS near ring per direction
            const uint64_t *bh = home + (int64_t)a.cap * B;
            const PilePtrs PB{(uint64_t *)bh, (uint64_t *)bh + a.ncap, (uint64_t *)bh + a.ncap + a.fcap, a.ncap, a.fcap,
                              s_heap + a.lring / 2, a.lring / 2};
            const PilePtrs PF{home, home + a.ncap, home + a.ncap + a.fcap, a.ncap, a.fcap, s_heap, a.lring / 2};
            const BidiSide D{{a.g.off, a.gr.off}, {a.g.col, a.gr.col}, {a.g.w, a.gr.w}, {gp, gp + N},
                             {touched, touched + N}, {PF, PB}};
            slot_search_bidi<B, HK>(D, a.h, a.g.relax_min != 0, N, s, t, a.maxit, a.delta0, S, C, stale);
            if (a.path_cap > 0) {
                slot_path_bidi(gp, gp + N, N, s, t, a.path + (int64_t)q * a.path_cap, a.path_cap, S);
                if (lane == 0) a.plen[q] = S.plen;
            }
            slot_reset_bidi(gp, gp + N, touched, touched + N, S);
            if (lane == 0) {
                a.cost[q] = (int32_t)(uint32_t)(S.meet >> 32);
                a.qerr[q] = S.err;
                a.qtime[2 * (int64_t)q + 1] = dev_clock_100mhz();
            }
        } else {
            if constexpr (ENG == ENG_PILES)
                slot_search_piles<B, HK>(a.g, a.h, gp, touched, P, N, s, t, a.maxit, a.delta0, S, C, stale);
            else
                slot_search<B, HK>(a.g, a.h, gp, touched, H, s_heap, N, s, t, a.maxit, S, C, stale);
            if (a.path_cap > 0) {
                slot_path(gp, N, s, t, a.path + (int64_t)q * a.path_cap, a.path_cap, S);
                if (lane == 0) a.plen[q] = S.plen;
            }
            slot_reset(gp, touched, S);
            if (lane == 0) {
                a.cost[q] = (int32_t)S.F;
                a.qerr[q] = S.err;
                a.qtime[2 * (int64_t)q + 1] = dev_clock_100mhz();
            }
        }
        if (__ballot(C.edges >= (1u << 30))) slot_flush_stats<B>(a.stats, C, stale);  // before 32 bits wrap
    }
    slot_flush_stats<B>(a.stats, C, stale);
}

__global__ void batch_fill_u64(uint64_t *p, int64_t n, uint64_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Resets the work counter and statistics before each launch (a kernel, not a
// memset node, so the sequence can be captured into a hipGraph).
__global__ void batch_reset(uint32_t *next, unsigned long long *stats)
{
    if (threadIdx.x == 0) *next = 0;
    if (threadIdx.x < 8) stats[threadIdx.x] = 0;
}

using BatchFn = void (*)(BatchArgs);

template <int B, int ENG>
BatchFn pick_b(int hk)
{
    switch (hk) {
    case DDP_H_ZERO: return batch_kernel<B, DDP_H_ZERO, ENG>;
    case DDP_H_ARRAY:
        if constexpr (ENG == ENG_BIDI) return nullptr;  // the potential needs h toward s as well as t
        else return batch_kernel<B, DDP_H_ARRAY, ENG>;
    case DDP_H_MANHATTAN: return batch_kernel<B, DDP_H_MANHATTAN, ENG>;
    default: return batch_kernel<B, DDP_H_EUCLID, ENG>;
    }
}

template <int ENG>
BatchFn pick_be(int B, int hk)
{
    if constexpr (ENG == ENG_BIDI) {  // two directions need half a block each: 256 lanes and up
        switch (B) {
        case 256: return pick_b<256, ENG>(hk);
        case 512: return pick_b<512, ENG>(hk);
        case 1024: return pick_b<1024, ENG>(hk);
        default: return nullptr;
        }
    }
    switch (B) {
    case 64: return pick_b<64, ENG>(hk);
    case 128: return pick_b<128, ENG>(hk);
    case 256: return pick_b<256, ENG>(hk);
    case 512: return pick_b<512, ENG>(hk);
    default: return pick_b<1024, ENG>(hk);
    }
}

BatchFn pick_batch(int B, int hk, int eng)
{
    if (eng == ENG_BIDI) return pick_be<ENG_BIDI>(B, hk);
    return eng == ENG_PILES ? pick_be<ENG_PILES>(B, hk) : pick_be<ENG_HEAPS>(B, hk);
}

}  // namespace

struct ddp_batch {
    ddp_ctx *ctx = nullptr;
    int device = 0;  // destroy uses this, never ctx (the context may be gone)
    const ddp_graph *g = nullptr;
    ddp_heur_t heur{};
    HeurView hv{};
    int32_t maxq = 0, B = 0, slots = 0, lcap = 0, cap = 0, maxit = 0, path_cap = 0, schedule = 0;
    int32_t eng = ENG_HEAPS;
    uint32_t ncap = 0, fcap = 0, delta0 = 1, lring = 0;  // pile engine sizes
    size_t lds = 0;                                       // dynamic LDS bytes per workgroup
    int32_t nq = 0;
    // device
    int32_t *qs = nullptr, *qt = nullptr, *order = nullptr, *cost = nullptr, *plen = nullptr, *path = nullptr,
            *touched = nullptr;
    uint32_t *next = nullptr, *qerr = nullptr;
    uint64_t *gp = nullptr, *home = nullptr, *qtime = nullptr;
    unsigned long long *stats = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t graph_nq = -1;
    float last_ms = 0;
};

static void batch_free(ddp_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    void *ptrs[] = {b->qs, b->qt, b->order, b->cost, b->plen, b->path, b->touched, b->next,
                    b->qerr, b->gp, b->home, b->qtime, b->stats};
    for (void *p : ptrs)
        if (p) hipFree(p);
}

extern "C" void ddp_batch_destroy(ddp_batch *b)
{
    if (!b) return;
    hipSetDevice(b->device);  // hipFree waits for work that uses the memory
    batch_free(b);
    delete b;
}

extern "C" int ddp_batch_create(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t max_queries,
                                const ddp_batch_opts_t *o, ddp_batch **out)
{
    if (!ctx || !g || !out) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_batch_create: null argument");
    *out = nullptr;
    DDP_CHECK(max_queries > 0, ctx, DDP_ERR_ARG, "ddp_batch_create: max_queries must be > 0");
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    HeurView hv;
    int rc = ddp_make_heur(g, h, -1, &hv);
    if (rc) return rc;
    const int32_t engine = o ? o->engine : DDP_BATCH_AUTO;
    DDP_CHECK(engine >= DDP_BATCH_AUTO && engine <= DDP_BATCH_BIDI, ctx, DDP_ERR_ARG, "unknown engine %d", engine);
    const int eng = engine == DDP_BATCH_HEAPS ? ENG_HEAPS : engine == DDP_BATCH_BIDI ? ENG_BIDI : ENG_PILES;
    const bool piles = eng != ENG_HEAPS;  // both pile engines
    int32_t B = (o && o->block) ? o->block : (piles ? 1024 : 256);
    DDP_CHECK(B == 64 || B == 128 || B == 256 || B == 512 || B == 1024, ctx, DDP_ERR_ARG,
              "block must be 64, 128, 256, 512 or 1024 (got %d)", B);
    if (eng == ENG_BIDI) {
        DDP_CHECK(hv.kind != DDP_H_ARRAY, ctx, DDP_ERR_ARG,
                  "the bidirectional engine needs a closed-form heuristic (DDP_H_ZERO, _MANHATTAN or _EUCLID)");
        DDP_CHECK(g->roff, ctx, DDP_ERR_STATE, "the bidirectional engine needs a graph uploaded with DDP_GRAPH_REVERSE");
        DDP_CHECK(B >= 256, ctx, DDP_ERR_ARG, "the bidirectional engine runs on blocks of 256, 512 or 1024 lanes");
    }
    const int DIRS = eng == ENG_BIDI ? 2 : 1;
    // deeper LDS heaps keep sift-downs out of HBM (about 10% on the C4 batch at
    // block 512, where occupancy is set by waves rather than LDS); the pile
    // engine keeps its queues in HBM and needs no dynamic LDS
    int32_t lcap = (o && o->lds_cap) ? o->lds_cap : std::max(4, std::min(32, 8192 / B));
    if (piles) lcap = 0;
    DDP_CHECK((int64_t)lcap * B * 8 <= 64 * 1024, ctx, DDP_ERR_ARG, "lds_cap %d x block %d exceeds 64 KiB", lcap, B);
    int32_t cap = (o && o->heap_cap) ? o->heap_cap : std::max(lcap, 2048);
    if (piles && !(o && o->heap_cap))  // room for an exhaustive search: 3N entries per slot and direction
        cap = (int32_t)std::min<int64_t>(std::max<int64_t>(cap, (3LL * g->N + B - 1) / B), INT32_MAX / 2);
    cap = std::max(cap, std::max(lcap, 16));
    // pile engine: the dynamic LDS holds the near ring (8 entries per lane,
    // 64 KiB at 1,024 lanes, which the one-slot-per-CU occupancy leaves free)
    const uint32_t lring = piles ? 8u * (uint32_t)B : 0u;  // bidirectional: half per direction
    const size_t lds = piles ? (size_t)lring * 8 : (size_t)lcap * B * 8;
    DDP_CHECK(pick_batch(B, hv.kind, eng), ctx, DDP_ERR_ARG, "no batch kernel for block %d and this engine", B);
    int per_cu = 0;
    DDP_HIP(ctx, hipFuncSetAttribute((const void *)pick_batch(B, hv.kind, eng),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DDP_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)pick_batch(B, hv.kind, eng), B,
                                                             lds));
    DDP_CHECK(per_cu > 0, ctx, DDP_ERR_ARG, "batch kernel does not fit on a CU with block %d", B);
    int64_t slots = (o && o->slots) ? o->slots : (int64_t)per_cu * ctx->prop.multiProcessorCount;
    slots = std::min<int64_t>(slots, max_queries);
    // admission by free HBM: keep 10% headroom
    size_t fr = 0, tot = 0;
    DDP_HIP(ctx, hipMemGetInfo(&fr, &tot));
    const size_t per_slot = DIRS * ((size_t)g->N * 12 + (size_t)cap * B * 8);
    const size_t fixed = (size_t)max_queries * (7 * 4 + 16 + (size_t)std::max(0, o ? o->path_cap : 0) * 4) + 4096;
    const size_t budget = fr > fixed ? (size_t)((fr - fixed) * 0.9) : 0;
    slots = std::min<int64_t>(slots, (int64_t)(budget / per_slot));
    DDP_CHECK(slots >= 1, ctx, DDP_ERR_NOMEM, "not enough free HBM for one query slot (%zu bytes needed, %zu free)",
              per_slot + fixed, fr);
    ddp_batch *b = new ddp_batch();
    b->ctx = ctx;
    b->device = ctx->device;
    b->g = g;
    b->heur = h;
    b->hv = hv;
    b->maxq = max_queries;
    b->B = B;
    b->slots = (int32_t)slots;
    b->lcap = lcap;
    b->cap = cap;
    b->maxit = (o && o->max_iters) ? o->max_iters : (int32_t)std::min<int64_t>(4LL * g->N + 1024, 1 << 30);
    b->path_cap = o ? std::max(0, o->path_cap) : 0;
    b->schedule = o ? o->schedule : 0;
    b->eng = eng;
    b->lring = lring;
    b->lds = lds;
    {
        // pile engine: the slot's cap * B queue entries hold a near ring (a
        // power of two, about a quarter) and two far buffers
        const uint64_t total = (uint64_t)cap * B;
        uint64_t nc = 1;
        while (nc * 2 <= total / 4) nc *= 2;
        b->ncap = (uint32_t)nc;
        b->fcap = (uint32_t)std::min<uint64_t>((total - nc) / 2, 0xFFFFFFFFull);
        b->delta0 = std::max<uint32_t>(1, g->meanw);
    }
    const size_t M = (size_t)max_queries;
    hipError_t e = hipSuccess;
    auto al = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 8));
    };
    al((void **)&b->qs, M * 4);
    al((void **)&b->qt, M * 4);
    al((void **)&b->order, M * 4);
    al((void **)&b->cost, M * 4);
    al((void **)&b->plen, M * 4);
    al((void **)&b->qerr, M * 4);
    al((void **)&b->qtime, M * 16);
    al((void **)&b->path, M * (size_t)b->path_cap * 4);
    al((void **)&b->next, 256);
    al((void **)&b->stats, 64);
    al((void **)&b->gp, (size_t)slots * DIRS * g->N * 8);
    al((void **)&b->touched, (size_t)slots * DIRS * g->N * 4);
    al((void **)&b->home, (size_t)slots * DIRS * cap * B * 8);
    if (e != hipSuccess) {
        batch_free(b);
        delete b;
        return ddp_hip_err(ctx, e, "batch allocation", __FILE__, __LINE__);
    }
    // one-time clear of every slot's (f, parent) array
    hipLaunchKernelGGL(batch_fill_u64, dim3(4096), dim3(256), 0, ctx->stream, b->gp, (int64_t)slots * DIRS * g->N,
                       (uint64_t)DDP_GP_INF);
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        batch_free(b);
        delete b;
        return ddp_hip_err(ctx, e, "batch slot clear", __FILE__, __LINE__);
    }
    *out = b;
    return DDP_OK;
}

extern "C" int ddp_batch_info(const ddp_batch *b, int32_t *slots, int32_t *block, int32_t *heap_cap,
                              int32_t *lds_cap)
{
    if (!b) return DDP_ERR_ARG;
    if (slots) *slots = b->slots;
    if (block) *block = b->B;
    if (heap_cap) *heap_cap = b->cap;
    if (lds_cap) *lds_cap = b->lcap;
    return DDP_OK;
}

extern "C" int ddp_batch_upload(ddp_batch *b, int32_t nq, const int32_t *src, const int32_t *dst)
{
    if (!b) return DDP_ERR_ARG;
    ddp_ctx *ctx = b->ctx;
    DDP_CHECK(nq >= 0 && nq <= b->maxq, ctx, DDP_ERR_ARG, "nq=%d outside [0,%d]", nq, b->maxq);
    DDP_CHECK(nq == 0 || (src && dst), ctx, DDP_ERR_ARG, "ddp_batch_upload: null query arrays");
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    if (b->hv.kind == DDP_H_ARRAY && b->g->h_target >= 0) {
        // an h array from ddp_bfs_hops estimates the distance to one target;
        // for any other target it may be inadmissible or mark reachable
        // nodes unreachable, so such queries are refused
        const int32_t ht = ddp_to_user(b->g, b->g->h_target);
        for (int32_t i = 0; i < nq; ++i)
            DDP_CHECK(dst[i] == ht, ctx, DDP_ERR_ARG,
                      "query %d: target %d, but the DDP_H_ARRAY heuristic was computed for target %d", i, dst[i], ht);
    }
    std::vector<int32_t> ms, md;
    if (!b->g->perm.empty()) {  // caller ids -> device ids (out-of-range ids stay, the kernel flags them)
        const int32_t N = b->g->N;
        ms.assign(src, src + nq);
        md.assign(dst, dst + nq);
        for (int32_t i = 0; i < nq; ++i) {
            if (ms[i] >= 0 && ms[i] < N) ms[i] = b->g->perm[ms[i]];
            if (md[i] >= 0 && md[i] < N) md[i] = b->g->perm[md[i]];
        }
        src = ms.data();
        dst = md.data();
    }
    std::vector<int32_t> ord(nq);
    std::iota(ord.begin(), ord.end(), 0);
    if (b->schedule == 0 && b->hv.kind == DDP_H_EUCLID && !b->g->host_xy.empty()) {
        // longest-estimate-first: straight-line distance orders the search sizes
        const auto &xy = b->g->host_xy;
        const int32_t N = b->g->N;
        std::vector<int64_t> est(nq, 0);
        for (int32_t i = 0; i < nq; ++i) {
            if (src[i] < 0 || src[i] >= N || dst[i] < 0 || dst[i] >= N) continue;
            const int64_t dx = xy[2 * (size_t)src[i]] - xy[2 * (size_t)dst[i]];
            const int64_t dy = xy[2 * (size_t)src[i] + 1] - xy[2 * (size_t)dst[i] + 1];
            est[i] = dx * dx + dy * dy;
        }
        std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return est[x] > est[y]; });
    }
    if (nq) {
        DDP_HIP(ctx, hipMemcpyAsync(b->qs, src, (size_t)nq * 4, hipMemcpyHostToDevice, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(b->qt, dst, (size_t)nq * 4, hipMemcpyHostToDevice, ctx->stream));
        DDP_HIP(ctx, hipMemcpyAsync(b->order, ord.data(), (size_t)nq * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->nq = nq;
    return DDP_OK;
}

static int batch_capture(ddp_batch *b)
{
    ddp_ctx *ctx = b->ctx;
    if (b->exec && b->graph_nq == b->nq) return DDP_OK;
    if (b->exec) {
        hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    BatchArgs a{};
    a.g = ddp_view(b->g);
    a.g.relax_min = b->g->minw >= 1;
    if (b->g->roff) {  // transposed rows (the bidirectional engine walks them backward from t)
        a.gr = a.g;
        a.gr.off = b->g->roff;
        a.gr.col = b->g->rcol;
        a.gr.w = b->g->rw;
    }
    a.h = b->hv;
    a.qs = b->qs;
    a.qt = b->qt;
    a.order = b->order;
    a.nq = b->nq;
    a.next = b->next;
    a.gp = b->gp;
    a.touched = b->touched;
    a.home = b->home;
    a.cost = b->cost;
    a.plen = b->plen;
    a.path = b->path;
    a.path_cap = b->path_cap;
    a.qerr = b->qerr;
    a.qtime = b->qtime;
    a.stats = b->stats;
    a.N = b->g->N;
    a.lcap = b->lcap;
    a.cap = b->cap;
    a.maxit = b->maxit;
    a.ncap = b->ncap;
    a.fcap = b->fcap;
    a.delta0 = b->delta0;
    a.lring = b->lring;
    const int32_t grid = std::max(1, std::min(b->slots, std::max(1, b->nq)));
    DDP_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, ctx->stream, b->next, b->stats);
    hipLaunchKernelGGL(pick_batch(b->B, b->hv.kind, b->eng), dim3(grid), dim3(b->B), b->lds, ctx->stream, a);
    hipError_t e = hipStreamEndCapture(ctx->stream, &b->graph);
    if (e != hipSuccess) return ddp_hip_err(ctx, e, "hipStreamEndCapture", __FILE__, __LINE__);
    DDP_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_nq = b->nq;
    return DDP_OK;
}

extern "C" int ddp_batch_launch(ddp_batch *b)
{
    if (!b) return DDP_ERR_ARG;
    ddp_ctx *ctx = b->ctx;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    int rc = batch_capture(b);
    if (rc) return rc;
    DDP_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    DDP_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    DDP_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    return DDP_OK;
}

extern "C" int ddp_batch_wait(ddp_batch *b, ddp_batch_stats_t *st)
{
    if (!b) return DDP_ERR_ARG;
    ddp_ctx *ctx = b->ctx;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    DDP_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    DDP_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    b->last_ms = ms;
    if (st) {
        std::memset(st, 0, sizeof *st);
        unsigned long long v[7];
        DDP_HIP(ctx, hipMemcpy(v, b->stats, sizeof v, hipMemcpyDeviceToHost));
        st->expansions = (int64_t)v[0];
        st->edges = (int64_t)v[1];
        st->improvements = (int64_t)v[2];
        st->stale_pops = (int64_t)v[3];
        st->pile_rebuilds = (int64_t)v[4];
        st->pile_refills = (int64_t)v[5];
        st->pile_scanned = (int64_t)v[6];
        st->slots = b->slots;
        st->block = b->B;
        st->kernel_ms = ms;
        std::vector<uint32_t> qe(b->nq);
        if (b->nq) DDP_HIP(ctx, hipMemcpy(qe.data(), b->qerr, (size_t)b->nq * 4, hipMemcpyDeviceToHost));
        for (uint32_t x : qe) st->errors += x != 0;
    }
    return DDP_OK;
}

extern "C" int ddp_batch_download(ddp_batch *b, int32_t *costs, int32_t *path_len, int32_t *paths,
                                  uint32_t *qerr, double *latency_ms)
{
    if (!b) return DDP_ERR_ARG;
    ddp_ctx *ctx = b->ctx;
    const size_t n = (size_t)b->nq;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    if (costs && n) DDP_HIP(ctx, hipMemcpyAsync(costs, b->cost, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (path_len && n) {
        if (b->path_cap > 0) DDP_HIP(ctx, hipMemcpyAsync(path_len, b->plen, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        else std::memset(path_len, 0, n * 4);
    }
    if (paths && n && b->path_cap > 0)
        DDP_HIP(ctx, hipMemcpyAsync(paths, b->path, n * b->path_cap * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (qerr && n) DDP_HIP(ctx, hipMemcpyAsync(qerr, b->qerr, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<uint64_t> tt;
    if (latency_ms && n) {
        tt.resize(2 * n);
        DDP_HIP(ctx, hipMemcpyAsync(tt.data(), b->qtime, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    }
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (paths && n && b->path_cap > 0 && !b->g->iperm.empty()) {  // device ids -> caller ids
        std::vector<int32_t> pl(n);
        DDP_HIP(ctx, hipMemcpyAsync(pl.data(), b->plen, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < n; ++i)
            for (int32_t k = 0; k < std::min(pl[i], b->path_cap); ++k)
                paths[i * b->path_cap + k] = ddp_to_user(b->g, paths[i * b->path_cap + k]);
    }
    if (latency_ms && n) {
        uint64_t t0 = ~0ull;
        for (size_t i = 0; i < n; ++i) t0 = std::min(t0, tt[2 * i]);
        // completion time of each query measured from the first query's start
        for (size_t i = 0; i < n; ++i) latency_ms[i] = (double)(tt[2 * i + 1] - t0) * 1e-5;
    }
    return DDP_OK;
}

extern "C" int ddp_astar_batch(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq, const int32_t *src,
                               const int32_t *dst, const ddp_batch_opts_t *opts, int32_t *costs,
                               int32_t *path_len, int32_t *paths, ddp_batch_stats_t *stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !g || !costs) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_astar_batch: null argument");
    if (nq == 0) return DDP_OK;
    ddp_batch *b = nullptr;
    int rc = ddp_batch_create(ctx, g, h, nq, opts, &b);
    if (rc) return rc;
    std::vector<uint32_t> qe(nq);
    if (!(rc = ddp_batch_upload(b, nq, src, dst)) && !(rc = ddp_batch_launch(b)) && !(rc = ddp_batch_wait(b, stats)))
        rc = ddp_batch_download(b, costs, path_len, paths, qe.data(), nullptr);
    ddp_batch_destroy(b);
    if (rc) return rc;
    int32_t nerr = 0, first = -1;
    uint32_t kind = 0;
    for (int32_t i = 0; i < nq; ++i)
        if (qe[i]) {
            if (first < 0) { first = i; kind = qe[i]; }
            ++nerr;
        }
    if (stats)
        stats->total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    DDP_CHECK(nerr == 0, ctx, (kind & QERR_HEAP) ? DDP_ERR_OVERFLOW : (kind & QERR_BADQ) ? DDP_ERR_ARG : DDP_ERR_TIMEOUT,
              "%d of %d queries failed (first: query %d, flags 0x%x: 1=heap_cap, 2=iterations, 4=bad id, 8=path)",
              nerr, nq, first, kind);
    return DDP_OK;
}
