// Online serving with continuous batching (see include/ddp_server.h).
//
// One persistent kernel: every workgroup is a slot that claims the next
// request from a ring in pinned, host-coherent memory as soon as it is free,
// so requests that arrive while others run are admitted as slots free up and
// their searches overlap on the device.  Query slots reuse the batched
// search (slot_search.h) on a versioned view of the graph.
//
// Graph versions.  Update batch k is applied by whichever slot claims it, in
// version order: each delete claims (CAS INF -> k) the delete-version word of
// the first instance of u->v still live, base row first and then inserted
// edges oldest first (DA_star/GA_merge.cu:1019-1037 semantics); inserts are
// appended to their source's list with links that carry k; then the applied
// version is published with release ordering.  A query at version V waits
// until V is applied, acquires it, and then reads only data written by
// versions <= V -- newer list entries are never dereferenced, and a stale copy
// of a newer delete version is as good as INF -- so its answer depends only on
// (s, t, V).
//
// Every device wait is bounded; idle workgroups exit after idle_ms and the host
// relaunches the kernel when new requests arrive.
//
// Grid mode at low load (opts.grid_k): a query that server_route_grid sends
// there (nothing else in flight, or arrivals slow enough for grid mode to
// keep up) is searched by the single-query GA* grid kernel (K lane heaps over
// the whole GPU, gastar.hip) instead of one slot.  Its ring entry is marked
// so that the serve kernel skips it, and its search runs on the server's
// stream after any serve kernel (which is told to stop: its running slots
// finish their requests and claim nothing published after the entry) and
// before the next launch; a one-lane kernel then writes its result into the
// ring like a slot would.  Updates are applied the same way, by a kernel of
// their own, unless an update is pending on the slot path or an open launch
// runs slot queries (serve_route.h).
#include <algorithm>
#include <array>
#include <atomic>
#include <tuple>
#include <chrono>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/ddp_server.h"
#include "bidi_search.h"
#include "compact.h"
#include "gastar.h"
#include "serve_route.h"

namespace {

struct Req {       // pinned, written by the host
    int32_t type;
    int32_t s, t;  // query endpoints
    int32_t version;
    int32_t del_off, del_n;  // updates: (u,v) pairs in the arena
    int32_t ins_off, ins_n;  // updates: (u,v,w,entry) quads, sorted by u
    int32_t rins_off;        // bidirectional engine: (v,entry) pairs, sorted by v (ins_n of them)
    int32_t relax_min;       // queries: every weight at this version is >= 1
    int32_t direct;          // handled by kernels of its own (a grid-mode query, or an update
                             // applied while no slot runs), skipped by the serve kernel
    int32_t stop_gen;        // direct entries: the launch told to stop when this entry was
                             // published (0: none ran); that launch claims nothing from here on
};

struct Res {       // pinned, written by the device
    uint32_t seq;  // ticket + 1 once valid
    int32_t cost;
    int32_t status;
    int32_t version;
    uint64_t t0, t1;
};

enum : uint32_t { SERR_WAIT = 16, SERR_LOG = 32 };
// ctrl words, one 64-byte line apart: claimed count, device copy of the host's
// published count, applied version, poller lease, sticky error, copy of the
// host's stop word, busy word.  Launches are numbered (gen) and the stop word
// holds the newest launch told to stop, so a launch exits once it is >= its
// own number and no flag is ever cleared while a kernel might still read it.
// The busy word counts slots running a request, plus kClosed once the launch
// closes for lack of work (cleared by the host before each launch).
// Counts are unsigned and compared by difference, so they may wrap.
enum : int {
    CT_HEAD = 0, CT_TAIL = 16, CT_APPLIED = 32, CT_POLL = 48, CT_ERR = 64, CT_STOP = 80, CT_BUSY = 96, CT_WORDS = 112
};
constexpr uint32_t kClosed = 0x80000000u;

struct ServeArgs {
    GraphView g;           // base rows + versioned view fields (ver set per query)
    GraphView gr;          // bidirectional engine: transposed base rows
    HeurView h;
    int32_t *dver;         // E, writable
    uint64_t *lhead;       // N, writable
    uint64_t *lnext;       // log capacity, writable
    int32_t *lv;
    uint32_t *lw;
    int32_t *ldel;
    int32_t *ltail;        // N, last entry of each list (updaters only)
    // bidirectional engine: delete version of each transposed entry, its
    // forward edge id, the source of each log entry and per-target lists
    int32_t *rdver;
    const int32_t *reid;
    int32_t *lu;
    uint64_t *rlhead, *rlnext;
    int32_t *rltail;
    int32_t log_cap;
    const Req *ring;       // pinned
    Res *res;              // pinned
    const int32_t *arena;  // pinned update payloads
    const uint32_t *tail;  // pinned: requests published
    const uint32_t *stop;  // pinned
    uint32_t ring_mask;
    uint32_t *ctrl;        // device: claimed count, applied version
    uint64_t *gp;
    int32_t *touched;
    uint64_t *home;
    unsigned long long *stats;
    int32_t N, lcap, cap, maxit;
    uint32_t ncap, fcap, delta0;  // pile engine: near ring, far buffers, first window
    uint32_t lring;               // pile engine: entries of the LDS near ring (dynamic LDS)
    uint64_t idle_ticks, wait_ticks;
    uint32_t gen;                 // this launch's number
};

enum { SENG_HEAPS = 0, SENG_PILES = 1, SENG_BIDI = 2 };

__device__ __forceinline__ int32_t sys_ld_i(const int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int B, bool BIDI>
__device__ void apply_update(const ServeArgs &a, const Req &rq, uint32_t *s_err)
{
    const int32_t V = rq.version;
    const int lane = threadIdx.x;
    // deletes: first instance still live at V-1 (its delete word is INF)
    for (int32_t i = lane; i < rq.del_n; i += B) {
        const int32_t u = sys_ld_i(a.arena + rq.del_off + 2 * i), v = sys_ld_i(a.arena + rq.del_off + 2 * i + 1);
        if (u < 0 || u >= a.N || v < 0 || v >= a.N) continue;
        bool done = false;
        for (int32_t j = a.g.off[u]; j < a.g.off[u + 1] && !done; ++j)
            if (a.g.col[j] == v && atomicCAS(a.dver + j, INT32_MAX, V) == INT32_MAX) {
                done = true;
                if constexpr (BIDI)  // the same edge in v's transposed row
                    for (int32_t r = a.gr.off[v]; r < a.gr.off[v + 1]; ++r)
                        if (a.reid[r] == j) {
                            a.rdver[r] = V;
                            break;
                        }
            }
        uint64_t link = __hip_atomic_load(a.lhead + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (!done && (uint32_t)link < (uint32_t)V) {
            const int32_t x = (int32_t)(link >> 32);
            link = __hip_atomic_load(a.lnext + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (a.lv[x] == v && atomicCAS(a.ldel + x, INT32_MAX, V) == INT32_MAX) done = true;
        }
    }
    block_sync_global();
    // inserts: one lane per source (entries arrive sorted by source, entry ids
    // assigned by the host in arrival order)
    for (int32_t i = lane; i < rq.ins_n; i += B) {
        const int32_t *q = a.arena + rq.ins_off + 4 * i;
        const int32_t u = sys_ld_i(q);
        if (i > 0 && sys_ld_i(q - 4) == u) continue;  // not the first of its source's group
        for (int32_t k = i; k < rq.ins_n; ++k) {
            const int32_t *e = a.arena + rq.ins_off + 4 * k;
            if (sys_ld_i(e) != u) break;
            const int32_t v = sys_ld_i(e + 1), x = sys_ld_i(e + 3);
            const uint32_t w = (uint32_t)sys_ld_i(e + 2);
            if (u < 0 || u >= a.N || v < 0 || v >= a.N || x < 0 || x >= a.log_cap) {
                atomicOr(s_err, (uint32_t)SERR_LOG);
                continue;
            }
            a.lv[x] = v;
            a.lw[x] = w;
            a.ldel[x] = INT32_MAX;
            a.lnext[x] = ~0ull;
            if constexpr (BIDI) {
                a.lu[x] = u;
                a.rlnext[x] = ~0ull;
            }
            const uint64_t link = ((uint64_t)(uint32_t)x << 32) | (uint32_t)V;
            const int32_t last = a.ltail[u];
            if (last < 0) a.lhead[u] = link;
            else a.lnext[last] = link;
            a.ltail[u] = x;
        }
    }
    block_sync_global();
    if constexpr (BIDI) {
        // the same entries on per-target lists, one lane per target (pairs
        // sorted by target, entry order kept within a target)
        for (int32_t i = lane; i < rq.ins_n; i += B) {
            const int32_t *q = a.arena + rq.rins_off + 2 * i;
            const int32_t v = sys_ld_i(q);
            if (i > 0 && sys_ld_i(q - 2) == v) continue;
            if (v < 0 || v >= a.N) continue;  // the forward pass flagged it
            for (int32_t k = i; k < rq.ins_n; ++k) {
                const int32_t *e = a.arena + rq.rins_off + 2 * k;
                if (sys_ld_i(e) != v) break;
                const int32_t x = sys_ld_i(e + 1);
                if (x < 0 || x >= a.log_cap) continue;
                const uint64_t link = ((uint64_t)(uint32_t)x << 32) | (uint32_t)V;
                const int32_t last = a.rltail[v];
                if (last < 0) a.rlhead[v] = link;
                else a.rlnext[last] = link;
                a.rltail[v] = x;
            }
        }
        block_sync_global();
    }
}

template <int B, int HK, int ENG>
__global__ void __launch_bounds__(B) serve_kernel(ServeArgs a)
{
    extern __shared__ uint64_t s_heap[];
    constexpr int SV = (ENG == SENG_PILES) ? B * pile_entries_per_lane<B>() : B;  // expansion width
    __shared__ std::conditional_t<ENG == SENG_BIDI, BidiLds<B>, SlotLds<B, SV>> S;
    __shared__ uint32_t s_ticket;
    __shared__ int32_t s_have;
    __shared__ uint32_t s_uerr;
    const int lane = threadIdx.x;
    const int64_t slot = blockIdx.x;
    const int32_t N = a.N;
    constexpr int DIRS = ENG == SENG_BIDI ? 2 : 1;
    uint64_t *gp = a.gp + slot * DIRS * N;
    int32_t *touched = a.touched + slot * DIRS * N;
    uint64_t *home = a.home + slot * DIRS * (int64_t)a.cap * B;  // this slot's queue memory
    const LaneHeap H{s_heap + lane, home + lane, B, a.lcap, (int64_t)B, a.cap};
    const PilePtrs P{home, home + a.ncap, home + a.ncap + a.fcap, a.ncap, a.fcap, s_heap, a.lring};
    ExpandCounters C;
    uint32_t stale = 0;
    for (;;) {
        // ---- claim the next published request (bounded idle wait).  Idle
        // slots watch device words only; one of them at a time holds the
        // poller lease and copies the host's published count (and the stop
        // flag) into them with plain relaxed loads of host memory, so idle
        // slots neither flood the host link nor invalidate caches that the
        // running searches depend on.
        //
        // A slot counts itself busy before it claims, and the launch closes
        // (every slot leaves) only by swapping a busy count of zero for
        // kClosed: slots never leave one by one while others still run, which
        // would strand later requests behind fewer slots until the last
        // running one finished and the host could relaunch.
        //
        // The next request is claimed only once the graph version it needs is
        // applied (entries handled by kernels of their own are skipped at
        // once), so no slot holds a claim while it waits: the version may come
        // from an update kernel queued behind this launch, which a stop or an
        // idle close then lets run.  A launch stopped for such an entry (its
        // stop_gen) claims nothing from there on: requests published after it
        // run after its kernels, in the next launch, instead of keeping this
        // one alive while that entry's kernels wait behind it.
        if (lane == 0) {
            uint32_t ticket = 0;
            bool have = false;
            const uint64_t t0 = dev_clock_100mhz();
            bool poller = false;
            int backoff = 0;  // non-pollers back off to ~15 us
            for (;;) {
                const uint32_t h = ld_u32(a.ctrl + CT_HEAD);
                const uint32_t tail = ld_u32(a.ctrl + CT_TAIL);
                const uint32_t bw = ld_u32(a.ctrl + CT_BUSY);
                if (bw & kClosed) break;
                bool ready = false;
                if ((int32_t)(tail - h) > 0) {
                    // a stale h reads an entry that its failed claim below ignores
                    __atomic_thread_fence(__ATOMIC_ACQUIRE);
                    const Req *hp = a.ring + (h & a.ring_mask);
                    ready = sys_ld_i(&hp->direct) != 0;
                    // this launch was stopped for this entry: what follows it
                    // starts after the entry's own kernels, in a later launch
                    if (ready && (uint32_t)sys_ld_i(&hp->stop_gen) == a.gen) break;
                    if (!ready) {
                        const int32_t ver = sys_ld_i(&hp->version);
                        const uint32_t need =
                            sys_ld_i(&hp->type) == DDP_REQ_UPDATE ? (uint32_t)ver - 1 : (uint32_t)ver;
                        ready = __hip_atomic_load(a.ctrl + CT_APPLIED, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >=
                                need;
                    }
                }
                if (ready) {
                    if (atomicCAS(a.ctrl + CT_BUSY, bw, bw + 1) != bw) continue;
                    if (atomicCAS(a.ctrl + CT_HEAD, h, h + 1) == h) {
                        ticket = h;
                        have = true;
                        break;
                    }
                    atomicSub(a.ctrl + CT_BUSY, 1u);
                    continue;
                }
                if ((int32_t)(ld_u32(a.ctrl + CT_STOP) - a.gen) >= 0) break;
                if (dev_clock_100mhz() - t0 > a.idle_ticks && bw == 0 && atomicCAS(a.ctrl + CT_BUSY, 0u, kClosed) == 0u)
                    break;
                if (!poller) poller = atomicCAS(a.ctrl + CT_POLL, 0u, 1u) == 0u;
                if (poller) {
                    // the lease holder is the only writer of the tail copy
                    const uint32_t ht = __hip_atomic_load(a.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    if ((int32_t)(ht - tail) > 0) st_u32(a.ctrl + CT_TAIL, ht);
                    const uint32_t sg = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    if ((int32_t)(sg - a.gen) >= 0) atomicExch(a.ctrl + CT_STOP, sg);
                    __builtin_amdgcn_s_sleep(16);
                } else {
                    backoff = backoff < 8 ? backoff + 1 : 8;
                    for (int k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(64);
                }
            }
            if (poller) atomicExch(a.ctrl + CT_POLL, 0u);
            // the request's fields were written before the published count:
            // one system-scope acquire per claim orders the reads below
            if (have) __atomic_thread_fence(__ATOMIC_ACQUIRE);
            s_ticket = ticket;
            s_have = have ? 1 : 0;
            s_uerr = 0;
        }
        __syncthreads();
        const uint32_t ticket = s_ticket;
        if (!s_have) break;
        const Req *rp = a.ring + (ticket & a.ring_mask);
        Req rq;
        rq.type = sys_ld_i(&rp->type);
        rq.s = sys_ld_i(&rp->s);
        rq.t = sys_ld_i(&rp->t);
        rq.version = sys_ld_i(&rp->version);
        rq.del_off = sys_ld_i(&rp->del_off);
        rq.del_n = sys_ld_i(&rp->del_n);
        rq.ins_off = sys_ld_i(&rp->ins_off);
        rq.ins_n = sys_ld_i(&rp->ins_n);
        rq.rins_off = sys_ld_i(&rp->rins_off);
        rq.relax_min = sys_ld_i(&rp->relax_min);
        if (sys_ld_i(&rp->direct)) {  // its own kernels write its result
            if (lane == 0) atomicSub(a.ctrl + CT_BUSY, 1u);
            __syncthreads();        // every lane has read s_ticket before lane 0 claims again
            continue;
        }
        const uint64_t tstart = dev_clock_100mhz();
        // ---- wait for the graph version this request needs (bounded)
        const uint32_t need = rq.type == DDP_REQ_UPDATE ? (uint32_t)rq.version - 1 : (uint32_t)rq.version;
        if (lane == 0) {
            while (__hip_atomic_load(a.ctrl + CT_APPLIED, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
                if (dev_clock_100mhz() - tstart > a.wait_ticks) {
                    s_uerr = SERR_WAIT;
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
        }
        // A failed update (a version wait that ran out, or a bad log entry)
        // leaves the graph without that version: the error is sticky and
        // every later request reports it instead of running.
        if (lane == 0) s_uerr |= ld_u32(a.ctrl + CT_ERR);
        __syncthreads();
        // every wave acquires the applied version before reading graph data
        (void)__hip_atomic_load(a.ctrl + CT_APPLIED, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        int32_t cost = DDP_INF_COST;
        uint32_t status = s_uerr;
        if (rq.type == DDP_REQ_UPDATE) {
            if (!status) apply_update<B, ENG == SENG_BIDI>(a, rq, &s_uerr);
            __syncthreads();
            status = s_uerr;
            if (lane == 0) {
                if (status) atomicOr(a.ctrl + CT_ERR, status);
                __threadfence();
                // published even after a failure (later requests then report the
                // sticky error instead of stalling); atomicMax keeps it from
                // moving backwards when an earlier version finishes late
                atomicMax(a.ctrl + CT_APPLIED, (uint32_t)rq.version);
            }
        } else if (!status) {
            GraphView G = a.g;
            G.ver = rq.version;
            G.relax_min = a.g.relax_min && rq.relax_min;
            if constexpr (ENG == SENG_BIDI) {
                uint64_t *bh = home + (int64_t)a.cap * B;
                const PilePtrs PF{home, home + a.ncap, home + a.ncap + a.fcap, a.ncap, a.fcap, s_heap, a.lring / 2};
                const PilePtrs PB{bh, bh + a.ncap, bh + a.ncap + a.fcap, a.ncap, a.fcap, s_heap + a.lring / 2,
                                  a.lring / 2};
                const BidiSide D{{a.g.off, a.gr.off},  {a.g.col, a.gr.col}, {a.g.w, a.gr.w},     {gp, gp + N},
                                 {touched, touched + N}, {PF, PB},          {a.dver, a.rdver}, rq.version,
                                 {a.lhead, a.rlhead},  {a.lnext, a.rlnext}, {a.lv, a.lu},       a.lw,
                                 a.ldel};
                slot_search_bidi<B, HK, true>(D, a.h, G.relax_min != 0, N, rq.s, rq.t, a.maxit, a.delta0, S, C,
                                              stale);
                cost = (int32_t)(uint32_t)(S.meet >> 32);
                status = S.err;
                slot_reset_bidi(gp, gp + N, touched, touched + N, S);
            } else {
                if constexpr (ENG == SENG_PILES)
                    slot_search_piles<B, HK>(G, a.h, gp, touched, P, N, rq.s, rq.t, a.maxit, a.delta0, S, C, stale);
                else
                    slot_search<B, HK>(G, a.h, gp, touched, H, s_heap, N, rq.s, rq.t, a.maxit, S, C, stale);
                cost = (int32_t)S.F;
                status = S.err;
                slot_reset(gp, touched, S);
            }
            if (__ballot(C.edges >= (1u << 30))) slot_flush_stats<B>(a.stats, C, stale);  // before 32 bits wrap
        }
        if (lane == 0) {
            Res *r = a.res + (ticket & a.ring_mask);
            __hip_atomic_store(&r->cost, cost, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&r->status, (int32_t)status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&r->version, rq.version, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&r->t0, tstart, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&r->t1, dev_clock_100mhz(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&r->seq, ticket + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            atomicSub(a.ctrl + CT_BUSY, 1u);
        }
        __syncthreads();
    }
    slot_flush_stats<B>(a.stats, C, stale);
}

__global__ void serve_fill_u64(uint64_t *p, int64_t n, uint64_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void serve_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Updates applied while no slot runs, by one workgroup launched for just that
// entry (stream order puts it after every earlier request's work, so the
// previous version is already applied).
template <bool BIDI>
__global__ void __launch_bounds__(1024) serve_update_kernel(ServeArgs a, uint32_t ticket)
{
    __shared__ uint32_t s_uerr;
    const int lane = threadIdx.x;
    const uint64_t tstart = dev_clock_100mhz();
    const Req *rp = a.ring + (ticket & a.ring_mask);
    Req rq{};
    rq.type = sys_ld_i(&rp->type);
    rq.version = sys_ld_i(&rp->version);
    rq.del_off = sys_ld_i(&rp->del_off);
    rq.del_n = sys_ld_i(&rp->del_n);
    rq.ins_off = sys_ld_i(&rp->ins_off);
    rq.ins_n = sys_ld_i(&rp->ins_n);
    rq.rins_off = sys_ld_i(&rp->rins_off);
    if (lane == 0) {
        s_uerr = ld_u32(a.ctrl + CT_ERR);
        if (__hip_atomic_load(a.ctrl + CT_APPLIED, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <
            (uint32_t)rq.version - 1)
            s_uerr |= SERR_WAIT;
    }
    __syncthreads();
    (void)__hip_atomic_load(a.ctrl + CT_APPLIED, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    if (!s_uerr) apply_update<1024, BIDI>(a, rq, &s_uerr);
    __syncthreads();
    if (lane == 0) {
        const uint32_t status = s_uerr;
        if (status) atomicOr(a.ctrl + CT_ERR, status);
        __threadfence();
        atomicMax(a.ctrl + CT_APPLIED, (uint32_t)rq.version);
        Res *r = a.res + (ticket & a.ring_mask);
        __hip_atomic_store(&r->cost, (int32_t)DDP_INF_COST, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&r->status, (int32_t)status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&r->version, rq.version, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&r->t0, tstart, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&r->t1, dev_clock_100mhz(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&r->seq, ticket + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Grid-mode queries: the device clock when the query's work starts, and its
// result written into the ring entry (GA_RES_F is the f of t, its cost, as
// h(t) = 0 for every closed-form heuristic).
__global__ void serve_grid_begin(uint64_t *stamp)
{
    if (threadIdx.x == 0) *stamp = dev_clock_100mhz();
}

__global__ void serve_grid_finish(const uint32_t *gctrl, const uint64_t *stamp, Res *r, uint32_t ticket,
                                  int32_t version)
{
    if (threadIdx.x != 0) return;
    const uint32_t err = gctrl[GA_RES_ERR];
    const int32_t status = ((err & (GA_ERR_HEAP | GA_ERR_PBUF)) ? 1 : 0) | ((err & GA_ERR_ITERS) ? 2 : 0) |
                           ((err & GA_ERR_BARRIER) ? 8 : 0);
    __hip_atomic_store(&r->cost, (int32_t)gctrl[GA_RES_F], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&r->status, status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&r->version, version, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&r->t0, *stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&r->t1, dev_clock_100mhz(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&r->seq, ticket + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

using ServeFn = void (*)(ServeArgs);

template <int B, int ENG>
ServeFn pick_s(int hk)
{
    switch (hk) {
    case DDP_H_ZERO: return serve_kernel<B, DDP_H_ZERO, ENG>;
    case DDP_H_ARRAY: return nullptr;  // refused at creation: one target per h array
    case DDP_H_MANHATTAN: return serve_kernel<B, DDP_H_MANHATTAN, ENG>;
    default: return serve_kernel<B, DDP_H_EUCLID, ENG>;
    }
}

template <int ENG>
ServeFn pick_se(int B, int hk)
{
    if constexpr (ENG == SENG_BIDI) {  // two directions need half a block each: 256 lanes and up
        switch (B) {
        case 256: return pick_s<256, ENG>(hk);
        case 512: return pick_s<512, ENG>(hk);
        case 1024: return pick_s<1024, ENG>(hk);
        default: return nullptr;
        }
    }
    switch (B) {
    case 64: return pick_s<64, ENG>(hk);
    case 128: return pick_s<128, ENG>(hk);
    case 256: return pick_s<256, ENG>(hk);
    case 512: return pick_s<512, ENG>(hk);
    default: return pick_s<1024, ENG>(hk);
    }
}

ServeFn pick_serve(int B, int hk, int eng)
{
    if (eng == SENG_BIDI) return pick_se<SENG_BIDI>(B, hk);
    return eng == SENG_PILES ? pick_se<SENG_PILES>(B, hk) : pick_se<SENG_HEAPS>(B, hk);
}

using clk = std::chrono::steady_clock;

}  // namespace

struct ddp_server {
    ddp_ctx *ctx = nullptr;
    int device = 0;
    const ddp_graph *g = nullptr;
    ddp_heur_t h{};
    HeurView hv{};
    int32_t B = 0, slots = 0, lcap = 0, cap = 0, maxit = 0;
    int32_t eng = SENG_PILES;
    uint32_t ncap = 0, fcap = 0, delta0 = 1, lring = 0;
    size_t lds = 0;  // dynamic LDS bytes per workgroup
    uint32_t ring = 0;
    int32_t log_cap = 0, log_used = 0;
    int32_t idle_ms = 2000;
    // host view of the graph (for nothing but sizes)
    int32_t N = 0, E = 0;
    // base rows after a compaction (owned; until the first one the caller's graph is read)
    MergedCsr own{}, own_r{};  // forward, transposed (bidirectional engine; eid = forward edge id)
    int32_t RE = 0;            // transposed entries
    bool base_relax_min = false;  // every base weight is >= 1
    int32_t compactions = 0;
    // device
    int32_t *dver = nullptr, *lv = nullptr, *ldel = nullptr, *ltail = nullptr, *touched = nullptr;
    int32_t *rdver = nullptr, *lu = nullptr, *rltail = nullptr;  // bidirectional engine
    uint64_t *rlhead = nullptr, *rlnext = nullptr;
    uint32_t *lw = nullptr, *ctrl = nullptr;
    uint64_t *lhead = nullptr, *lnext = nullptr, *gp = nullptr, *home = nullptr;
    unsigned long long *stats = nullptr;
    // pinned
    Req *ring_h = nullptr;
    Res *res_h = nullptr;
    int32_t *arena_h = nullptr;
    uint32_t *flags_h = nullptr;  // [0] tail, [1] stop
    Req *ring_d = nullptr;
    Res *res_d = nullptr;
    int32_t *arena_d = nullptr;
    uint32_t *flags_d = nullptr;
    size_t arena_cap = 0, arena_head = 0;       // int32 units, ring
    std::vector<std::tuple<uint32_t, size_t, size_t>> arena_marks;  // (ticket, start, end) in use
    // bookkeeping
    uint32_t submitted = 0, collected = 0;
    int32_t version = 0;
    int32_t zero_w_version = INT32_MAX;  // first version with an inserted weight-0 edge
    // per ring entry (request k lives at k & (ring - 1)); counts wrap modulo 2^32
    std::vector<uint64_t> tags;
    std::vector<double> submit_ms;
    std::vector<int32_t> types;
    std::vector<uint8_t> delivered;  // handed to the caller, not yet retired
    clk::time_point t_create;
    hipEvent_t done_ev = nullptr;
    bool running = false;
    uint32_t gen = 0;              // number of the newest launch
    // grid mode for low-load queries (server_route_grid)
    bool grid_on = false;
    GaGridPlan gplan{};
    GaGridState gstate{};
    uint64_t *gstamp = nullptr;    // device clock at the start of the current grid query
    hipEvent_t direct_ev = nullptr;  // after the newest direct request's kernels
    std::vector<uint8_t> direct_of;  // per ring entry: handled by kernels of its own
    int64_t requests = 0, grid_queries = 0, direct_updates = 0;
    RouteState route;  // what the grid-mode routing goes by (serve_route.h)
};

static double ms_now(const ddp_server *sv)
{
    return std::chrono::duration<double, std::milli>(clk::now() - sv->t_create).count();
}

static void server_free(ddp_server *sv)
{
    void *dev[] = {sv->dver, sv->lv, sv->ldel, sv->ltail, sv->touched, sv->lw, sv->ctrl, sv->lhead,
                   sv->lnext, sv->gp, sv->home, sv->stats, sv->rdver, sv->lu, sv->rltail, sv->rlhead, sv->rlnext};
    for (void *p : dev)
        if (p) hipFree(p);
    ddp_merged_free(&sv->own);
    ddp_merged_free(&sv->own_r);
    void *host[] = {sv->ring_h, sv->res_h, sv->arena_h, sv->flags_h};
    for (void *p : host)
        if (p) hipHostFree(p);
    if (sv->done_ev) hipEventDestroy(sv->done_ev);
    ddp_ga_free(&sv->gstate);
    if (sv->gstamp) hipFree(sv->gstamp);
    if (sv->direct_ev) hipEventDestroy(sv->direct_ev);
}

// Base rows (the caller's graph, or the merged rows after a compaction) with
// the versioned fields; queries set ver and relax_min.
static GraphView server_view(const ddp_server *sv)
{
    GraphView v = ddp_view(sv->g);
    if (sv->own.off) {
        v.off = sv->own.off;
        v.col = sv->own.col;
        v.w = sv->own.w;
    }
    v.relax_min = sv->base_relax_min;
    v.dver = sv->dver;
    v.lhead = sv->lhead;
    v.lnext = sv->lnext;
    v.lv = sv->lv;
    v.lw = sv->lw;
    v.ldel = sv->ldel;
    return v;
}

static ServeArgs server_args(const ddp_server *sv)
{
    ServeArgs a{};
    a.g = server_view(sv);
    if (sv->eng == SENG_BIDI) {
        a.gr = a.g;
        a.gr.off = sv->own_r.off ? sv->own_r.off : sv->g->roff;
        a.gr.col = sv->own_r.off ? sv->own_r.col : sv->g->rcol;
        a.gr.w = sv->own_r.off ? sv->own_r.w : sv->g->rw;
        a.reid = sv->own_r.off ? sv->own_r.eid : sv->g->reid;
        a.rdver = sv->rdver;
        a.lu = sv->lu;
        a.rlhead = sv->rlhead;
        a.rlnext = sv->rlnext;
        a.rltail = sv->rltail;
    }
    a.h = sv->hv;
    a.dver = sv->dver;
    a.lhead = sv->lhead;
    a.lnext = sv->lnext;
    a.lv = sv->lv;
    a.lw = sv->lw;
    a.ldel = sv->ldel;
    a.ltail = sv->ltail;
    a.log_cap = sv->log_cap;
    a.ring = sv->ring_d;
    a.res = sv->res_d;
    a.arena = sv->arena_d;
    a.tail = sv->flags_d;
    a.stop = sv->flags_d + 1;
    a.ring_mask = sv->ring - 1;
    a.ctrl = sv->ctrl;
    a.gp = sv->gp;
    a.touched = sv->touched;
    a.home = sv->home;
    a.stats = sv->stats;
    a.N = sv->N;
    a.lcap = sv->lcap;
    a.cap = sv->cap;
    a.maxit = sv->maxit;
    a.ncap = sv->ncap;
    a.fcap = sv->fcap;
    a.delta0 = sv->delta0;
    a.lring = sv->lring;
    a.idle_ticks = (uint64_t)sv->idle_ms * 100000ull;
    a.wait_ticks = 100000ull * 60000;  // a version wait longer than a minute is an error
    return a;
}

static int server_launch(ddp_server *sv)
{
    ddp_ctx *ctx = sv->ctx;
    ServeArgs a = server_args(sv);
    a.gen = ++sv->gen;
    void *args[] = {&a};
    DDP_HIP(ctx, hipMemsetAsync(sv->ctrl + CT_BUSY, 0, 4, ctx->stream));  // the last launch closed it
    DDP_HIP(ctx, hipLaunchKernel((const void *)pick_serve(sv->B, sv->hv.kind, sv->eng), dim3(sv->slots), dim3(sv->B),
                                 args, sv->lds, ctx->stream));
    DDP_HIP(ctx, hipEventRecord(sv->done_ev, ctx->stream));
    sv->running = true;
    return DDP_OK;
}

// Relaunches the kernel if every workgroup has exited while requests wait.
static int server_ensure(ddp_server *sv)
{
    if (sv->running) {
        const hipError_t q = hipEventQuery(sv->done_ev);
        if (q == hipErrorNotReady) return DDP_OK;
        if (q != hipSuccess) return ddp_hip_err(sv->ctx, q, "serve kernel", __FILE__, __LINE__);
        sv->running = false;
    }
    // with every workgroup gone, a request without a result was never claimed
    // (grid-mode queries get theirs from their own kernels)
    for (uint32_t k = sv->collected; k != sv->submitted; ++k)
        if (!sv->direct_of[k & (sv->ring - 1)] &&
            __atomic_load_n(&sv->res_h[k & (sv->ring - 1)].seq, __ATOMIC_ACQUIRE) != k + 1)
            return server_launch(sv);
    return DDP_OK;
}

// Requests without a result yet (serve_route.h).
static RouteInFlight server_in_flight(const ddp_server *sv)
{
    RouteInFlight f;
    const uint32_t mask = sv->ring - 1;
    const double now = ms_now(sv);
    double youngest = 0;
    for (uint32_t k = sv->collected; k != sv->submitted; ++k) {
        if (sv->delivered[k & mask] || __atomic_load_n(&sv->res_h[k & mask].seq, __ATOMIC_ACQUIRE) == k + 1)
            continue;
        if (sv->direct_of[k & mask]) {
            ++f.direct;
        } else if (sv->types[k & mask] == DDP_REQ_UPDATE) {
            ++f.slot_u;
        } else {
            ++f.slot_q;
            youngest = std::max(youngest, sv->submit_ms[k & mask]);
        }
    }
    f.slot_age = f.slot_q ? now - youngest : 0;
    f.slots_open = sv->running && __atomic_load_n(&sv->flags_h[1], __ATOMIC_RELAXED) != sv->gen;
    return f;
}

// Grid mode for the next query (route_grid).  Routing a query there stops the
// serve kernel (its running slots finish first): requests sent to slots later
// wait for the grid queue.
static bool server_route_grid(ddp_server *sv)
{
    return sv->grid_on && route_grid(sv->route, server_in_flight(sv));
}

static bool server_route_direct_update(const ddp_server *sv)
{
    return sv->grid_on && route_direct_update(server_in_flight(sv));
}

extern "C" int ddp_server_create(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, const ddp_server_opts_t *o,
                                 ddp_server **out)
{
    if (!ctx || !g || !out) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_server_create: null argument");
    *out = nullptr;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    HeurView hv;
    int rc = ddp_make_heur(g, h, -1, &hv);
    if (rc) return rc;
    DDP_CHECK(hv.kind != DDP_H_ARRAY, ctx, DDP_ERR_ARG,
              "serving needs a per-query heuristic: DDP_H_ZERO, DDP_H_MANHATTAN or DDP_H_EUCLID");
    const int32_t engine = o ? o->engine : DDP_SERVER_AUTO;
    DDP_CHECK(engine >= DDP_SERVER_AUTO && engine <= DDP_SERVER_BIDI, ctx, DDP_ERR_ARG, "unknown engine %d", engine);
    DDP_CHECK(engine != DDP_SERVER_BIDI || (g->roff && g->reid), ctx, DDP_ERR_STATE,
              "the bidirectional engine needs a graph uploaded with DDP_GRAPH_REVERSE");
    const int32_t grid_k = o ? o->grid_k : 0;
    DDP_CHECK(grid_k >= 0, ctx, DDP_ERR_ARG, "grid_k must be >= 0 (got %d)", grid_k);
    ddp_server *sv = new ddp_server();
    sv->ctx = ctx;
    sv->device = ctx->device;
    sv->g = g;
    sv->h = h;
    sv->hv = hv;
    sv->N = g->N;
    sv->E = g->E;
    sv->base_relax_min = g->minw >= 1;
    sv->eng = engine == DDP_SERVER_HEAPS ? SENG_HEAPS : engine == DDP_SERVER_BIDI ? SENG_BIDI : SENG_PILES;
    const bool piles = sv->eng != SENG_HEAPS;
    const int DIRS = sv->eng == SENG_BIDI ? 2 : 1;
    sv->RE = g->RE;
    // 1024-lane pile slots served the C4 mix with the best capacity and tail latency
    sv->B = (o && o->block) ? o->block : (piles ? 1024 : 256);
    sv->lcap = piles ? 0 : std::max(4, std::min(32, 4096 / std::max(64, sv->B)));
    sv->cap = (o && o->heap_cap) ? o->heap_cap : 2048;
    if (piles && !(o && o->heap_cap))  // room for an exhaustive search: 3N entries per slot and direction
        sv->cap = (int32_t)std::min<int64_t>(std::max<int64_t>(sv->cap, (3LL * g->N + sv->B - 1) / sv->B),
                                             INT32_MAX / 2);
    sv->cap = std::max(sv->cap, 16);
    {
        const uint64_t total = (uint64_t)sv->cap * sv->B;
        uint64_t nc = 1;
        while (nc * 2 <= total / 4) nc *= 2;
        sv->ncap = (uint32_t)nc;
        sv->fcap = (uint32_t)std::min<uint64_t>((total - nc) / 2, 0xFFFFFFFFull);
        sv->delta0 = std::max<uint32_t>(1, g->meanw);
    }
    sv->ring = (o && o->ring) ? (uint32_t)o->ring : 65536u;
    sv->log_cap = (o && o->max_inserts) ? o->max_inserts : (1 << 20);
    sv->idle_ms = (o && o->idle_ms) ? o->idle_ms : 2000;
    sv->maxit = (int32_t)std::min<int64_t>(4LL * g->N + 1024, 1 << 30);
    sv->t_create = clk::now();
    auto fail = [&](int code) {
        server_free(sv);
        delete sv;
        return code;
    };
    if (!(sv->B == 64 || sv->B == 128 || sv->B == 256 || sv->B == 512 || sv->B == 1024) ||
        (sv->ring & (sv->ring - 1)) || sv->ring < 64)
        return fail(ddp_set_err(ctx, DDP_ERR_ARG, "block must be 64..1024 and ring a power of two >= 64"));
    const void *fn = (const void *)pick_serve(sv->B, hv.kind, sv->eng);
    if (!fn) return fail(ddp_set_err(ctx, DDP_ERR_ARG, "no serving kernel for block %d and this engine", sv->B));
    if (grid_k > 0) {
        // the grid-mode state is allocated first: slots take what HBM is left
        ddp_astar_opts_t ao{};
        ao.K = grid_k;
        if ((rc = ddp_ga_plan(ctx, g, &ao, &sv->gplan))) return fail(rc);
        // room for 8N queued entries over the K heaps (a hub-heavy graph pushes
        // each node several times over a full search)
        sv->gplan.cap = (int32_t)std::max<int64_t>(sv->gplan.cap, (8LL * g->N + sv->gplan.K - 1) / sv->gplan.K);
        if ((rc = ddp_ga_alloc_owned(ctx, g->N, sv->gplan, &sv->gstate))) return fail(rc);
        hipError_t ge = hipMalloc((void **)&sv->gstamp, 64);
        if (ge == hipSuccess) ge = hipEventCreateWithFlags(&sv->direct_ev, hipEventDisableTiming);
        if (ge != hipSuccess) return fail(ddp_hip_err(ctx, ge, "grid-mode state", __FILE__, __LINE__));
        sv->grid_on = true;
    }
    int per_cu = 0;
    sv->lring = piles ? 8u * (uint32_t)sv->B : 0u;  // LDS near ring, as for batches (half per direction: bidi)
    sv->lds = piles ? (size_t)sv->lring * 8 : (size_t)sv->lcap * sv->B * 8;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sv->lds);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, sv->B, sv->lds);
    if (e != hipSuccess || per_cu < 1) return fail(ddp_set_err(ctx, DDP_ERR_HIP, "serve kernel occupancy query failed"));
    int64_t slots = (o && o->slots) ? o->slots : (int64_t)per_cu * ctx->prop.multiProcessorCount;
    slots = std::min<int64_t>(slots, (int64_t)per_cu * ctx->prop.multiProcessorCount);
    // admission by free HBM: graph-version state first, then as many slots as fit (10% headroom)
    size_t fr = 0, tot = 0;
    e = hipMemGetInfo(&fr, &tot);
    const size_t fixed = (size_t)g->E * 4 + (size_t)g->N * 12 + (size_t)sv->log_cap * 24 + (1 << 20) +
                         (DIRS == 2 ? (size_t)g->RE * 4 + (size_t)g->N * 12 + (size_t)sv->log_cap * 12 : 0);
    const size_t per_slot = DIRS * ((size_t)g->N * 12 + (size_t)sv->cap * sv->B * 8);
    const size_t budget = fr > fixed ? (size_t)((fr - fixed) * 0.9) : 0;
    slots = std::min<int64_t>(slots, (int64_t)(budget / per_slot));
    if (e != hipSuccess || slots < 1)
        return fail(ddp_set_err(ctx, DDP_ERR_NOMEM, "not enough free HBM for one serving slot"));
    sv->slots = (int32_t)slots;
    auto al = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 64));
    };
    al((void **)&sv->dver, (size_t)g->E * 4);
    al((void **)&sv->lhead, (size_t)g->N * 8);
    al((void **)&sv->ltail, (size_t)g->N * 4);
    al((void **)&sv->lnext, (size_t)sv->log_cap * 8);
    al((void **)&sv->lv, (size_t)sv->log_cap * 4);
    al((void **)&sv->lw, (size_t)sv->log_cap * 4);
    al((void **)&sv->ldel, (size_t)sv->log_cap * 4);
    al((void **)&sv->ctrl, CT_WORDS * 4);
    al((void **)&sv->stats, 64);
    al((void **)&sv->gp, (size_t)slots * DIRS * g->N * 8);
    al((void **)&sv->touched, (size_t)slots * DIRS * g->N * 4);
    al((void **)&sv->home, (size_t)slots * DIRS * sv->cap * sv->B * 8);
    if (DIRS == 2) {
        al((void **)&sv->rdver, (size_t)g->RE * 4);
        al((void **)&sv->lu, (size_t)sv->log_cap * 4);
        al((void **)&sv->rlhead, (size_t)g->N * 8);
        al((void **)&sv->rlnext, (size_t)sv->log_cap * 8);
        al((void **)&sv->rltail, (size_t)g->N * 4);
    }
    sv->arena_cap = (size_t)1 << 22;
    auto hal = [&](void **hp, void **dp, size_t bytes) {
        if (e == hipSuccess) e = hipHostMalloc(hp, bytes, hipHostMallocCoherent | hipHostMallocMapped);
        if (e == hipSuccess) std::memset(*hp, 0, bytes);
        if (e == hipSuccess) e = hipHostGetDevicePointer(dp, *hp, 0);
    };
    hal((void **)&sv->ring_h, (void **)&sv->ring_d, sizeof(Req) * sv->ring);
    hal((void **)&sv->res_h, (void **)&sv->res_d, sizeof(Res) * sv->ring);
    hal((void **)&sv->arena_h, (void **)&sv->arena_d, sv->arena_cap * 4);
    hal((void **)&sv->flags_h, (void **)&sv->flags_d, 64);
    sv->delivered.assign(sv->ring, 0);
    sv->direct_of.assign(sv->ring, 0);
    sv->tags.assign(sv->ring, 0);
    sv->types.assign(sv->ring, 0);
    sv->submit_ms.assign(sv->ring, 0.0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->done_ev, hipEventDisableTiming);
    if (e != hipSuccess) return fail(ddp_hip_err(ctx, e, "serving allocation", __FILE__, __LINE__));
    hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->dver, (int64_t)g->E, INT32_MAX);
    hipLaunchKernelGGL(serve_fill_u64, dim3(1024), dim3(256), 0, ctx->stream, sv->lhead, (int64_t)g->N, ~0ull);
    hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->ltail, (int64_t)g->N, -1);
    hipLaunchKernelGGL(serve_fill_u64, dim3(4096), dim3(256), 0, ctx->stream, sv->gp, (int64_t)slots * DIRS * g->N,
                       (uint64_t)DDP_GP_INF);
    if (DIRS == 2) {
        hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->rdver, (int64_t)g->RE, INT32_MAX);
        hipLaunchKernelGGL(serve_fill_u64, dim3(1024), dim3(256), 0, ctx->stream, sv->rlhead, (int64_t)g->N, ~0ull);
        hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->rltail, (int64_t)g->N, -1);
    }
    if (e == hipSuccess) e = hipMemsetAsync(sv->ctrl, 0, CT_WORDS * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(sv->stats, 0, 64, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ddp_hip_err(ctx, e, "serving setup", __FILE__, __LINE__));
    *out = sv;
    return DDP_OK;
}

extern "C" int ddp_server_info(const ddp_server *sv, int32_t *slots, int32_t *block, int32_t *version)
{
    if (!sv) return DDP_ERR_ARG;
    if (slots) *slots = sv->slots;
    if (block) *block = sv->B;
    if (version) *version = sv->version;
    return DDP_OK;
}

extern "C" int ddp_server_stats(const ddp_server *sv, ddp_server_stats_t *out)
{
    if (!sv || !out) return DDP_ERR_ARG;
    out->requests = sv->requests;
    out->grid_queries = sv->grid_queries;
    out->direct_updates = sv->direct_updates;
    out->launches = (int32_t)sv->gen;
    out->compactions = sv->compactions;
    return DDP_OK;
}

static void publish(ddp_server *sv, const Req &r, uint64_t tag, int32_t type)
{
    const uint32_t k = sv->submitted & (sv->ring - 1);
    sv->ring_h[k] = r;
    sv->direct_of[k] = r.direct ? 1 : 0;
    sv->requests += 1;
    sv->grid_queries += (r.direct && type == DDP_REQ_QUERY) ? 1 : 0;
    sv->direct_updates += (r.direct && type == DDP_REQ_UPDATE) ? 1 : 0;
    Res *res = sv->res_h + k;
    // cleared to the ticket itself: never equal to ticket + 1, even when it wraps
    __atomic_store_n(&res->seq, sv->submitted, __ATOMIC_RELAXED);
    sv->tags[k] = tag;
    sv->types[k] = type;
    sv->submit_ms[k] = ms_now(sv);
    route_note_submit(sv->route, sv->submit_ms[k]);
    ++sv->submitted;
    __atomic_store_n(&sv->flags_h[0], sv->submitted, __ATOMIC_RELEASE);
}

// A direct request none of whose kernels could be queued: its result says so.
static void direct_failed(ddp_server *sv, uint32_t ticket, int32_t version, int32_t status)
{
    Res *res = sv->res_h + (ticket & (sv->ring - 1));
    res->cost = DDP_INF_COST;
    res->status = status;
    res->version = version;
    res->t0 = res->t1 = 0;
    __atomic_store_n(&res->seq, ticket + 1, __ATOMIC_RELEASE);
}

// A query routed to grid mode.  The entry is published (its result word
// cleared) before any of its kernels is queued, so the result they write
// cannot be overwritten.  Grid queries queue on the stream one after another.
static int server_grid_query(ddp_server *sv, Req r, uint64_t tag)
{
    ddp_ctx *ctx = sv->ctx;
    HeurView hq;  // the heuristic aimed at this query's target
    int rc = ddp_make_heur(sv->g, sv->h, r.t, &hq);
    if (rc) return rc;
    int32_t hs = 0;
    if ((rc = ddp_heur_at(sv->g, hq, r.s, &hs))) return rc;
    GraphView gv = server_view(sv);
    gv.ver = r.version;
    gv.relax_min = sv->base_relax_min && r.relax_min;
    // an idle serve kernel would hold the stream for idle_ms: tell it to stop
    // (its running slots finish their requests and claim none after this one)
    if (sv->running) __atomic_store_n(&sv->flags_h[1], sv->gen, __ATOMIC_RELEASE);
    r.direct = 1;
    r.stop_gen = sv->running ? (int32_t)sv->gen : 0;
    const uint32_t ticket = sv->submitted;
    publish(sv, r, tag, DDP_REQ_QUERY);
    hipLaunchKernelGGL(serve_grid_begin, dim3(1), dim3(64), 0, ctx->stream, sv->gstamp);
    rc = ddp_ga_launch(ctx, gv, hq, sv->gplan, sv->gstate, sv->N, r.s, (uint32_t)hs, r.t);
    if (!rc)
        hipLaunchKernelGGL(serve_grid_finish, dim3(1), dim3(64), 0, ctx->stream, sv->gstate.ctrl, sv->gstamp,
                           sv->res_d + (ticket & (sv->ring - 1)), ticket, r.version);
    const hipError_t e = rc ? hipSuccess : hipGetLastError();
    if (rc || e != hipSuccess) {
        direct_failed(sv, ticket, r.version, 8);
        return rc ? rc : ddp_hip_err(ctx, e, "grid-mode query", __FILE__, __LINE__);
    }
    DDP_HIP(ctx, hipEventRecord(sv->direct_ev, ctx->stream));
    return DDP_OK;
}

// An update applied while no slot runs: one workgroup for just this entry.
static int server_direct_update(ddp_server *sv, uint32_t ticket, int32_t version)
{
    ddp_ctx *ctx = sv->ctx;
    ServeArgs a = server_args(sv);
    void *args[] = {&a, &ticket};
    const void *fn =
        sv->eng == SENG_BIDI ? (const void *)serve_update_kernel<true> : (const void *)serve_update_kernel<false>;
    hipError_t e = hipLaunchKernel(fn, dim3(1), dim3(1024), args, 0, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(sv->direct_ev, ctx->stream);
    if (e != hipSuccess) {
        direct_failed(sv, ticket, version, SERR_WAIT);
        return ddp_hip_err(ctx, e, "direct update", __FILE__, __LINE__);
    }
    return DDP_OK;
}

extern "C" int ddp_server_submit_query(ddp_server *sv, int32_t s, int32_t t, uint64_t tag)
{
    if (!sv) return DDP_ERR_ARG;
    ddp_ctx *ctx = sv->ctx;
    DDP_CHECK(s >= 0 && s < sv->N && t >= 0 && t < sv->N, ctx, DDP_ERR_ARG, "query %d->%d outside [0,%d)", s, t, sv->N);
    if (sv->submitted - sv->collected >= sv->ring) return ddp_set_err(ctx, DDP_ERR_NOMEM, "request ring full");
    Req r{};
    r.type = DDP_REQ_QUERY;
    r.s = ddp_to_dev(sv->g, s);
    r.t = ddp_to_dev(sv->g, t);
    r.version = sv->version;
    r.relax_min = sv->version < sv->zero_w_version;  // the base graph's own minimum is in a.g
    if (r.s != r.t && server_route_grid(sv)) return server_grid_query(sv, r, tag);
    publish(sv, r, tag, DDP_REQ_QUERY);
    return server_ensure(sv);
}

// Folds the inserted-edge log into the base rows at the current version
// (merge_graph / mergeDiff, DA_star/DA_dyn_insert_del_nocycle.cu:940-980,
// 2236-2284): every submitted request is finished first and the kernel is
// stopped, since queries read the log and the delete versions; the merged rows
// of the current version V become the base rows (their delete versions INF)
// and the log starts empty.  A query at version >= V reads the same graph
// through the new rows, and later updates number their versions on from V.
static int server_compact(ddp_server *sv)
{
    ddp_ctx *ctx = sv->ctx;
    int rc = ddp_server_drain(sv, 600000);
    if (rc) return rc;
    if (sv->running) {  // idle slots exit on the stop word
        __atomic_store_n(&sv->flags_h[1], sv->gen, __ATOMIC_RELEASE);
        DDP_HIP(ctx, hipEventSynchronize(sv->done_ev));
        sv->running = false;
    }
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));  // and any grid-mode query
    const GraphView base = server_view(sv);
    const MergeSrc m{base.off, base.col, base.w, sv->N, sv->dver, sv->version, sv->lhead, sv->lnext,
                     sv->lv, sv->lw, sv->ldel};
    MergedCsr f, fr;
    if ((rc = ddp_merge_rows(ctx, m, &f))) return rc;
    if (sv->eng == SENG_BIDI && (rc = ddp_transpose_rows(ctx, f, true, &fr))) {
        ddp_merged_free(&f);
        return rc;
    }
    int32_t *dver = nullptr, *rdver = nullptr;
    hipError_t e = hipMalloc((void **)&dver, (size_t)std::max(f.E, 1) * 4);
    if (e == hipSuccess && sv->eng == SENG_BIDI) e = hipMalloc((void **)&rdver, (size_t)std::max(fr.E, 1) * 4);
    if (e != hipSuccess) {
        if (dver) hipFree(dver);
        ddp_merged_free(&f);
        ddp_merged_free(&fr);
        return ddp_hip_err(ctx, e, "compaction: delete versions", __FILE__, __LINE__);
    }
    hipFree(sv->dver);
    sv->dver = dver;
    ddp_merged_free(&sv->own);
    sv->own = f;
    sv->E = f.E;
    if (sv->eng == SENG_BIDI) {
        hipFree(sv->rdver);
        sv->rdver = rdver;
        ddp_merged_free(&sv->own_r);
        sv->own_r = fr;
        sv->RE = fr.E;
        hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->rdver, (int64_t)fr.E, INT32_MAX);
        hipLaunchKernelGGL(serve_fill_u64, dim3(1024), dim3(256), 0, ctx->stream, sv->rlhead, (int64_t)sv->N, ~0ull);
        hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->rltail, (int64_t)sv->N, -1);
    }
    // a zero-weight insert at or before this version is now a base edge
    sv->base_relax_min = sv->base_relax_min && sv->zero_w_version > sv->version;
    hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->dver, (int64_t)f.E, INT32_MAX);
    hipLaunchKernelGGL(serve_fill_u64, dim3(1024), dim3(256), 0, ctx->stream, sv->lhead, (int64_t)sv->N, ~0ull);
    hipLaunchKernelGGL(serve_fill_i32, dim3(1024), dim3(256), 0, ctx->stream, sv->ltail, (int64_t)sv->N, -1);
    DDP_HIP(ctx, hipGetLastError());
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    sv->log_used = 0;
    sv->compactions += 1;
    return DDP_OK;
}

extern "C" int ddp_server_compact(ddp_server *sv)
{
    if (!sv) return DDP_ERR_ARG;
    DDP_HIP(sv->ctx, hipSetDevice(sv->ctx->device));
    return server_compact(sv);
}

extern "C" int ddp_server_submit_update(ddp_server *sv, const ddp_update_t *lines, int32_t n, uint64_t tag)
{
    if (!sv || (n && !lines)) return DDP_ERR_ARG;
    ddp_ctx *ctx = sv->ctx;
    if (sv->submitted - sv->collected >= sv->ring) return ddp_set_err(ctx, DDP_ERR_NOMEM, "request ring full");
    std::vector<int32_t> dels;
    std::vector<std::array<int32_t, 3>> ins;
    for (int32_t i = 0; i < n; ++i) {
        const ddp_update_t &l = lines[i];
        DDP_CHECK((l.flag == 0 || l.flag == 1) && l.u >= 0 && l.u < sv->N && l.v >= 0 && l.v < sv->N, ctx, DDP_ERR_ARG,
                  "update line %d invalid", i);
        const int32_t u = ddp_to_dev(sv->g, l.u), v = ddp_to_dev(sv->g, l.v);
        if (l.flag == 0) {
            dels.push_back(u);
            dels.push_back(v);
        } else {
            ins.push_back({u, v, (int32_t)l.w});
            if (l.w == 0) sv->zero_w_version = std::min(sv->zero_w_version, sv->version + 1);
        }
    }
    DDP_CHECK((int64_t)ins.size() <= sv->log_cap, ctx, DDP_ERR_CAPACITY,
              "update batch of %zu inserts exceeds the inserted-edge log (%d entries); raise max_inserts",
              ins.size(), sv->log_cap);
    if (sv->log_used + (int64_t)ins.size() > sv->log_cap) {
        // the log is full: fold it into the rows (waits for every request in flight)
        int rc = server_compact(sv);
        if (rc) return rc;
    }
    // entry ids in arrival order, then grouped by source (stable)
    std::vector<std::array<int32_t, 4>> q(ins.size());
    for (size_t k = 0; k < ins.size(); ++k) q[k] = {ins[k][0], ins[k][1], ins[k][2], sv->log_used + (int32_t)k};
    std::stable_sort(q.begin(), q.end(), [](const auto &x, const auto &y) { return x[0] < y[0]; });
    // bidirectional engine: the same entries grouped by target (stable in entry order)
    std::vector<std::array<int32_t, 2>> rq;
    if (sv->eng == SENG_BIDI) {
        rq.resize(ins.size());
        for (size_t k = 0; k < ins.size(); ++k) rq[k] = {ins[k][1], sv->log_used + (int32_t)k};
        std::stable_sort(rq.begin(), rq.end(), [](const auto &x, const auto &y) { return x[0] < y[0]; });
    }
    const size_t need = dels.size() + 4 * q.size() + 2 * rq.size();
    DDP_CHECK(need < sv->arena_cap / 2, ctx, DDP_ERR_ARG, "update batch too large for the payload arena");
    // The arena is a ring: a payload is free again once its update has been
    // collected (the device is done with it).  Refuse (caller polls and
    // retries) if the new payload would overlap one still in use.
    size_t pos = sv->arena_head;
    if (pos + need > sv->arena_cap) pos = 0;
    while (!sv->arena_marks.empty() && (int32_t)(std::get<0>(sv->arena_marks.front()) - sv->collected) < 0)
        sv->arena_marks.erase(sv->arena_marks.begin());
    for (const auto &m : sv->arena_marks)
        if (pos < std::get<2>(m) && std::get<1>(m) < pos + need)
            return ddp_set_err(ctx, DDP_ERR_NOMEM, "update payload arena busy; poll and retry");
    std::memcpy(sv->arena_h + pos, dels.data(), dels.size() * 4);
    for (size_t k = 0; k < q.size(); ++k) std::memcpy(sv->arena_h + pos + dels.size() + 4 * k, q[k].data(), 16);
    const size_t roff = pos + dels.size() + 4 * q.size();
    for (size_t k = 0; k < rq.size(); ++k) std::memcpy(sv->arena_h + roff + 2 * k, rq[k].data(), 8);
    // With nothing in flight on the slot path, the update is applied by a
    // kernel of its own: no slot launch, so later queries can stay in grid mode.
    const bool direct = server_route_direct_update(sv);
    Req r{};
    r.type = DDP_REQ_UPDATE;
    r.direct = direct ? 1 : 0;
    r.version = sv->version + 1;
    r.del_off = (int32_t)pos;
    r.del_n = (int32_t)(dels.size() / 2);
    r.ins_off = (int32_t)(pos + dels.size());
    r.ins_n = (int32_t)q.size();
    r.rins_off = (int32_t)roff;
    sv->arena_head = pos + need;
    sv->arena_marks.emplace_back(sv->submitted, pos, pos + need);
    sv->log_used += (int32_t)q.size();
    sv->version += 1;
    const uint32_t ticket = sv->submitted;
    if (direct && sv->running) {  // idle: stop it
        __atomic_store_n(&sv->flags_h[1], sv->gen, __ATOMIC_RELEASE);
        r.stop_gen = (int32_t)sv->gen;
    }
    publish(sv, r, tag, DDP_REQ_UPDATE);
    if (direct) return server_direct_update(sv, ticket, r.version);
    return server_ensure(sv);
}

extern "C" int ddp_server_poll(ddp_server *sv, ddp_server_result_t *out, int32_t max, int32_t *n)
{
    if (!sv || !n || (max > 0 && !out)) return DDP_ERR_ARG;
    *n = 0;
    // Results are handed out as they finish, in any order; a ring entry is
    // retired (reusable) once it and every earlier one have been handed out.
    const uint32_t mask = sv->ring - 1;
    for (uint32_t k = sv->collected; k != sv->submitted && *n < max; ++k) {
        if (sv->delivered[k & mask]) continue;
        Res *r = sv->res_h + (k & mask);
        if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) != k + 1) continue;
        ddp_server_result_t &o = out[(*n)++];
        o.tag = sv->tags[k & mask];
        o.type = sv->types[k & mask];
        o.cost = r->cost;
        o.version = r->version;
        o.status = r->status;
        o.submit_ms = sv->submit_ms[k & mask];
        o.done_ms = ms_now(sv);
        o.device_ms = (double)(r->t1 - r->t0) * 1e-5;
        sv->delivered[k & mask] = 1;
        if (o.type == DDP_REQ_QUERY && o.status == 0 && sv->direct_of[k & mask])  // for the grid-mode routing
            route_note_grid(sv->route, o.device_ms);
    }
    while (sv->collected != sv->submitted && sv->delivered[sv->collected & mask]) {
        sv->delivered[sv->collected & mask] = 0;
        ++sv->collected;
    }
    return server_ensure(sv);
}

extern "C" int ddp_server_drain(ddp_server *sv, int32_t timeout_ms)
{
    if (!sv) return DDP_ERR_ARG;
    const auto t0 = clk::now();
    while (sv->collected != sv->submitted) {
        // results are kept for the caller: only wait here, do not consume
        const uint32_t k = sv->collected;
        uint32_t done = k;
        while (done != sv->submitted &&
               (sv->delivered[done & (sv->ring - 1)] ||
                __atomic_load_n(&sv->res_h[done & (sv->ring - 1)].seq, __ATOMIC_ACQUIRE) == done + 1))
            ++done;
        if (done == sv->submitted) break;
        int rc = server_ensure(sv);
        if (rc) return rc;
        if (std::chrono::duration<double, std::milli>(clk::now() - t0).count() > timeout_ms)
            return ddp_set_err(sv->ctx, DDP_ERR_TIMEOUT, "drain: %u requests still running", sv->submitted - done);
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    return DDP_OK;
}

extern "C" void ddp_server_destroy(ddp_server *sv)
{
    if (!sv) return;
    hipSetDevice(sv->device);
    if (sv->flags_h) __atomic_store_n(&sv->flags_h[1], sv->gen, __ATOMIC_RELEASE);
    if (sv->running && sv->done_ev) hipEventSynchronize(sv->done_ev);
    if (sv->direct_ev) hipEventSynchronize(sv->direct_ev);  // a grid-mode query still running
    server_free(sv);
    delete sv;
}

extern "C" int ddp_server_replay(ddp_server *sv, int32_t n, const double *arrival_ms, const int32_t *type,
                                 const int32_t *s, const int32_t *t, const int32_t *upd_off, const ddp_update_t *upd,
                                 ddp_server_result_t *out)
{
    if (!sv || n < 0 || (n && (!arrival_ms || !type || !out))) return DDP_ERR_ARG;
    ddp_ctx *ctx = sv->ctx;
    std::vector<ddp_server_result_t> got(4096);
    std::vector<char> have(n, 0);
    int32_t received = 0;
    auto collect = [&]() -> int {
        int32_t k = 0;
        int rc = ddp_server_poll(sv, got.data(), (int32_t)got.size(), &k);
        for (int32_t i = 0; i < k; ++i) {
            const uint64_t idx = got[i].tag;
            if (idx < (uint64_t)n && !have[idx]) {
                out[idx] = got[i];
                have[idx] = 1;
                ++received;
            }
        }
        return rc;
    };
    const double start = ms_now(sv);
    for (int32_t i = 0; i < n; ++i) {
        while (ms_now(sv) - start < arrival_ms[i]) {
            int rc = collect();
            if (rc) return rc;
        }
        for (;;) {
            int rc = type[i] == DDP_REQ_UPDATE
                         ? ddp_server_submit_update(sv, upd + upd_off[i], upd_off[i + 1] - upd_off[i], (uint64_t)i)
                         : ddp_server_submit_query(sv, s[i], t[i], (uint64_t)i);
            if (rc == DDP_OK) break;
            if (rc != DDP_ERR_NOMEM) return rc;
            if ((rc = collect())) return rc;  // ring full: make room
        }
    }
    const auto t0 = clk::now();
    while (received < n) {
        int rc = collect();
        if (rc) return rc;
        if (std::chrono::duration<double>(clk::now() - t0).count() > 600)
            return ddp_set_err(ctx, DDP_ERR_TIMEOUT, "replay: %d of %d requests unfinished", n - received, n);
    }
    // times relative to the replay start; a request's submit time is its
    // scheduled arrival, so latencies include any wait before it entered the
    // ring (a full ring, a busy payload arena)
    for (int32_t i = 0; i < n; ++i) {
        out[i].submit_ms = arrival_ms[i];
        out[i].done_ms -= start;
    }
    return DDP_OK;
}
