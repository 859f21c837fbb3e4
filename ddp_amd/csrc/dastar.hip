// Dynamic A* (DA*): repair one query's search state after a batch of edge
// inserts and deletes, then resume the GA* grid kernel from its retained
// queues.
//
// Invariant kept between batches (the one A* itself maintains): every finite
// g is the cost of a real path, and every node that is not in a queue has
// relaxed all of its live out-edges with its current g.  A batch restores it:
//   1. deletes (matched on the host, first live instance of u->v as in
//      DA_star/GA_merge.cu:1019-1037) become tombstones in HBM;
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
// The graph is base rows with tombstones plus delta rows holding the inserts
// in arrival order, so the row of u reads exactly as GA_merge.cu's merged row.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/ddp_dastar.h"
#include "gastar.h"
#include "search_expand.h"

namespace {

using clk = std::chrono::steady_clock;
float ms_since(clk::time_point t0) { return std::chrono::duration<float, std::milli>(clk::now() - t0).count(); }

enum : int { RC_NEXT = 0, RC_MARKED = 1, RC_ROOTS = 2, RC_REPULLED = 3, RC_PUSHES = 4, RC_WORDS = 16 };

__global__ void scatter_tombstones(int32_t *col, const int32_t *pos, int32_t n)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) col[pos[i]] = -1;
}

// Roots: deleted u->v where v's parent is u.  Claims v for this epoch.
__global__ void seed_roots(const int32_t *pairs, int32_t n, const uint64_t *gp, int32_t *mark, int32_t epoch,
                           int32_t *front, int32_t *marked, uint32_t *rctrl)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t u = pairs[2 * i], v = pairs[2 * i + 1];
    const uint64_t x = ld_gp(gp + v);
    if ((uint32_t)(x >> 32) == DDP_INF_F || (uint32_t)x != (uint32_t)u) return;
    if (atomicMax(mark + v, epoch) == epoch) return;
    const uint32_t k = atomicAdd(rctrl + RC_NEXT, 1u);
    front[k] = v;
    marked[atomicAdd(rctrl + RC_MARKED, 1u)] = v;
    atomicAdd(rctrl + RC_ROOTS, 1u);
}

// One level of the parent-subtree walk: children y of frontier node x are the
// live out-neighbours whose parent is x.
__global__ void subtree_level(GraphView G, const uint64_t *gp, int32_t *mark, int32_t epoch, const int32_t *cur,
                              int32_t ncur, int32_t *nxt, int32_t *marked, uint32_t *rctrl)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncur) return;
    const int32_t x = cur[i];
    for (int pass = 0; pass < 2; ++pass) {
        const int32_t *off = pass ? G.doff : G.off;
        const int32_t *col = pass ? G.dcol : G.col;
        if (!off) continue;
        for (int32_t j = off[x]; j < off[x + 1]; ++j) {
            const int32_t y = col[j];
            if (y < 0) continue;
            if ((uint32_t)ld_gp(gp + y) != (uint32_t)x) continue;
            if (atomicMax(mark + y, epoch) == epoch) continue;
            nxt[atomicAdd(rctrl + RC_NEXT, 1u)] = y;
            marked[atomicAdd(rctrl + RC_MARKED, 1u)] = y;
        }
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
__global__ void repull(GraphView R, const int32_t *h, uint64_t *gp, const int32_t *mark, int32_t epoch,
                       const int32_t *marked, int32_t n, RepairPush P)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t x = -1;
    uint32_t best = 0xFFFFFFFFu, bp = 0xFFFFFFFFu;
    if (i < n) {
        x = marked[i];
        const uint32_t hx = (uint32_t)h[x];
        for (int pass = 0; pass < 2; ++pass) {
            const int32_t *off = pass ? R.doff : R.off;
            const int32_t *col = pass ? R.dcol : R.col;
            const uint32_t *w = pass ? R.dw : R.w;
            if (!off) continue;
            for (int32_t j = off[x]; j < off[x + 1]; ++j) {
                const int32_t p = col[j];
                if (p < 0 || p == x || mark[p] == epoch) continue;
                const uint32_t fp = (uint32_t)(ld_gp(gp + p) >> 32);
                if (fp == DDP_INF_F) continue;
                const uint32_t cand = (fp - (uint32_t)h[p]) + w[j] + hx;  // unsigned, as GA_v2.cu:137
                if (cand < DDP_INF_F && (cand < best || (cand == best && (uint32_t)p < bp))) {
                    best = cand;
                    bp = (uint32_t)p;
                }
            }
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
    int32_t N = 0, s = 0, t = 0;
    // host copy of the current graph version
    std::vector<int32_t> off, col, roff, rcol, reid;  // off/roff have N+1 entries
    std::vector<uint32_t> w, rw;
    struct DEdge {
        int32_t u, v;
        uint32_t w;
    };
    std::vector<DEdge> delta;  // live inserts in arrival order
    std::vector<int32_t> doff, dcol, didx, droff, drcol;
    std::vector<uint32_t> dw, drw;
    // device
    ddp_graph *g = nullptr;  // base rows (tombstones scattered in), transposed rows too
    int32_t *d_h = nullptr;
    int32_t *d_doff = nullptr, *d_dcol = nullptr, *d_droff = nullptr, *d_drcol = nullptr;
    uint32_t *d_dw = nullptr, *d_drw = nullptr;
    size_t c_doff = 0, c_dcol = 0, c_droff = 0, c_drcol = 0, c_dw = 0, c_drw = 0;
    int32_t *d_mark = nullptr, *d_front = nullptr, *d_marked = nullptr;
    uint32_t *d_rctrl = nullptr;
    int32_t *d_work = nullptr;  // tombstone positions / pairs / inserts
    size_t c_work = 0;
    int32_t epoch = 0;
    // search
    GaGridPlan P{};
    GaGridState S{};
    HeurView hv{};
    int32_t rr = 0;
    uint32_t F = DDP_INF_F;
};

static void dastar_free(ddp_dastar *d)
{
    ddp_ga_free(&d->S);
    void *ptrs[] = {d->d_h, d->d_doff, d->d_dcol, d->d_droff, d->d_drcol, d->d_dw, d->d_drw,
                    d->d_mark, d->d_front, d->d_marked, d->d_rctrl, d->d_work};
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

static GraphView fwd_view(const ddp_dastar *d)
{
    GraphView v = ddp_view(d->g);
    if (!d->delta.empty()) {
        v.doff = d->d_doff;
        v.dcol = d->d_dcol;
        v.dw = d->d_dw;
    }
    return v;
}

static GraphView rev_view(const ddp_dastar *d)
{
    GraphView v{};
    v.off = d->g->roff;
    v.col = d->g->rcol;
    v.w = d->g->rw;
    v.N = d->N;
    if (!d->delta.empty()) {
        v.doff = d->d_droff;
        v.dcol = d->d_drcol;
        v.dw = d->d_drw;
    }
    return v;
}

// Rebuilds the delta rows (forward: grouped by source, arrival order kept;
// reverse: grouped by target) from the live inserts and uploads them.
static int rebuild_delta(ddp_dastar *d)
{
    const int32_t N = d->N;
    const size_t D = d->delta.size();
    d->doff.assign(N + 1, 0);
    d->droff.assign(N + 1, 0);
    for (const auto &e : d->delta) {
        d->doff[e.u + 1]++;
        d->droff[e.v + 1]++;
    }
    for (int32_t i = 0; i < N; ++i) {
        d->doff[i + 1] += d->doff[i];
        d->droff[i + 1] += d->droff[i];
    }
    d->dcol.assign(D, 0);
    d->dw.assign(D, 0);
    d->didx.assign(D, 0);
    d->drcol.assign(D, 0);
    d->drw.assign(D, 0);
    std::vector<int32_t> f(d->doff.begin(), d->doff.end() - 1), r(d->droff.begin(), d->droff.end() - 1);
    for (size_t k = 0; k < D; ++k) {
        const auto &e = d->delta[k];
        const int32_t p = f[e.u]++;
        d->dcol[p] = e.v;
        d->dw[p] = e.w;
        d->didx[p] = (int32_t)k;
        const int32_t q = r[e.v]++;
        d->drcol[q] = e.u;
        d->drw[q] = e.w;
    }
    ddp_ctx *ctx = d->ctx;
    int rc;
    if ((rc = dev_upload(ctx, &d->d_doff, &d->c_doff, d->doff.data(), (size_t)N + 1)) ||
        (rc = dev_upload(ctx, &d->d_dcol, &d->c_dcol, d->dcol.data(), D)) ||
        (rc = dev_upload(ctx, &d->d_dw, &d->c_dw, d->dw.data(), D)) ||
        (rc = dev_upload(ctx, &d->d_droff, &d->c_droff, d->droff.data(), (size_t)N + 1)) ||
        (rc = dev_upload(ctx, &d->d_drcol, &d->c_drcol, d->drcol.data(), D)) ||
        (rc = dev_upload(ctx, &d->d_drw, &d->c_drw, d->drw.data(), D)))
        return rc;
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
    // host copy + stable transpose (DA_star/genOpCSR.cpp order) with forward edge ids
    d->off.assign(off, off + N);
    d->off.push_back(E);
    d->col.assign(col, col + E);
    d->w.assign(w, w + E);
    d->roff.assign(N + 1, 0);
    for (int32_t j = 0; j < E; ++j)
        if (col[j] >= 0) d->roff[col[j] + 1]++;
    for (int32_t i = 0; i < N; ++i) d->roff[i + 1] += d->roff[i];
    const int32_t RE = d->roff[N];
    d->rcol.assign(RE, 0);
    d->rw.assign(RE, 0);
    d->reid.assign(RE, 0);
    {
        std::vector<int32_t> f(d->roff.begin(), d->roff.end() - 1);
        for (int32_t u = 0; u < N; ++u)
            for (int32_t j = d->off[u]; j < d->off[u + 1]; ++j) {
                if (col[j] < 0) continue;
                const int32_t p = f[col[j]]++;
                d->rcol[p] = u;
                d->rw[p] = w[j];
                d->reid[p] = j;
            }
    }
    hipError_t e = hipSuccess;
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(dst, std::max<size_t>(bytes, 16));
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    };
    up((void **)&d->g->roff, d->roff.data(), (size_t)(N + 1) * 4);
    up((void **)&d->g->rcol, d->rcol.data(), (size_t)RE * 4);
    up((void **)&d->g->rw, d->rw.data(), (size_t)RE * 4);
    d->g->RE = RE;
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
        (rc = ddp_ga_alloc_owned(ctx, N, d->P, &d->S))) {
        dastar_free(d);
        delete d;
        return rc;
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
    // ---- 1. host: match deletes (first live instance, base row then inserts)
    std::vector<int32_t> ftomb, rtomb, pairs, ins;
    std::vector<char> dead_delta(d->delta.size(), 0);
    for (int32_t i = 0; i < n; ++i) {
        const ddp_update_t &l = lines[i];
        if (l.flag == 1) {
            ++S.inserts;
            continue;
        }
        ++S.deletes;
        bool found = false;
        for (int32_t j = d->off[l.u]; j < d->off[l.u + 1] && !found; ++j)
            if (d->col[j] == l.v) {
                d->col[j] = -1;
                ftomb.push_back(j);
                for (int32_t r = d->roff[l.v]; r < d->roff[l.v + 1]; ++r)
                    if (d->reid[r] == j) {
                        d->rcol[r] = -1;
                        rtomb.push_back(r);
                        break;
                    }
                found = true;
            }
        if (!found && !d->doff.empty())
            for (int32_t p = d->doff[l.u]; p < d->doff[l.u + 1] && !found; ++p) {
                const int32_t k = d->didx[p];
                if (!dead_delta[k] && d->delta[k].v == l.v) {
                    dead_delta[k] = 1;
                    found = true;
                }
            }
        if (found) {
            ++S.deletes_effective;
            pairs.push_back(l.u);
            pairs.push_back(l.v);
        }
    }
    // ---- 2. host: drop deleted inserts, append the new ones, rebuild delta rows
    {
        std::vector<ddp_dastar::DEdge> nd;
        nd.reserve(d->delta.size() + S.inserts);
        for (size_t k = 0; k < d->delta.size(); ++k)
            if (!dead_delta[k]) nd.push_back(d->delta[k]);
        for (int32_t i = 0; i < n; ++i)
            if (lines[i].flag == 1) {
                nd.push_back({lines[i].u, lines[i].v, lines[i].w});
                ins.push_back(lines[i].u);
                ins.push_back(lines[i].v);
                ins.push_back((int32_t)lines[i].w);
            }
        d->delta.swap(nd);
    }
    int rc = rebuild_delta(d);
    if (rc) return rc;
    // device work list: [ftomb | rtomb | pairs | ins]
    std::vector<int32_t> work;
    work.reserve(ftomb.size() + rtomb.size() + pairs.size() + ins.size());
    work.insert(work.end(), ftomb.begin(), ftomb.end());
    work.insert(work.end(), rtomb.begin(), rtomb.end());
    work.insert(work.end(), pairs.begin(), pairs.end());
    work.insert(work.end(), ins.begin(), ins.end());
    if ((rc = dev_upload(ctx, &d->d_work, &d->c_work, work.data(), work.size()))) return rc;
    const int32_t *w_ft = d->d_work, *w_rt = w_ft + ftomb.size(), *w_pairs = w_rt + rtomb.size(),
                  *w_ins = w_pairs + pairs.size();
    const int T = 256;
    auto grid = [&](size_t k) { return dim3((unsigned)std::max<size_t>(1, (k + T - 1) / T)); };
    if (!ftomb.empty())
        hipLaunchKernelGGL(scatter_tombstones, grid(ftomb.size()), dim3(T), 0, ctx->stream, d->g->col, w_ft,
                           (int32_t)ftomb.size());
    if (!rtomb.empty())
        hipLaunchKernelGGL(scatter_tombstones, grid(rtomb.size()), dim3(T), 0, ctx->stream, d->g->rcol, w_rt,
                           (int32_t)rtomb.size());
    DDP_HIP(ctx, hipGetLastError());
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    S.host_ms = ms_since(t0);
    int32_t iters = 0;
    float kms = 0;
    if (mode == DDP_DASTAR_RECOMPUTE) {
        if ((rc = fresh_search(d, &iters, &kms))) return rc;
    } else {
        DDP_CHECK(mode == DDP_DASTAR_REPAIR, ctx, DDP_ERR_ARG, "unknown DA* mode %d", mode);
        const auto t1 = clk::now();
        // ---- 3. invalidate parent subtrees of deleted tree edges
        d->epoch += 1;
        uint32_t rctrl[RC_WORDS] = {0};
        DDP_HIP(ctx, hipMemsetAsync(d->d_rctrl, 0, RC_WORDS * 4, ctx->stream));
        if ((rc = ddp_ga_prepare_resume(ctx, d->S))) return rc;
        const GraphView F = fwd_view(d), R = rev_view(d);
        int32_t *cur = d->d_front, *nxt = d->d_front + N;
        const int32_t npairs = (int32_t)(pairs.size() / 2);
        if (npairs)
            hipLaunchKernelGGL(seed_roots, grid(npairs), dim3(T), 0, ctx->stream, w_pairs, npairs, d->S.gp, d->d_mark,
                               d->epoch, cur, d->d_marked, d->d_rctrl);
        DDP_HIP(ctx, hipMemcpyAsync(rctrl, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int32_t ncur = (int32_t)rctrl[RC_NEXT];
        S.roots = (int32_t)rctrl[RC_ROOTS];
        while (ncur > 0) {
            DDP_HIP(ctx, hipMemsetAsync(d->d_rctrl + RC_NEXT, 0, 4, ctx->stream));
            hipLaunchKernelGGL(subtree_level, grid(ncur), dim3(T), 0, ctx->stream, F, d->S.gp, d->d_mark, d->epoch,
                               cur, ncur, nxt, d->d_marked, d->d_rctrl);
            DDP_HIP(ctx, hipMemcpyAsync(rctrl, d->d_rctrl, RC_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
            DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            ncur = (int32_t)rctrl[RC_NEXT];
            std::swap(cur, nxt);
        }
        S.invalidated = (int32_t)rctrl[RC_MARKED];
        // ---- 4. re-pull invalidated nodes, 5. relax inserted edges
        RepairPush P{d->S.pbuf + d->P.pcap, d->S.ctrl + 2 * GA_SLOT_WORDS, d->P.pcap, d->d_rctrl, d->t};
        if (S.invalidated)
            hipLaunchKernelGGL(repull, grid(S.invalidated), dim3(T), 0, ctx->stream, R, d->d_h, d->S.gp, d->d_mark,
                               d->epoch, d->d_marked, S.invalidated, P);
        const int32_t nins = (int32_t)(ins.size() / 3);
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
    S.total_ms = ms_since(t0);
    *cost = (int32_t)d->F;
    if (st) *st = S;
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
    int64_t live = (int64_t)d->delta.size();
    for (int32_t c : d->col) live += c >= 0;
    *E = (int32_t)live;
    return DDP_OK;
}

extern "C" int ddp_dastar_export(const ddp_dastar *d, int32_t *off, int32_t *col, uint32_t *w)
{
    if (!d || !off || !col || !w) return DDP_ERR_ARG;
    int32_t k = 0;
    for (int32_t u = 0; u < d->N; ++u) {
        off[u] = k;
        for (int32_t j = d->off[u]; j < d->off[u + 1]; ++j)
            if (d->col[j] >= 0) {
                col[k] = d->col[j];
                w[k] = d->w[j];
                ++k;
            }
        if (!d->doff.empty())
            for (int32_t p = d->doff[u]; p < d->doff[u + 1]; ++p) {
                col[k] = d->dcol[p];
                w[k] = d->dw[p];
                ++k;
            }
    }
    return DDP_OK;
}
