// Context, graph upload, heuristics and small shared kernels.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <thread>

#include "device_common.h"

static thread_local std::string g_create_err;

int ddp_set_err(ddp_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else g_create_err = buf;
    return code;
}

int ddp_hip_err(ddp_ctx *ctx, hipError_t e, const char *what, const char *file, int line)
{
    const int code = (e == hipErrorOutOfMemory) ? DDP_ERR_NOMEM : DDP_ERR_HIP;
    return ddp_set_err(ctx, code, "HIP error %d (%s) in %s at %s:%d", (int)e,
                       hipGetErrorString(e), what, file, line);
}

extern "C" const char *ddp_create_error(void) { return g_create_err.c_str(); }

extern "C" const char *ddp_last_error(const ddp_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int ddp_ctx_create(int device, ddp_ctx **out)
{
    if (!out) return ddp_set_err(nullptr, DDP_ERR_ARG, "ddp_ctx_create: out is null");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return ddp_set_err(nullptr, DDP_ERR_NODEVICE, "no HIP device available (%s)",
                           hipGetErrorString(e));
    if (device < 0 || device >= n)
        return ddp_set_err(nullptr, DDP_ERR_ARG, "device %d out of range [0,%d)", device, n);
    ddp_ctx *c = new ddp_ctx();
    c->device = device;
    c->ws_ptr.assign(WS_COUNT_, nullptr);
    c->ws_size.assign(WS_COUNT_, 0);
    if ((e = hipSetDevice(device)) != hipSuccess ||
        (e = hipGetDeviceProperties(&c->prop, device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        ddp_set_err(nullptr, DDP_ERR_HIP, "context setup failed: %s", hipGetErrorString(e));
        delete c;
        return DDP_ERR_HIP;
    }
    if (std::strncmp(c->prop.gcnArchName, "gfx950", 6) != 0) {
        ddp_set_err(nullptr, DDP_ERR_NODEVICE, "device %d is %s; this library is built for gfx950",
                    device, c->prop.gcnArchName);
        hipStreamDestroy(c->stream);
        delete c;
        return DDP_ERR_NODEVICE;
    }
    *out = c;
    return DDP_OK;
}

extern "C" void ddp_ctx_destroy(ddp_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (void *p : ctx->ws_ptr)
        if (p) hipFree(p);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int ddp_ws(ddp_ctx *ctx, int slot, size_t bytes, void **out)
{
    if (bytes == 0) bytes = 256;
    if (ctx->ws_size[slot] < bytes) {
        if (ctx->ws_ptr[slot]) {
            DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            DDP_HIP(ctx, hipFree(ctx->ws_ptr[slot]));
            ctx->ws_ptr[slot] = nullptr;
            ctx->ws_size[slot] = 0;
        }
        void *p = nullptr;
        DDP_HIP(ctx, hipMalloc(&p, bytes));
        ctx->ws_ptr[slot] = p;
        ctx->ws_size[slot] = bytes;
    }
    *out = ctx->ws_ptr[slot];
    return DDP_OK;
}

extern "C" int ddp_device_info(ddp_ctx *ctx, ddp_device_info_t *info)
{
    if (!ctx || !info) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_device_info: null argument");
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    DDP_HIP(ctx, hipMemGetInfo(&fr, &tot));
    info->compute_units = ctx->prop.multiProcessorCount;
    info->lds_bytes_per_cu = (int32_t)ctx->prop.maxSharedMemoryPerMultiProcessor;
    info->wave_size = ctx->prop.warpSize;
    info->l2_bytes = ctx->prop.l2CacheSize;
    info->hbm_total_bytes = (int64_t)tot;
    info->hbm_free_bytes = (int64_t)fr;
    std::snprintf(info->arch, sizeof info->arch, "%s", ctx->prop.gcnArchName);
    return DDP_OK;
}

extern "C" int ddp_ctx_synchronize(ddp_ctx *ctx)
{
    if (!ctx) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_ctx_synchronize: null context");
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    DDP_HIP(ctx, hipDeviceSynchronize());
    return DDP_OK;
}

__global__ void copy_kernel(const int4 *__restrict__ a, int4 *__restrict__ b, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

extern "C" int ddp_copy_bandwidth(ddp_ctx *ctx, int64_t bytes, int32_t reps, float *ms, float *gbps)
{
    if (!ctx || bytes < 16 || reps < 1) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_copy_bandwidth: bad argument");
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t n = bytes / 16;
    int4 *a = nullptr, *b = nullptr;
    DDP_HIP(ctx, hipMalloc((void **)&a, (size_t)n * 16));
    hipError_t e = hipMalloc((void **)&b, (size_t)n * 16);
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, (size_t)n * 16, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b, 0, (size_t)n * 16, ctx->stream);
    float best = 1e30f;
    const int grid = ctx->prop.multiProcessorCount * 16;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        hipEventRecord(ctx->ev0, ctx->stream);
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, ctx->stream, a, b, n);
        hipEventRecord(ctx->ev1, ctx->stream);
        e = hipEventSynchronize(ctx->ev1);
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
        best = std::min(best, t);
    }
    hipFree(a);
    if (b) hipFree(b);
    if (e != hipSuccess) return ddp_hip_err(ctx, e, "copy kernel", __FILE__, __LINE__);
    if (ms) *ms = best;
    if (gbps) *gbps = (float)(2.0 * (double)n * 16 / (best * 1e-3) / 1e9);
    return DDP_OK;
}

// ------------------------------------------------------------------ graph

extern "C" void ddp_graph_destroy(ddp_graph *g)
{
    if (!g) return;
    hipSetDevice(g->device);  // hipFree waits for work that uses the memory
    void *ptrs[] = {g->off, g->col, g->w, g->roff, g->rcol, g->reid, g->rw, g->h, g->xy, g->d_iperm};
    for (void *p : ptrs)
        if (p) hipFree(p);
    delete g;
}

template <class T>
static int upload(ddp_ctx *ctx, T **dst, const T *src, size_t n)
{
    DDP_HIP(ctx, hipMalloc((void **)dst, std::max<size_t>(n, 1) * sizeof(T)));
    if (n) DDP_HIP(ctx, hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return DDP_OK;
}

// Position of (x, y) along the Hilbert curve through a 2^bits x 2^bits grid
// (the classic xy-to-distance walk: one quadrant digit per level, rotating
// the coordinates so the curve stays continuous).
static uint64_t hilbert_key(uint64_t x, uint64_t y, int bits)
{
    const uint64_t n = 1ull << bits;
    uint64_t d = 0;
    for (uint64_t sq = n >> 1; sq > 0; sq >>= 1) {
        const uint64_t rx = (x & sq) ? 1 : 0, ry = (y & sq) ? 1 : 0;
        d += sq * sq * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

// New order of the nodes (iperm[new] = old): along a Hilbert curve through
// the coordinates when there are any, else breadth-first over out-edges from
// node 0, then from the lowest unvisited id, and so on.  Either way nodes that
// are adjacent in the graph land on nearby ids, so a search's random reads of
// per-node words (g, parent, coordinates, rows) share cache lines.
static void relabel_order(int32_t N, const std::vector<int32_t> &o, const int32_t *col, const int32_t *xy,
                          std::vector<int32_t> &iperm)
{
    iperm.resize(N);
    if (xy) {
        int64_t mnx = INT64_MAX, mny = INT64_MAX, mxx = INT64_MIN, mxy = INT64_MIN;
        for (int32_t i = 0; i < N; ++i) {
            mnx = std::min<int64_t>(mnx, xy[2 * i]);
            mxx = std::max<int64_t>(mxx, xy[2 * i]);
            mny = std::min<int64_t>(mny, xy[2 * i + 1]);
            mxy = std::max<int64_t>(mxy, xy[2 * i + 1]);
        }
        const uint64_t span = (uint64_t)std::max(mxx - mnx, mxy - mny);
        int bits = 1;
        while (bits < 32 && (span >> bits) != 0) ++bits;
        std::vector<std::pair<uint64_t, int32_t>> k(N);
        for (int32_t i = 0; i < N; ++i)
            k[i] = {hilbert_key((uint64_t)(xy[2 * i] - mnx), (uint64_t)(xy[2 * i + 1] - mny), bits), i};
        std::sort(k.begin(), k.end());
        for (int32_t i = 0; i < N; ++i) iperm[i] = k[i].second;
        return;
    }
    std::vector<char> seen(N, 0);
    int32_t head = 0, tail = 0;
    for (int32_t r = 0; r < N; ++r) {
        if (seen[r]) continue;
        seen[r] = 1;
        iperm[tail++] = r;
        while (head < tail) {
            const int32_t u = iperm[head++];
            for (int32_t j = o[u]; j < o[u + 1]; ++j) {
                const int32_t c = col[j];
                if (c >= 0 && !seen[c]) {
                    seen[c] = 1;
                    iperm[tail++] = c;
                }
            }
        }
    }
}

extern "C" int ddp_graph_upload(ddp_ctx *ctx, int32_t N, int32_t E, const int32_t *off,
                                const int32_t *col, const uint32_t *w, uint32_t flags,
                                ddp_graph **out)
{
    return ddp_graph_upload_ex(ctx, N, E, off, col, w, nullptr, flags, out);
}

extern "C" int ddp_graph_upload_ex(ddp_ctx *ctx, int32_t N, int32_t E, const int32_t *off,
                                   const int32_t *col, const uint32_t *w, const int32_t *xy,
                                   uint32_t flags, ddp_graph **out)
{
    if (!ctx || !out) return ddp_set_err(ctx, DDP_ERR_ARG, "ddp_graph_upload: null argument");
    *out = nullptr;
    DDP_CHECK(N > 0 && E >= 0, ctx, DDP_ERR_ARG, "ddp_graph_upload: bad sizes N=%d E=%d", N, E);
    DDP_CHECK(off && (E == 0 || (col && w)), ctx, DDP_ERR_ARG, "ddp_graph_upload: null array");
    DDP_CHECK(off[0] == 0, ctx, DDP_ERR_FORMAT, "ddp_graph_upload: off[0] = %d, expected 0", off[0]);
    DDP_CHECK(!(flags & ~(DDP_GRAPH_REVERSE | DDP_GRAPH_REORDER)), ctx, DDP_ERR_ARG,
              "ddp_graph_upload: unknown flags 0x%x", flags);
    // validate offsets and ids; compute min weight and max degree
    std::vector<int32_t> o(N + 1);
    int32_t maxdeg = 0;
    for (int32_t i = 0; i < N; ++i) {
        o[i] = off[i];
        const int32_t end = (i + 1 < N) ? off[i + 1] : E;
        DDP_CHECK(off[i] <= end && end <= E, ctx, DDP_ERR_FORMAT,
                  "ddp_graph_upload: row %d has offsets [%d,%d) outside [0,%d]", i, off[i], end, E);
        maxdeg = std::max(maxdeg, end - off[i]);
    }
    o[N] = E;
    uint32_t minw = 0xFFFFFFFFu;
    uint64_t wsum = 0, nlive = 0;
    for (int32_t j = 0; j < E; ++j) {
        DDP_CHECK(col[j] < N, ctx, DDP_ERR_FORMAT, "ddp_graph_upload: col[%d] = %d >= N = %d", j,
                  col[j], N);
        if (col[j] >= 0) {
            minw = std::min(minw, w[j]);
            wsum += w[j];
            ++nlive;
        }
    }
    if (minw == 0xFFFFFFFFu) minw = 0;
    // relabel: rows move to their new ids with their edges in the same order
    // (so "first matching instance" rules are unaffected), columns are mapped
    std::vector<int32_t> perm, iperm, o2, col2, xy2;
    std::vector<uint32_t> w2;
    if (flags & DDP_GRAPH_REORDER) {
        relabel_order(N, o, col, xy, iperm);
        perm.resize(N);
        for (int32_t i = 0; i < N; ++i) perm[iperm[i]] = i;
        o2.resize((size_t)N + 1);
        col2.resize(E);
        w2.resize(E);
        int32_t pos = 0;
        for (int32_t i = 0; i < N; ++i) {
            o2[i] = pos;
            const int32_t u = iperm[i];
            for (int32_t j = o[u]; j < o[u + 1]; ++j, ++pos) {
                col2[pos] = col[j] >= 0 ? perm[col[j]] : col[j];
                w2[pos] = w[j];
            }
        }
        o2[N] = E;
        o.swap(o2);
        col = col2.data();
        w = w2.data();
    }
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    ddp_graph *g = new ddp_graph();
    g->ctx = ctx;
    g->device = ctx->device;
    g->N = N;
    g->E = E;
    g->minw = minw;
    g->meanw = nlive ? (uint32_t)std::min<uint64_t>(wsum / nlive, 0x7FFFFFFFull) : 0;
    g->maxdeg = maxdeg;
    g->perm.swap(perm);
    g->iperm.swap(iperm);
    int rc;
    if ((rc = upload(ctx, &g->off, o.data(), (size_t)N + 1)) || (rc = upload(ctx, &g->col, col, (size_t)E)) ||
        (rc = upload(ctx, &g->w, w, (size_t)E)) ||
        (!g->iperm.empty() && (rc = upload(ctx, &g->d_iperm, g->iperm.data(), (size_t)N)))) {
        ddp_graph_destroy(g);
        return rc;
    }
    if (flags & DDP_GRAPH_REVERSE) {
        // Stable transpose (DA_star/genOpCSR.cpp:55-74 order: predecessors by
        // ascending source id, parallel edges in forward order).
        std::vector<int32_t> ro(N + 1, 0);
        for (int32_t j = 0; j < E; ++j)
            if (col[j] >= 0) ro[col[j] + 1]++;
        for (int32_t i = 0; i < N; ++i) ro[i + 1] += ro[i];
        const int32_t RE = ro[N];
        std::vector<int32_t> rc_(RE), re_(RE);
        std::vector<uint32_t> rw_(RE);
        std::vector<int32_t> fill(ro.begin(), ro.end() - 1);
        for (int32_t u = 0; u < N; ++u)
            for (int32_t j = o[u]; j < o[u + 1]; ++j) {
                const int32_t c = col[j];
                if (c < 0) continue;
                const int32_t p = fill[c]++;
                rc_[p] = u;
                rw_[p] = w[j];
                re_[p] = j;
            }
        g->RE = RE;
        if ((rc = upload(ctx, &g->roff, ro.data(), (size_t)N + 1)) ||
            (rc = upload(ctx, &g->rcol, rc_.data(), (size_t)RE)) ||
            (rc = upload(ctx, &g->rw, rw_.data(), (size_t)RE)) ||
            (rc = upload(ctx, &g->reid, re_.data(), (size_t)RE))) {
            ddp_graph_destroy(g);
            return rc;
        }
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ddp_graph_destroy(g);
        return ddp_hip_err(ctx, e, "graph upload", __FILE__, __LINE__);
    }
    if (xy && (rc = ddp_graph_set_coords(g, xy))) {
        ddp_graph_destroy(g);
        return rc;
    }
    *out = g;
    return DDP_OK;
}

extern "C" int ddp_graph_info(const ddp_graph *g, int32_t *N, int32_t *E, uint32_t *min_weight,
                              int32_t *max_out_degree)
{
    if (!g) return DDP_ERR_ARG;
    if (N) *N = g->N;
    if (E) *E = g->E;
    if (min_weight) *min_weight = g->minw;
    if (max_out_degree) *max_out_degree = g->maxdeg;
    return DDP_OK;
}

extern "C" int ddp_graph_set_h(ddp_graph *g, const int32_t *h)
{
    if (!g || !h) return ddp_set_err(g ? g->ctx : nullptr, DDP_ERR_ARG, "ddp_graph_set_h: null argument");
    ddp_ctx *ctx = g->ctx;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<int32_t> hp;
    if (!g->iperm.empty()) {  // caller order -> device order
        hp.resize(g->N);
        for (int32_t i = 0; i < g->N; ++i) hp[i] = h[g->iperm[i]];
        h = hp.data();
    }
    if (!g->h) DDP_HIP(ctx, hipMalloc((void **)&g->h, (size_t)g->N * 4));
    g->h_target = -1;  // a caller's array: its target is not known here
    DDP_HIP(ctx, hipMemcpyAsync(g->h, h, (size_t)g->N * 4, hipMemcpyHostToDevice, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DDP_OK;
}

extern "C" int ddp_graph_set_coords(ddp_graph *g, const int32_t *xy)
{
    if (!g || !xy) return ddp_set_err(g ? g->ctx : nullptr, DDP_ERR_ARG, "ddp_graph_set_coords: null argument");
    ddp_ctx *ctx = g->ctx;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<int32_t> xp;
    if (!g->iperm.empty()) {  // caller order -> device order
        xp.resize(2 * (size_t)g->N);
        for (int32_t i = 0; i < g->N; ++i) {
            xp[2 * (size_t)i] = xy[2 * (size_t)g->iperm[i]];
            xp[2 * (size_t)i + 1] = xy[2 * (size_t)g->iperm[i] + 1];
        }
        xy = xp.data();
    }
    if (!g->xy) DDP_HIP(ctx, hipMalloc((void **)&g->xy, (size_t)g->N * 8));
    DDP_HIP(ctx, hipMemcpyAsync(g->xy, xy, (size_t)g->N * 8, hipMemcpyHostToDevice, ctx->stream));
    g->host_xy.assign(xy, xy + 2 * (size_t)g->N);
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DDP_OK;
}

extern "C" int ddp_graph_set_grid(ddp_graph *g, int32_t width)
{
    if (!g) return DDP_ERR_ARG;
    DDP_CHECK(width > 0 && g->N % width == 0, g->ctx, DDP_ERR_ARG,
              "ddp_graph_set_grid: width %d does not divide N = %d", width, g->N);
    DDP_CHECK(g->perm.empty(), g->ctx, DDP_ERR_STATE,
              "ddp_graph_set_grid: the grid heuristic reads row and column from the node id, which "
              "DDP_GRAPH_REORDER has changed");
    g->grid_w = width;
    return DDP_OK;
}

GraphView ddp_view(const ddp_graph *g)
{
    GraphView v{};
    v.off = g->off;
    v.col = g->col;
    v.w = g->w;
    v.N = g->N;
    return v;
}

int ddp_make_heur(const ddp_graph *g, ddp_heur_t h, int32_t t, HeurView *out)
{
    ddp_ctx *ctx = g->ctx;
    HeurView v{};
    v.kind = h.kind;
    v.scale = h.scale > 0 ? h.scale : 1;
    switch (h.kind) {
    case DDP_H_ZERO:
        break;
    case DDP_H_ARRAY:
        DDP_CHECK(g->h, ctx, DDP_ERR_STATE, "DDP_H_ARRAY requested but no heuristic set (ddp_graph_set_h / ddp_bfs_hops)");
        v.h = g->h;
        break;
    case DDP_H_MANHATTAN:
        DDP_CHECK(g->grid_w > 0, ctx, DDP_ERR_STATE, "DDP_H_MANHATTAN requested but no grid set (ddp_graph_set_grid)");
        v.width = g->grid_w;
        if (t >= 0) {
            v.ty = t / g->grid_w;
            v.tx = t % g->grid_w;
        }
        break;
    case DDP_H_EUCLID:
        DDP_CHECK(g->xy, ctx, DDP_ERR_STATE, "DDP_H_EUCLID requested but no coordinates set (ddp_graph_set_coords)");
        v.xy = g->xy;
        if (t >= 0) {
            v.tx = g->host_xy[2 * (size_t)t];
            v.ty = g->host_xy[2 * (size_t)t + 1];
        }
        break;
    default:
        return ddp_set_err(ctx, DDP_ERR_ARG, "unknown heuristic kind %d", h.kind);
    }
    *out = v;
    return DDP_OK;
}

int ddp_heur_at(const ddp_graph *g, const HeurView &hv, int32_t v, int32_t *out)
{
    ddp_ctx *ctx = g->ctx;
    switch (hv.kind) {
    case DDP_H_ZERO:
        *out = 0;
        return DDP_OK;
    case DDP_H_ARRAY: {
        int32_t x = 0;
        DDP_HIP(ctx, hipMemcpyAsync(&x, g->h + v, 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *out = x;
        return DDP_OK;
    }
    case DDP_H_MANHATTAN: {
        int32_t r = v / hv.width, c = v % hv.width;
        *out = (std::abs(r - hv.ty) + std::abs(c - hv.tx)) * hv.scale;
        return DDP_OK;
    }
    default: {
        int64_t dx = (int64_t)g->host_xy[2 * (size_t)v] - hv.tx;
        int64_t dy = (int64_t)g->host_xy[2 * (size_t)v + 1] - hv.ty;
        *out = (int32_t)isqrt_sum_sq(dx, dy) * hv.scale;
        return DDP_OK;
    }
    }
}

// ---------------------------------------------------------- path extraction

// One lane walks parent pointers t -> s (bounded by N steps), writing t..s.
__global__ void walk_path_kernel(const uint64_t *gp, int32_t N, int32_t s, int32_t t,
                                 int32_t *buf, int32_t cap, int32_t *len)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int32_t v = t, n = 0;
    for (int32_t step = 0; step <= N; ++step) {
        if (n < cap) buf[n] = v;
        ++n;
        if (v == s) break;
        const uint32_t p = (uint32_t)gp[v];
        if (p == 0xFFFFFFFFu || (int32_t)p >= N) { n = -1; break; }
        v = (int32_t)p;
    }
    if (n > 0 && v != s) n = -1;  // no chain back to s
    *len = n;
}

int ddp_extract_path(ddp_ctx *ctx, const uint64_t *gp_dev, int32_t N, int32_t s, int32_t t,
                     int32_t *path_out, int32_t cap, int32_t *len)
{
    void *ws = nullptr;
    const int32_t dcap = std::min<int32_t>(cap, N);
    int rc = ddp_ws(ctx, WS_PATH, ((size_t)dcap + 64) * 4, &ws);
    if (rc) return rc;
    int32_t *dbuf = (int32_t *)ws + 64, *dlen = (int32_t *)ws;
    hipLaunchKernelGGL(walk_path_kernel, dim3(1), dim3(64), 0, ctx->stream, gp_dev, N, s, t, dbuf,
                       dcap, dlen);
    DDP_HIP(ctx, hipGetLastError());
    int32_t n = 0;
    DDP_HIP(ctx, hipMemcpyAsync(&n, dlen, 4, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DDP_CHECK(n >= 0, ctx, DDP_ERR_STATE, "parent chain from %d does not reach %d", t, s);
    if (n > cap) {
        *len = -n;
        return DDP_OK;
    }
    std::vector<int32_t> tmp(n);
    if (n) DDP_HIP(ctx, hipMemcpyAsync(tmp.data(), dbuf, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int32_t i = 0; i < n; ++i) path_out[i] = tmp[n - 1 - i];
    *len = n;
    return DDP_OK;
}
