// Reverse-BFS heuristic (bfs_hx).
//
// Reference (GA_star/bfs_hx.cu:11-36, 111-138): one thread per frontier node
// walks its row serially; duplicates are appended when two threads see
// dist = -1 at once; the host launches one kernel per level and copies the
// next frontier size back each time.  Here one persistent cooperative kernel
// runs all levels: each workgroup takes B frontier nodes at a time and walks
// their rows cooperatively (prefix sum of degrees, B edges per step), a node
// joins the next frontier only if its CAS from -1 succeeds (no duplicates),
// and appends are aggregated per wave (ballot + one atomic).  Levels are
// separated by a grid barrier; the frontier count of level L lives in one of
// three rotating words so no extra barrier is needed to reset it.
#include <algorithm>
#include <mutex>

#include "device_common.h"

namespace {

enum : int { BC_CNT = 0, BC_LEVELS = 8, BC_ERR = 9, BC_BAR = 32, BC_WORDS = 64 };

struct BfsArgs {
    const int32_t *off, *col;
    int32_t N;
    int32_t *dist;
    int32_t *front;  // 2 * N
    uint32_t *ctrl;
    uint64_t barrier_ticks;
};

template <int B>
__global__ void __launch_bounds__(B) bfs_kernel(BfsArgs a)
{
    __shared__ int32_t s_incl[B], s_rowb[B];
    __shared__ int32_t s_tmp[16];
    __shared__ int32_t s_flag;
    const int lane = threadIdx.x;
    uint32_t *ctrl = a.ctrl;
    uint32_t gen = ld_u32(ctrl + BC_BAR + 1);
    int32_t level = 0;
    for (;; ++level) {
        const uint32_t n = ld_u32(ctrl + BC_CNT + (level % 3));
        if (n == 0) break;
        if (blockIdx.x == 0 && lane == 0) st_u32(ctrl + BC_CNT + ((level + 2) % 3), 0);
        const int32_t *cur = a.front + (int64_t)(level & 1) * a.N;
        int32_t *nxt = a.front + (int64_t)((level + 1) & 1) * a.N;
        uint32_t *ncnt = ctrl + BC_CNT + ((level + 1) % 3);
        const int32_t nl = level + 1;
        for (int64_t base = (int64_t)blockIdx.x * B; base < n; base += (int64_t)gridDim.x * B) {
            const int64_t i = base + lane;
            int32_t deg = 0, b = 0;
            if (i < n) {
                const int32_t v = __hip_atomic_load(cur + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                b = a.off[v];
                deg = a.off[v + 1] - b;
            }
            int32_t total;
            const int32_t incl = block_incl_scan<B>(deg, s_tmp, total);
            s_incl[lane] = incl;
            s_rowb[lane] = b;
            __syncthreads();
            for (int32_t eb = 0; eb < total; eb += B) {
                const int32_t e = eb + lane;
                bool add = false;
                int32_t c = -1;
                if (e < total) {
                    int32_t lo = 0, hi = B - 1;
                    while (lo < hi) {
                        const int32_t mid = (lo + hi) >> 1;
                        if (s_incl[mid] > e) hi = mid;
                        else lo = mid + 1;
                    }
                    const int32_t k = e - (lo ? s_incl[lo - 1] : 0);
                    c = a.col[s_rowb[lo] + k];
                    if (c >= 0 && a.dist[c] < 0) add = atomicCAS(a.dist + c, -1, nl) == -1;
                }
                const uint64_t mask = __ballot(add);
                if (mask) {
                    const int wl = lane & 63;
                    const int leader = __builtin_ctzll(mask);
                    uint32_t pos = 0;
                    if (wl == leader) pos = atomicAdd(ncnt, (uint32_t)__popcll(mask));
                    pos = __shfl(pos, leader, 64);
                    if (add) nxt[pos + lane_rank(mask)] = c;
                }
            }
            __syncthreads();
        }
        if (!grid_barrier(ctrl + BC_BAR, gridDim.x, gen, a.barrier_ticks, &s_flag)) {
            atomicOr(ctrl + BC_ERR, 1u);
            break;
        }
    }
    if (blockIdx.x == 0 && lane == 0) ctrl[BC_LEVELS] = (uint32_t)level;
}

__global__ void bfs_init(int32_t *dist, int32_t N, int32_t *front, uint32_t *ctrl, int32_t t)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) dist[i] = -1;
    if (blockIdx.x == 0 && threadIdx.x < BC_WORDS) ctrl[threadIdx.x] = 0;
}

__global__ void bfs_seed(int32_t *dist, int32_t *front, uint32_t *ctrl, int32_t t)
{
    if (threadIdx.x == 0) {
        dist[t] = 0;
        front[0] = t;
        ctrl[BC_CNT + 0] = 1;
    }
}

__global__ void bfs_scale(const int32_t *dist, int32_t *out, int32_t N, int32_t scale)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int32_t d = dist[i];
        out[i] = d < 0 ? -1 : d * scale;
    }
}

// bfs_scale into the caller's numbering: out[iperm[i]] = hops(i) * scale.
__global__ void bfs_scale_perm(const int32_t *dist, const int32_t *iperm, int32_t *out, int32_t N, int32_t scale)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int32_t d = dist[i];
        out[iperm[i]] = d < 0 ? -1 : d * scale;
    }
}

constexpr int kBfsBlock = 256;

// Runs the BFS from device node t into dist (N entries) with front (2N
// entries) and ctrl (BC_WORDS words) as scratch; waits for it and returns the
// level count.  Allocates nothing.
int bfs_run(ddp_ctx *ctx, const ddp_graph *g, int32_t t, int32_t reverse, int32_t *dist, int32_t *front,
            uint32_t *ctrl, int32_t *levels)
{
    constexpr int B = kBfsBlock;
    const int32_t N = g->N;
    int per_cu = 0;
    DDP_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)bfs_kernel<B>, B, 0));
    DDP_CHECK(per_cu > 0, ctx, DDP_ERR_HIP, "bfs kernel does not fit on a CU");
    const int32_t need = (N + B - 1) / B;
    const int32_t grid = std::max(1, std::min(need, std::min(per_cu, 2) * ctx->prop.multiProcessorCount));
    // the control words are cleared by a kernel, not a memset node
    hipLaunchKernelGGL(bfs_init, dim3(1024), dim3(256), 0, ctx->stream, dist, N, front, ctrl, t);
    hipLaunchKernelGGL(bfs_seed, dim3(1), dim3(64), 0, ctx->stream, dist, front, ctrl, t);
    BfsArgs a{reverse ? g->roff : g->off, reverse ? g->rcol : g->col, N, dist, front, ctrl,
              100ull * 1000 * 1000 * 20};
    void *args[] = {&a};
    DDP_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    // Ordinary launch: the grid barrier is software and only needs every
    // workgroup resident at once, which the occupancy-sized grid gives on an
    // idle device (a cooperative launch would also reserve the GPU's global
    // wave-sync unit, which only one process may hold).  Every barrier wait is
    // bounded, so a non-resident grid ends with DDP_ERR_TIMEOUT, not a hang.
    DDP_HIP(ctx, hipLaunchKernel((const void *)bfs_kernel<B>, dim3(grid), dim3(B), args, 0, ctx->stream));
    DDP_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t res[2];
    DDP_HIP(ctx, hipMemcpyAsync(res, ctrl + BC_LEVELS, 8, hipMemcpyDeviceToHost, ctx->stream));
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DDP_CHECK(res[1] == 0, ctx, DDP_ERR_TIMEOUT, "bfs: grid barrier wait timed out");
    *levels = (int32_t)res[0];
    return DDP_OK;
}

int bfs_check(ddp_ctx *ctx, const ddp_graph *g, int32_t t, int32_t reverse, int32_t scale, const char *who)
{
    if (!ctx || !g) return ddp_set_err(ctx, DDP_ERR_ARG, "%s: null argument", who);
    DDP_CHECK(t >= 0 && t < g->N, ctx, DDP_ERR_ARG, "%s: t=%d outside [0,%d)", who, t, g->N);
    DDP_CHECK(!reverse || g->roff, ctx, DDP_ERR_STATE, "%s: reverse=1 needs a graph uploaded with DDP_GRAPH_REVERSE",
              who);
    DDP_CHECK(scale >= 1, ctx, DDP_ERR_ARG, "%s: scale must be >= 1", who);
    return DDP_OK;
}

// ------------------------------------------------- allocation-free variant
//
// Device memory of ddp_bfs_hops_noalloc: the distance array, the two frontier
// buffers and the control words of one BFS over at most kBfsScratchNodes
// nodes, 12 bytes per node (24 MiB + 256 bytes).  Every GPU the library runs
// on holds its own copy until the process ends.  In host code the name is not
// a device address: kernels get scratch_dev(device), the address
// hipGetSymbolAddress returns for that device.
constexpr int32_t kBfsScratchNodes = 1 << 21;
static __device__ __align__(256) int32_t scratch[3 * (size_t)kBfsScratchNodes + BC_WORDS];

constexpr int kMaxDevices = 64;
static std::mutex s_scratch_mu;
static int32_t *s_scratch_ptr[kMaxDevices];

// The address of this device's copy of scratch, looked up once per device
// with that device current.
static int scratch_dev(ddp_ctx *ctx, int32_t **out)
{
    DDP_CHECK(ctx->device >= 0 && ctx->device < kMaxDevices, ctx, DDP_ERR_ARG, "device %d out of range",
              ctx->device);
    std::lock_guard<std::mutex> lock(s_scratch_mu);
    if (!s_scratch_ptr[ctx->device]) {
        void *p = nullptr;
        DDP_HIP(ctx, hipSetDevice(ctx->device));
        DDP_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        DDP_CHECK(p != nullptr, ctx, DDP_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        s_scratch_ptr[ctx->device] = (int32_t *)p;
    }
    *out = s_scratch_ptr[ctx->device];
    return DDP_OK;
}

}  // namespace

extern "C" int ddp_bfs_hops(ddp_ctx *ctx, ddp_graph *g, int32_t t, int32_t reverse, int32_t scale,
                            int32_t store_as_h, int32_t *h_out, int32_t *levels_out)
{
    int rc = bfs_check(ctx, g, t, reverse, scale, "ddp_bfs_hops");
    if (rc) return rc;
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    const int32_t N = g->N;
    t = ddp_to_dev(g, t);
    void *p;
    if ((rc = ddp_ws(ctx, WS_BFS_FRONT, (size_t)N * 12 + 256, &p))) return rc;
    int32_t *dist = (int32_t *)p;
    int32_t *front = dist + N;
    if ((rc = ddp_ws(ctx, WS_BFS_CTRL, BC_WORDS * 4, &p))) return rc;
    uint32_t *ctrl = (uint32_t *)p;
    int32_t levels = 0;
    if ((rc = bfs_run(ctx, g, t, reverse, dist, front, ctrl, &levels))) return rc;
    if (levels_out) *levels_out = levels;
    int32_t *scaled = dist;
    if (scale != 1 || store_as_h) {
        if (store_as_h) {
            if (!g->h) DDP_HIP(ctx, hipMalloc((void **)&g->h, (size_t)N * 4));
            scaled = g->h;
            g->h_target = t;
        } else {
            scaled = front;  // frontier buffers are free now
        }
        hipLaunchKernelGGL(bfs_scale, dim3(1024), dim3(256), 0, ctx->stream, dist, scaled, N, scale);
    }
    if (h_out && g->iperm.empty()) {
        DDP_HIP(ctx, hipMemcpyAsync(h_out, scaled, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
    } else if (h_out) {  // device order -> caller order
        std::vector<int32_t> tmp(N);
        DDP_HIP(ctx, hipMemcpyAsync(tmp.data(), scaled, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
        DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int32_t i = 0; i < N; ++i) h_out[g->iperm[i]] = tmp[i];
    }
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DDP_OK;
}

extern "C" int ddp_bfs_hops_noalloc(ddp_ctx *ctx, const ddp_graph *g, int32_t t, int32_t reverse, int32_t scale,
                                    int32_t *h_out_dev, int32_t *levels_out)
{
    int rc = bfs_check(ctx, g, t, reverse, scale, "ddp_bfs_hops_noalloc");
    if (rc) return rc;
    DDP_CHECK(h_out_dev, ctx, DDP_ERR_ARG, "ddp_bfs_hops_noalloc: null output");
    DDP_CHECK(g->N <= kBfsScratchNodes, ctx, DDP_ERR_CAPACITY,
              "ddp_bfs_hops_noalloc: N = %d exceeds the %d nodes its static device memory holds", g->N,
              kBfsScratchNodes);
    DDP_HIP(ctx, hipSetDevice(ctx->device));
    int32_t *base = nullptr;
    if ((rc = scratch_dev(ctx, &base))) return rc;
    const int32_t N = g->N;
    int32_t *dist = base, *front = base + N;  // dist N, front 2N, ctrl BC_WORDS
    uint32_t *ctrl = (uint32_t *)(base + 3 * (size_t)N);
    int32_t levels = 0;
    if ((rc = bfs_run(ctx, g, ddp_to_dev(g, t), reverse, dist, front, ctrl, &levels))) return rc;
    if (levels_out) *levels_out = levels;
    if (g->d_iperm)
        hipLaunchKernelGGL(bfs_scale_perm, dim3(1024), dim3(256), 0, ctx->stream, dist, (const int32_t *)g->d_iperm,
                           h_out_dev, N, scale);
    else
        hipLaunchKernelGGL(bfs_scale, dim3(1024), dim3(256), 0, ctx->stream, dist, h_out_dev, N, scale);
    DDP_HIP(ctx, hipGetLastError());
    DDP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DDP_OK;
}
