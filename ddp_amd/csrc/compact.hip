// Compaction of an edge-update log back into CSR rows (the role of
// merge_graph / mergeDiff in DA_star/DA_dyn_insert_del_nocycle.cu:940-980,
// 2015-2102, 2236-2284 and DA_star/GA_merge.cu:1042-1089), on the device.
//
// The source is a graph version as DA* and the server keep it: base rows whose
// edges may be deleted (a negative column, or a delete version at or below
// the version being compacted) plus inserted edges threaded onto per-source
// lists, oldest first.  The merged row of u at version `ver` is u's surviving
// base edges in row order followed by its live inserted edges in insert order
// -- exactly the row GA_merge.cu builds, so "first live instance" deletes
// keep their meaning across a compaction.  The reference rebuilds each row
// with a one-thread launch and a device-to-host copy per node; here one wave
// per row counts and fills the rows of the whole graph in two passes around a
// device-wide scan of the row lengths.
#include <algorithm>

#include "compact.h"
#include "device_common.h"

namespace {

constexpr int kT = 256;  // threads per block of the row passes (4 waves)

__device__ __forceinline__ bool base_live(const MergeSrc &m, int32_t j)
{
    return m.col[j] >= 0 && (!m.dver || m.dver[j] > m.ver);
}

// Next live list entry at or after link (the link already read), or -1.
__device__ __forceinline__ int32_t list_next(const MergeSrc &m, uint64_t &link)
{
    while ((uint32_t)link <= (uint32_t)m.ver) {
        const int32_t x = (int32_t)(link >> 32);
        link = m.lnext[x];
        if (m.ldel[x] > m.ver) return x;
    }
    return -1;
}

// One wave per row: cnt[u] = length of u's merged row.
__global__ void merge_count(MergeSrc m, int32_t *cnt)
{
    const int64_t wave = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kT) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t u = wave; u < m.N; u += nwaves) {
        const int32_t b = m.off[u], e = m.off[u + 1];
        uint32_t n = 0;
        for (int32_t j = b + lane; j < e; j += 64) n += base_live(m, j) ? 1u : 0u;
        n = (uint32_t)wave_sum_u64(n);
        if (lane == 0 && m.lhead) {
            uint64_t link = m.lhead[u];
            while (list_next(m, link) >= 0) ++n;
        }
        if (lane == 0) cnt[u] = (int32_t)n;
    }
}

// One wave per row: the surviving base edges in row order (ballot + mbcnt
// keeps their order), then the live inserted edges in insert order.
__global__ void merge_fill(MergeSrc m, const int32_t *noff, int32_t *ncol, uint32_t *nw)
{
    const int64_t wave = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kT) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t u = wave; u < m.N; u += nwaves) {
        const int32_t b = m.off[u], e = m.off[u + 1];
        int32_t pos = noff[u];
        for (int32_t j0 = b; j0 < e; j0 += 64) {
            const int32_t j = j0 + lane;
            const bool live = j < e && base_live(m, j);
            const uint64_t mask = __ballot(live);
            if (live) {
                const int32_t p = pos + (int32_t)lane_rank(mask);
                ncol[p] = m.col[j];
                nw[p] = m.w[j];
            }
            pos += __popcll(mask);
        }
        if (lane == 0 && m.lhead) {
            uint64_t link = m.lhead[u];
            for (int32_t x; (x = list_next(m, link)) >= 0; ++pos) {
                ncol[pos] = m.lv[x];
                nw[pos] = m.lw[x];
            }
        }
    }
}

// Exclusive scan of n counts, three passes: per-block sums, a scan of those
// sums by one block, and a pass adding each block's offset.  A block covers
// kScanItems counts (kT threads x 4).
constexpr int kScanItems = kT * 4;

__global__ void scan_local(const int32_t *in, int64_t n, int64_t *out, int64_t *bsum)
{
    __shared__ int64_t s[kT];
    const int64_t base = (int64_t)blockIdx.x * kScanItems + threadIdx.x * 4;
    int64_t v[4], run = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        run += v[k];
    }
    s[threadIdx.x] = run;
    __syncthreads();
    for (int d = 1; d < kT; d <<= 1) {  // Hillis-Steele over the block's thread sums
        const int64_t x = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    int64_t ex = s[threadIdx.x] - run;
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == kT - 1) bsum[blockIdx.x] = s[kT - 1];
}

__global__ void scan_sums(int64_t *bsum, int64_t nb, int64_t *total)
{
    __shared__ int64_t s[kT];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kT) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < nb ? bsum[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < kT; d <<= 1) {
            const int64_t x = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += x;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
        const int64_t blk = s[kT - 1];
        __syncthreads();
        carry += blk;
    }
    if (threadIdx.x == 0) *total = carry;
}

// Adds the block offsets and narrows to int32 (the caller checked that the
// total fits); out32[n] = total.
__global__ void scan_finish(const int64_t *ex, const int64_t *bsum, int64_t n, const int64_t *total, int32_t *out32)
{
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i <= n; i += (int64_t)gridDim.x * kT)
        out32[i] = (int32_t)(i < n ? ex[i] + bsum[i / kScanItems] : *total);
}

// Transposed rows: in-degrees, then one atomic slot per edge.  The order of
// predecessors inside a transposed row is therefore not fixed; nothing that
// reads these rows depends on it (the best in-edge, or the instance whose
// forward id matches, is found by a full scan of the row).
__global__ void rev_count(const int32_t *ncol, int64_t E, int32_t *rcnt)
{
    for (int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x; j < E; j += (int64_t)gridDim.x * kT)
        atomicAdd(rcnt + ncol[j], 1);
}

__global__ void rev_fill(const int32_t *noff, const int32_t *ncol, const uint32_t *nw, int32_t N, int32_t *cursor,
                         int32_t *rcol, uint32_t *rw, int32_t *reid)
{
    const int64_t wave = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kT) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t u = wave; u < N; u += nwaves)
        for (int32_t j = noff[u] + lane; j < noff[u + 1]; j += 64) {
            const int32_t p = atomicAdd(cursor + ncol[j], 1);
            rcol[p] = (int32_t)u;
            rw[p] = nw[j];
            if (reid) reid[p] = j;
        }
}

int grid_for(ddp_ctx *ctx, int64_t items_per_block, int64_t n)
{
    const int64_t want = (n + items_per_block - 1) / items_per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)ctx->prop.multiProcessorCount * 16));
}

}  // namespace

int ddp_scan_counts(ddp_ctx *ctx, const int32_t *cnt, int64_t n, int32_t *out, int64_t *total)
{
    const int64_t nb = std::max<int64_t>(1, (n + kScanItems - 1) / kScanItems);
    int64_t *ex = nullptr, *bsum = nullptr, *dtot = nullptr;
    hipError_t e = hipMalloc((void **)&ex, (size_t)std::max<int64_t>(n, 1) * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&bsum, (size_t)nb * 8 + 8);
    if (e == hipSuccess) {
        dtot = bsum + nb;
        hipLaunchKernelGGL(scan_local, dim3((unsigned)nb), dim3(kT), 0, ctx->stream, cnt, n, ex, bsum);
        hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kT), 0, ctx->stream, bsum, nb, dtot);
        e = hipMemcpyAsync(total, dtot, 8, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && *total <= INT32_MAX) {
        hipLaunchKernelGGL(scan_finish, dim3(grid_for(ctx, kT, n + 1)), dim3(kT), 0, ctx->stream, ex, bsum, n, dtot,
                           out);
        e = hipStreamSynchronize(ctx->stream);
    }
    if (ex) hipFree(ex);
    if (bsum) hipFree(bsum);
    if (e != hipSuccess) return ddp_hip_err(ctx, e, "row-length scan", __FILE__, __LINE__);
    DDP_CHECK(*total <= INT32_MAX, ctx, DDP_ERR_CAPACITY, "compacted graph would hold %lld edges (int32 ids)",
              (long long)*total);
    return DDP_OK;
}

int ddp_merge_rows(ddp_ctx *ctx, const MergeSrc &m, MergedCsr *out)
{
    *out = MergedCsr{};
    const int32_t N = m.N;
    int32_t *cnt = nullptr;
    hipError_t e = hipMalloc((void **)&cnt, (size_t)N * 4 + 4);
    if (e != hipSuccess) return ddp_hip_err(ctx, e, "merge: counts", __FILE__, __LINE__);
    hipLaunchKernelGGL(merge_count, dim3(grid_for(ctx, kT / 64, N)), dim3(kT), 0, ctx->stream, m, cnt);
    int rc = DDP_OK;
    e = hipMalloc((void **)&out->off, (size_t)(N + 1) * 4);
    if (e != hipSuccess) rc = ddp_hip_err(ctx, e, "merge: offsets", __FILE__, __LINE__);
    int64_t E = 0;
    if (!rc) rc = ddp_scan_counts(ctx, cnt, N, out->off, &E);
    hipFree(cnt);
    if (rc) {
        ddp_merged_free(out);
        return rc;
    }
    out->N = N;
    out->E = (int32_t)E;
    e = hipMalloc((void **)&out->col, (size_t)std::max<int64_t>(E, 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&out->w, (size_t)std::max<int64_t>(E, 1) * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(merge_fill, dim3(grid_for(ctx, kT / 64, N)), dim3(kT), 0, ctx->stream, m,
                           (const int32_t *)out->off, out->col, out->w);
        e = hipStreamSynchronize(ctx->stream);
    }
    if (e != hipSuccess) {
        ddp_merged_free(out);
        return ddp_hip_err(ctx, e, "merge: fill", __FILE__, __LINE__);
    }
    return DDP_OK;
}

int ddp_transpose_rows(ddp_ctx *ctx, const MergedCsr &f, bool with_eid, MergedCsr *out)
{
    *out = MergedCsr{};
    const int32_t N = f.N;
    const int64_t E = f.E;
    int32_t *cnt = nullptr;
    hipError_t e = hipMalloc((void **)&cnt, (size_t)N * 4 + 4);
    if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, (size_t)N * 4, ctx->stream);
    if (e != hipSuccess) {
        if (cnt) hipFree(cnt);
        return ddp_hip_err(ctx, e, "transpose: counts", __FILE__, __LINE__);
    }
    hipLaunchKernelGGL(rev_count, dim3(grid_for(ctx, kT, E)), dim3(kT), 0, ctx->stream, f.col, E, cnt);
    int rc = DDP_OK;
    e = hipMalloc((void **)&out->off, (size_t)(N + 1) * 4);
    if (e != hipSuccess) rc = ddp_hip_err(ctx, e, "transpose: offsets", __FILE__, __LINE__);
    int64_t RE = 0;
    if (!rc) rc = ddp_scan_counts(ctx, cnt, N, out->off, &RE);
    if (rc) {
        hipFree(cnt);
        ddp_merged_free(out);
        return rc;
    }
    out->N = N;
    out->E = (int32_t)RE;
    e = hipMemcpyAsync(cnt, out->off, (size_t)N * 4, hipMemcpyDeviceToDevice, ctx->stream);  // write cursors
    if (e == hipSuccess) e = hipMalloc((void **)&out->col, (size_t)std::max<int64_t>(RE, 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&out->w, (size_t)std::max<int64_t>(RE, 1) * 4);
    if (e == hipSuccess && with_eid) e = hipMalloc((void **)&out->eid, (size_t)std::max<int64_t>(RE, 1) * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(rev_fill, dim3(grid_for(ctx, kT / 64, N)), dim3(kT), 0, ctx->stream,
                           (const int32_t *)f.off, (const int32_t *)f.col, (const uint32_t *)f.w, N, cnt, out->col,
                           out->w, out->eid);
        e = hipStreamSynchronize(ctx->stream);
    }
    hipFree(cnt);
    if (e != hipSuccess) {
        ddp_merged_free(out);
        return ddp_hip_err(ctx, e, "transpose: fill", __FILE__, __LINE__);
    }
    return DDP_OK;
}

void ddp_merged_free(MergedCsr *c)
{
    void *ptrs[] = {c->off, c->col, c->w, c->eid};
    for (void *p : ptrs)
        if (p) hipFree(p);
    *c = MergedCsr{};
}
