// Device-side building blocks shared by the search kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "ddp_internal.h"
#include "isqrt.h"

#define DDP_INF_F 0x7FFFFFFFu                      // Cx "infinity" (INT_MAX)
#define DDP_GP_INF 0x7FFFFFFFFFFFFFFFull           // packed (f=INT_MAX, parent=-1)

// ------------------------------------------------------------ coherent access
// The packed (f, parent) array is updated with device-scope atomics that are
// performed in L2; plain loads could be served from a stale vector-L1 line.
// These loads bypass L1 (relaxed, agent scope).
__device__ __forceinline__ uint64_t ld_gp(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_u32(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_u32(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------- heuristics

// A heuristic value in two steps, so that a lane can issue the loads for
// several nodes before the first value is computed: heur_load reads what h(v)
// needs from HBM (the h array entry, or v's coordinates), heur_from finishes.
// (No initialisers: a lane that skips heur_load never reads its HeurIn, and
// leaving it undefined spares the compiler a merge that would wait on the load.)
struct HeurIn {
    int32_t a, b;
};

template <int HK>
__device__ __forceinline__ HeurIn heur_load(const HeurView &H, int32_t v)
{
    HeurIn in{0, 0};
    if constexpr (HK == DDP_H_ARRAY) {
        in.a = H.h[v];
    } else if constexpr (HK == DDP_H_EUCLID) {
        const int2 p = reinterpret_cast<const int2 *>(H.xy)[v];
        in.a = p.x;
        in.b = p.y;
    }
    return in;
}

template <int HK>
__device__ __forceinline__ int32_t heur_from(const HeurView &H, int32_t v, HeurIn in)
{
    if constexpr (HK == DDP_H_ZERO) {
        return 0;
    } else if constexpr (HK == DDP_H_ARRAY) {
        return in.a;
    } else if constexpr (HK == DDP_H_MANHATTAN) {
        int32_t r = v / H.width, c = v - r * H.width;
        int32_t dr = r - H.ty, dc = c - H.tx;
        dr = dr < 0 ? -dr : dr;
        dc = dc < 0 ? -dc : dc;
        return (dr + dc) * H.scale;
    } else {
        return (int32_t)isqrt_sum_sq((int64_t)in.a - H.tx, (int64_t)in.b - H.ty) * H.scale;
    }
}

template <int HK>
__device__ __forceinline__ int32_t heur_eval(const HeurView &H, int32_t v)
{
    return heur_from<HK>(H, v, heur_load<HK>(H, v));
}

// ---------------------------------------------------------- relaxation
// Strict-decrease update of the packed (f << 32 | parent) word: succeeds only
// if cand < current f.  Parent changes only with a strict f decrease, so the
// parent pointers stay acyclic even with zero-weight edges.  Returns true on
// success and the previous word in *prev.
__device__ __forceinline__ bool relax_gp(uint64_t *gp, int32_t c, uint32_t cand,
                                         uint32_t parent, uint64_t *prev)
{
    uint64_t cur = ld_gp(gp + c);
    const uint64_t nv = ((uint64_t)cand << 32) | parent;
    while ((uint32_t)(cur >> 32) > cand) {
        uint64_t old = atomicCAS((unsigned long long *)(gp + c), (unsigned long long)cur,
                                 (unsigned long long)nv);
        if (old == cur) { *prev = cur; return true; }
        cur = old;
    }
    return false;
}

// ---------------------------------------------------------------- heaps
// One binary min-heap per lane ("queue").  Entry key = f << 32 | node, so
// equal-f entries pop in node-id order.  Entries [0, lcap) live in LDS at
// lds[i * lstride]; the rest spill to HBM at home[i * hstride].
struct LaneHeap {
    uint64_t *lds;
    uint64_t *home;
    int32_t lstride;
    int32_t lcap;
    int64_t hstride;
    int32_t cap;

    __device__ __forceinline__ uint64_t get(int32_t i) const
    {
        return i < lcap ? lds[i * lstride] : home[(int64_t)i * hstride];
    }
    __device__ __forceinline__ void set(int32_t i, uint64_t v) const
    {
        if (i < lcap) lds[i * lstride] = v;
        else home[(int64_t)i * hstride] = v;
    }
};

__device__ __forceinline__ bool heap_push(const LaneHeap &H, int32_t &size, uint64_t key)
{
    if (size >= H.cap) return false;
    int32_t i = size++;
    while (i > 0) {
        int32_t p = (i - 1) >> 1;
        uint64_t kp = H.get(p);
        if (kp <= key) break;
        H.set(i, kp);
        i = p;
    }
    H.set(i, key);
    return true;
}

__device__ __forceinline__ bool heap_pop(const LaneHeap &H, int32_t &size, uint64_t &top)
{
    if (size <= 0) return false;
    top = H.get(0);
    --size;
    if (size == 0) return true;
    uint64_t last = H.get(size);
    int32_t i = 0;
    for (;;) {
        int32_t l = 2 * i + 1;
        if (l >= size) break;
        uint64_t km = H.get(l);
        int32_t m = l;
        if (l + 1 < size) {
            uint64_t kr = H.get(l + 1);
            if (kr < km) { km = kr; m = l + 1; }
        }
        if (km >= last) break;
        H.set(i, km);
        i = m;
    }
    H.set(i, last);
    return true;
}

// ----------------------------------------------------- wave / block helpers
// Wave-wide inclusive prefix sum with DPP row shifts and row broadcasts (no
// LDS traffic): shifts by 1, 2, 4 and 8 scan each 16-lane row, then lane 15
// of rows 0 and 2 is added into rows 1 and 3, and lane 31 into rows 2 and 3.
// Lanes whose source is outside the row, or whose row is masked off, add 0.
__device__ __forceinline__ int32_t wave_incl_scan(int32_t v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t o = __shfl_xor(v, d, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Block-wide inclusive scan; `tmp` must hold B/64 ints of LDS.  Returns the
// block total too.  With TAIL_SYNC it ends with a barrier, so `tmp` may be
// reused right after; without, the caller must pass another barrier before
// anything writes `tmp` again.
template <int B, bool TAIL_SYNC = true>
__device__ __forceinline__ int32_t block_incl_scan(int32_t v, int32_t *tmp, int32_t &total)
{
    int32_t x = wave_incl_scan(v);
    if constexpr (B == 64) {
        total = __builtin_amdgcn_readlane(x, 63);
        return x;
    } else {
        const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 63) tmp[wid] = x;
        __syncthreads();
        int32_t add = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < B / 64; ++k) {
            int32_t s = tmp[k];
            if (k < wid) add += s;
            tot += s;
        }
        total = tot;
        if constexpr (TAIL_SYNC) __syncthreads();
        return x + add;
    }
}

template <int B>
__device__ __forceinline__ uint32_t block_min_u32(uint32_t v, uint32_t *tmp)
{
    uint32_t m = wave_min_u32(v);
    if constexpr (B == 64) {
        return m;
    } else {
        const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 0) tmp[wid] = m;
        __syncthreads();
        uint32_t r = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < B / 64; ++k) r = tmp[k] < r ? tmp[k] : r;
        __syncthreads();
        return r;
    }
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-aggregated append of up to R values per lane: one atomicAdd on *count
// (LDS or global) reserves room for every selected slot of the wave; slot r
// of a lane lands after the wave's selected slots r' < r and after slot r of
// the lanes below it.  pos[r] = UINT32_MAX for unselected slots.  Every lane
// of the wave must call it.
template <int R>
__device__ __forceinline__ void wave_append(const bool (&sel)[R], uint32_t *count, uint32_t (&pos)[R])
{
    uint64_t m[R];
    uint32_t off[R], tot = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        m[r] = __ballot(sel[r]);
        off[r] = tot;
        tot += (uint32_t)__popcll(m[r]);
    }
    uint32_t base = 0;
    if (tot) {
        if ((threadIdx.x & 63) == 0) base = atomicAdd(count, tot);
        base = __builtin_amdgcn_readfirstlane(base);  // lane 0's value (every lane is active)
    }
#pragma unroll
    for (int r = 0; r < R; ++r) pos[r] = sel[r] ? base + off[r] + lane_rank(m[r]) : 0xFFFFFFFFu;
}

// On gfx950 __syncthreads() is a bare s_barrier: a wave's global stores may
// still be in flight when another wave of the workgroup passes the barrier
// and issues an atomic (performed in L2) on the same address.  Waiting for
// this wave's vector-memory counter first makes its stores visible in L2 to
// every wave that passes the barrier.
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ void block_sync_global()
{
    wait_vmem();
    __syncthreads();
}

// 100 MHz constant-rate device clock, for bounded waits.
__device__ __forceinline__ uint64_t dev_clock_100mhz()
{
    return __builtin_amdgcn_s_memrealtime();
}

// Sense-counting grid barrier over `nblocks` co-resident workgroups.
// bar[0] = arrival count, bar[1] = generation.  Every wait is bounded by
// `limit_ticks` of the 100 MHz clock; returns false if the wait timed out.
__device__ __forceinline__ bool grid_barrier(uint32_t *bar, uint32_t nblocks, uint32_t &gen,
                                             uint64_t limit_ticks, int32_t *s_flag)
{
    // every wave's stores reach L2 before thread 0's agent-scope release
    block_sync_global();
    if (threadIdx.x == 0) {
        bool ok = true;
        __threadfence();
        uint32_t arrived = atomicAdd(&bar[0], 1u);
        if (arrived == nblocks - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = dev_clock_100mhz();
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (dev_clock_100mhz() - t0 > limit_ticks) { ok = false; break; }
            }
        }
        __threadfence();
        *s_flag = ok ? 1 : 0;
    }
    gen += 1;
    __syncthreads();
    return *s_flag != 0;
}
