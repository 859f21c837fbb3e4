// floor(sqrt(x)) for the closed-form Euclidean heuristic, exact for every
// x < 2^63 (squared distances of int32 coordinates).  Shared by host and
// device code; the host-only build (g++) sees plain inline functions.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define DDP_HOST_DEVICE __host__ __device__
#else
#define DDP_HOST_DEVICE
#endif

// Given an estimate r within one of floor(sqrt(x)) (and r < 2^32 - 1),
// returns the floor: one step up or down brings r*r <= x < (r+1)*(r+1).
DDP_HOST_DEVICE inline uint32_t isqrt_settle(uint64_t x, uint32_t r)
{
    const uint64_t r2 = (uint64_t)r * r;
    if (r2 > x) return r - 1;
    if (r2 + 2ull * r + 1 <= x) return r + 1;
    return r;
}

// Any x < 2^63: double estimate, then integer steps to the floor.
DDP_HOST_DEVICE inline uint32_t isqrt_u64_wide(uint64_t x)
{
    uint64_t r = (uint64_t)sqrt((double)x);
    while (r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return (uint32_t)r;
}

DDP_HOST_DEVICE inline uint32_t isqrt_u64(uint64_t x)
{
    if (x < (1ull << 44)) {
        // float(x) is off by at most 2^-24 relative and the square root adds
        // at most one ulp (2^-23): |estimate - sqrt(x)| < 2^22 * 1.25 * 2^-23
        // < 1, so the truncated estimate is within one of the floor
        return isqrt_settle(x, (uint32_t)sqrtf((float)x));
    }
    return isqrt_u64_wide(x);
}

// Single-precision square root with at most one ulp of error: the hardware
// instruction on the device, the correctly rounded library call on the host.
DDP_HOST_DEVICE inline float sqrt_1ulp(float v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(v);
#else
    return sqrtf(v);
#endif
}

// floor(sqrt(dx*dx + dy*dy)) for coordinate differences of int32 points.
// With |dx|, |dy| < 2^21 every step is exact or bounded: the float
// conversions are exact, dy*dy and the fused multiply-add round once each
// (relative error below 2^-23 together), halved by the square root, which
// adds at most one ulp; the estimate is then off by less than
// 2^22 * 1.5 * 2^-23 < 1 and one integer step settles it.  SQRT is the
// single-precision square root used (at most one ulp of error).
template <float (*SQRT)(float) = sqrt_1ulp>
DDP_HOST_DEVICE inline uint32_t isqrt_sum_sq(int64_t dx, int64_t dy)
{
    if ((uint64_t)(dx + (1 << 21)) < (1ull << 22) && (uint64_t)(dy + (1 << 21)) < (1ull << 22)) {
        const float fx = (float)(int32_t)dx, fy = (float)(int32_t)dy;
        const uint32_t r = (uint32_t)SQRT(fmaf(fx, fx, fy * fy));
        const uint32_t ax = (uint32_t)(dx < 0 ? -dx : dx), ay = (uint32_t)(dy < 0 ? -dy : dy);
        return isqrt_settle((uint64_t)ax * ax + (uint64_t)ay * ay, r);
    }
    return isqrt_u64_wide((uint64_t)(dx * dx + dy * dy));
}
