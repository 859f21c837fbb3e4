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

DDP_HOST_DEVICE inline uint32_t isqrt_u64(uint64_t x)
{
    if (x < (1ull << 44)) {
        // float(x) is off by at most 2^-24 relative and the square root adds
        // at most one ulp (2^-23): |estimate - sqrt(x)| < 2^22 * 1.25 * 2^-23
        // < 1, so the truncated estimate is within one of the floor
        return isqrt_settle(x, (uint32_t)sqrtf((float)x));
    }
    // double estimate then integer fix-up
    uint64_t r = (uint64_t)sqrt((double)x);
    while (r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return (uint32_t)r;
}
