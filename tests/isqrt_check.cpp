// Checks isqrt_u64 / isqrt_settle / isqrt_sum_sq (ddp_amd/csrc/isqrt.h) against the defining
// property r*r <= x < (r+1)*(r+1), evaluated in 128-bit arithmetic, on every
// perfect square neighbourhood below 2^44 (where the single-precision path
// runs), across its boundary, and on pseudo-random x < 2^63.  Prints the
// number of failures and exits non-zero if there are any.
#include <cstdio>
#include <cstdint>
#include <cmath>
#include <initializer_list>

#include "isqrt.h"

static float sqrt_exact(float v) { return sqrtf(v); }
static float sqrt_up(float v) { return nextafterf(sqrtf(v), INFINITY); }
static float sqrt_down(float v) { return nextafterf(sqrtf(v), 0.0f); }

static bool ok(uint64_t x, uint64_t r)
{
    const unsigned __int128 a = (unsigned __int128)r * r, b = (unsigned __int128)(r + 1) * (r + 1);
    return a <= x && (unsigned __int128)x < b;
}

int main()
{
    long bad = 0, n_checked = 0;
    auto check = [&](uint64_t x) {
        ++n_checked;
        if (!ok(x, isqrt_u64(x))) {
            if (bad < 10) std::printf("isqrt_u64(%llu) = %u\n", (unsigned long long)x, isqrt_u64(x));
            ++bad;
        }
    };
    // neighbourhoods of every square up to (2^22 + 4)^2, past the 2^44 switch
    for (uint64_t n = 0; n <= (1ull << 22) + 4; ++n) {
        const uint64_t q = n * n;
        if (q) check(q - 1);
        check(q);
        check(q + 1);
        check(q + n);
        check(q + 2 * n);
    }
    for (uint64_t x = (1ull << 44) - 4096; x < (1ull << 44) + 4096; ++x) check(x);
    // settle: every estimate within one of the floor
    for (uint64_t n = 1; n < (1ull << 20); n += 7) {
        for (uint64_t x : {n * n, n * n + n, n * n + 2 * n}) {
            for (int d = -1; d <= 1; ++d) {
                ++n_checked;
                if (!ok(x, isqrt_settle(x, (uint32_t)(n + d)))) ++bad;
            }
        }
    }
    // isqrt_sum_sq: the single-precision path (|dx|, |dy| < 2^21) with the
    // square root exact, one ulp high and one ulp low, across its boundary
    auto check2 = [&](int64_t dx, int64_t dy) {
        const uint64_t x = (uint64_t)(dx * dx + dy * dy);
        const uint32_t r[3] = {isqrt_sum_sq<sqrt_exact>(dx, dy), isqrt_sum_sq<sqrt_up>(dx, dy),
                               isqrt_sum_sq<sqrt_down>(dx, dy)};
        for (uint32_t v : r) {
            ++n_checked;
            if (!ok(x, v)) {
                if (bad < 10) std::printf("isqrt_sum_sq(%lld, %lld) = %u\n", (long long)dx, (long long)dy, v);
                ++bad;
            }
        }
    };
    for (int64_t k = 0; k <= (1 << 21) + 2; ++k) {
        check2(k, 0);
        check2(-k, 1);
        check2(k, k);
        check2(k, -(k >> 1));
    }
    // Pythagorean triples (m^2 - n^2, 2mn) and their neighbours: sums that are
    // perfect squares, one less, one more
    for (int64_t m = 2; m < 1400; ++m)
        for (int64_t n = 1; n < m; n += 3) {
            const int64_t a = m * m - n * n, b2 = 2 * m * n;
            for (int64_t da = -1; da <= 1; ++da) check2(a + da, b2), check2(-b2, a + da);
        }
    uint64_t y = 777;
    for (int i = 0; i < 2000000; ++i) {
        y += 0x9E3779B97F4A7C15ull;
        uint64_t v = y;
        v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
        v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
        v ^= v >> 31;
        const int sh = 10 + (i % 23);  // magnitudes up to 2^32, both paths
        check2((int64_t)(int32_t)(v & 0xFFFFFFFF) >> (32 - sh), (int64_t)(int32_t)(v >> 32) >> (32 - sh));
    }
    // pseudo-random, spread over every magnitude below 2^63
    uint64_t z = 12345;
    for (int i = 0; i < 4000000; ++i) {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t v = z;
        v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
        v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
        v ^= v >> 31;
        check(v >> (1 + (i % 63)));
    }
    std::printf("checked %ld failures %ld\n", n_checked, bad);
    return bad ? 1 : 0;
}
