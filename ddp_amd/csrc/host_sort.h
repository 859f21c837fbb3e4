// Host-side grouping of a DA* batch's inserts by node (dastar.hip): no HIP
// dependency, so it is also built and checked on the CPU.
#pragma once

#include <cstdint>
#include <vector>

// Stable LSD radix sort of m (key, value) pairs by key (keys < n), 11 bits a
// pass; the last pass writes them interleaved (key, value) into out.  tk and
// tv are scratch kept between batches.
inline void radix_pairs(const int32_t *key, const int32_t *val, int32_t m, int32_t n, int32_t *out,
                        std::vector<int32_t> &tk, std::vector<int32_t> &tv)
{
    int bits = 1;
    while (bits < 31 && (1ll << bits) < (int64_t)n) ++bits;
    const int passes = (bits + 10) / 11;
    tk.resize(2 * (size_t)m);
    tv.resize(2 * (size_t)m);
    const int32_t *ck = key, *cv = val;
    int32_t *bk[2] = {tk.data(), tk.data() + m}, *bv[2] = {tv.data(), tv.data() + m};
    for (int p = 0; p < passes; ++p) {
        const int sh = 11 * p;
        uint32_t cnt[2049] = {0};
        for (int32_t i = 0; i < m; ++i) cnt[(((uint32_t)ck[i] >> sh) & 2047u) + 1]++;
        for (int i = 0; i < 2048; ++i) cnt[i + 1] += cnt[i];
        if (p == passes - 1) {
            for (int32_t i = 0; i < m; ++i) {
                const uint32_t j = cnt[((uint32_t)ck[i] >> sh) & 2047u]++;
                out[2 * j] = ck[i];
                out[2 * j + 1] = cv[i];
            }
        } else {
            int32_t *nk = bk[p & 1], *nv = bv[p & 1];
            for (int32_t i = 0; i < m; ++i) {
                const uint32_t j = cnt[((uint32_t)ck[i] >> sh) & 2047u]++;
                nk[j] = ck[i];
                nv[j] = cv[i];
            }
            ck = nk;
            cv = nv;
        }
    }
}
