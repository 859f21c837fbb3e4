// DA*'s host-side insert grouping (ddp_amd/csrc/host_sort.h): radix_pairs must
// equal a stable sort by key, for empty, single and duplicate-heavy batches and
// key ranges of one, two and three 11-bit passes, and never write past 2m.
#include <algorithm>
#include <cstdio>
#include <random>
#include <utility>
#include <vector>

#include "host_sort.h"

int main()
{
    std::mt19937 rng(7);
    int bad = 0, cases = 0;
    std::vector<int32_t> tk, tv;  // scratch reused across calls, as between batches
    for (int n : {1, 2, 5, 2048, 2049, 1000000, 1 << 22, 1 << 30}) {
        for (int m : {0, 1, 2, 7, 1000, 20000}) {
            std::vector<int32_t> k(m), v(m), out(2 * (size_t)m + 2, -7);
            for (int i = 0; i < m; ++i) {
                k[i] = (int32_t)(rng() % (uint32_t)n);
                if (i % 5 == 0 && i) k[i] = k[i - 1];  // runs of equal keys
                v[i] = 1000 + i;                       // arrival order
            }
            radix_pairs(k.data(), v.data(), m, n, out.data(), tk, tv);
            std::vector<std::pair<int32_t, int32_t>> ref(m);
            for (int i = 0; i < m; ++i) ref[i] = {k[i], v[i]};
            std::stable_sort(ref.begin(), ref.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
            ++cases;
            for (int i = 0; i < m; ++i)
                if (out[2 * i] != ref[i].first || out[2 * i + 1] != ref[i].second) {
                    std::printf("mismatch n=%d m=%d at %d\n", n, m, i);
                    ++bad;
                    break;
                }
            if (out[2 * (size_t)m] != -7 || out[2 * (size_t)m + 1] != -7) {
                std::printf("write past the end n=%d m=%d\n", n, m);
                ++bad;
            }
        }
    }
    std::printf("cases %d failures %d\n", cases, bad);
    return bad != 0;
}
