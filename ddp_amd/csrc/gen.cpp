// Seeded synthetic inputs for the benchmark configurations.
//
// The reference's generators are unseeded (GA_star/randomGraph.py,
// maze/create_maze.cpp:70) and allow zero weights (GA_star/randomGraph.py:12);
// these use std::mt19937_64 (a fixed, standardised sequence) with explicit
// integer sampling so every run on every machine gets the same inputs.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <random>
#include <thread>
#include <vector>

#include "../../include/ddp_capi.h"
#include "../../include/ddp_io.h"
#include "hgraph.h"

namespace {

struct Rng {
    std::mt19937_64 e;
    explicit Rng(uint64_t s) : e(s) {}
    // uniform in [0, n), Lemire's multiply-shift with rejection
    uint64_t below(uint64_t n)
    {
        uint64_t x = e();
        __uint128_t m = (__uint128_t)x * n;
        uint64_t l = (uint64_t)m;
        if (l < n) {
            const uint64_t t = (0 - n) % n;
            while (l < t) {
                x = e();
                m = (__uint128_t)x * n;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
    double unit() { return (double)(e() >> 11) * 0x1.0p-53; }
};

uint64_t mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint32_t isqrt64(uint64_t x)
{
    uint64_t r = (uint64_t)std::sqrt((double)x);
    while (r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return (uint32_t)r;
}

// CSR from (src, dst, w) with rows in id order and arrival order inside a row.
ddp_hgraph *csr_from_edges(int32_t N, const std::vector<int32_t> &src, const std::vector<int32_t> &dst,
                           const std::vector<uint32_t> &w)
{
    ddp_hgraph *g = new ddp_hgraph();
    g->N = N;
    g->E = (int32_t)src.size();
    std::vector<int64_t> cnt((size_t)N + 1, 0);
    for (int32_t s : src) cnt[s + 1]++;
    for (int32_t i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
    g->off.resize(N);
    for (int32_t i = 0; i < N; ++i) g->off[i] = (int32_t)cnt[i];
    g->col.resize(src.size());
    g->w.resize(src.size());
    for (size_t k = 0; k < src.size(); ++k) {
        const int64_t p = cnt[src[k]]++;
        g->col[p] = dst[k];
        g->w[p] = w[k];
    }
    return g;
}

bool reachable(const ddp_hgraph *g, int32_t s, int32_t t, std::vector<int32_t> &mark, int32_t stamp)
{
    std::vector<int32_t> st{s};
    mark[s] = stamp;
    while (!st.empty()) {
        const int32_t u = st.back();
        st.pop_back();
        if (u == t) return true;
        const int32_t b = g->off[u], e = (u + 1 < g->N) ? g->off[u + 1] : g->E;
        for (int32_t j = b; j < e; ++j) {
            const int32_t c = g->col[j];
            if (c >= 0 && mark[c] != stamp) {
                mark[c] = stamp;
                st.push_back(c);
            }
        }
    }
    return false;
}

}  // namespace

extern "C" int ddp_gen_random(int32_t N, int32_t deg, uint32_t wmin, uint32_t wmax, uint64_t seed,
                              ddp_hgraph **out)
{
    if (!out || N < 2 || deg < 0 || wmin > wmax || (int64_t)N * deg > INT32_MAX) return DDP_ERR_ARG;
    Rng r(seed);
    ddp_hgraph *g = new ddp_hgraph();
    g->N = N;
    g->E = N * deg;
    g->off.resize(N);
    g->col.resize((size_t)g->E);
    g->w.resize((size_t)g->E);
    for (int32_t u = 0; u < N; ++u) {
        g->off[u] = u * deg;
        for (int32_t k = 0; k < deg; ++k) {
            int32_t v = (int32_t)r.below((uint64_t)N - 1);  // uniform over [0,N) \ {u}
            if (v >= u) ++v;
            g->col[(size_t)u * deg + k] = v;
            g->w[(size_t)u * deg + k] = wmin + (uint32_t)r.below((uint64_t)wmax - wmin + 1);
        }
    }
    *out = g;
    return DDP_OK;
}

extern "C" int ddp_gen_grid_maze(int32_t width, int32_t height, int32_t p_block_permille, uint64_t seed,
                                 ddp_hgraph **out)
{
    if (!out || width < 1 || height < 1 || (int64_t)width * height > INT32_MAX / 4 || p_block_permille < 0 ||
        p_block_permille > 1000)
        return DDP_ERR_ARG;
    Rng r(seed);
    const int32_t N = width * height;
    std::vector<uint8_t> blocked((size_t)N);
    for (int32_t i = 0; i < N; ++i) blocked[i] = r.below(1000) < (uint64_t)p_block_permille;
    // carved monotone corridor (0,0) -> (height-1, width-1)
    int32_t rr = 0, cc = 0;
    blocked[0] = 0;
    while (rr != height - 1 || cc != width - 1) {
        if (rr == height - 1) ++cc;
        else if (cc == width - 1) ++rr;
        else if (r.below(2)) ++rr;
        else ++cc;
        blocked[(size_t)rr * width + cc] = 0;
    }
    ddp_hgraph *g = new ddp_hgraph();
    g->N = N;
    g->off.resize(N);
    g->col.reserve((size_t)N * 3);
    g->w.reserve((size_t)N * 3);
    for (int32_t i = 0; i < height; ++i)
        for (int32_t j = 0; j < width; ++j) {
            const int32_t v = i * width + j;
            g->off[v] = (int32_t)g->col.size();
            if (blocked[v]) continue;
            // neighbour order up, down, right, left (maze/create_maze.cpp:93-111)
            const int32_t nb[4] = {i > 0 ? v - width : -1, i < height - 1 ? v + width : -1,
                                   j < width - 1 ? v + 1 : -1, j > 0 ? v - 1 : -1};
            for (int32_t c : nb)
                if (c >= 0 && !blocked[c]) {
                    g->col.push_back(c);
                    g->w.push_back(1);
                }
        }
    g->E = (int32_t)g->col.size();
    *out = g;
    return DDP_OK;
}

extern "C" int ddp_gen_geometric(int32_t N, double mean_degree, int32_t scale, uint64_t seed, ddp_hgraph **out)
{
    if (!out || N < 2 || mean_degree <= 0 || scale < 1) return DDP_ERR_ARG;
    Rng r(seed);
    std::vector<int32_t> xy(2 * (size_t)N);
    for (int32_t i = 0; i < N; ++i) {
        xy[2 * (size_t)i] = (int32_t)r.below((uint64_t)scale + 1);
        xy[2 * (size_t)i + 1] = (int32_t)r.below((uint64_t)scale + 1);
    }
    // pi r^2 N = mean degree on the unit square; R in lattice units
    const double R2d = mean_degree * (double)scale * (double)scale / (M_PI * (double)N);
    const int64_t R2 = (int64_t)std::floor(R2d);
    const int32_t R = (int32_t)isqrt64((uint64_t)R2) + 1;
    const int32_t cells = scale / R + 1;
    std::vector<int32_t> cstart((size_t)cells * cells + 1, 0), cpts(N);
    auto cell_of = [&](int32_t i) { return (xy[2 * (size_t)i + 1] / R) * cells + xy[2 * (size_t)i] / R; };
    for (int32_t i = 0; i < N; ++i) cstart[cell_of(i) + 1]++;
    for (size_t c = 0; c < (size_t)cells * cells; ++c) cstart[c + 1] += cstart[c];
    {
        std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
        for (int32_t i = 0; i < N; ++i) cpts[fill[cell_of(i)]++] = i;
    }
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<int32_t>> rows_col(nt);
    std::vector<std::vector<uint32_t>> rows_w(nt);
    std::vector<std::vector<int32_t>> rows_deg(nt);
    std::vector<std::thread> th;
    for (unsigned k = 0; k < nt; ++k)
        th.emplace_back([&, k] {
            const int32_t lo = (int32_t)((int64_t)N * k / nt), hi = (int32_t)((int64_t)N * (k + 1) / nt);
            std::vector<std::pair<int32_t, uint32_t>> nb;
            for (int32_t u = lo; u < hi; ++u) {
                nb.clear();
                const int32_t x = xy[2 * (size_t)u], y = xy[2 * (size_t)u + 1];
                const int32_t cx = x / R, cy = y / R;
                for (int32_t dy = -1; dy <= 1; ++dy)
                    for (int32_t dx = -1; dx <= 1; ++dx) {
                        const int32_t ax = cx + dx, ay = cy + dy;
                        if (ax < 0 || ay < 0 || ax >= cells || ay >= cells) continue;
                        const size_t c = (size_t)ay * cells + ax;
                        for (int32_t p = cstart[c]; p < cstart[c + 1]; ++p) {
                            const int32_t v = cpts[p];
                            if (v == u) continue;
                            const int64_t ddx = xy[2 * (size_t)v] - x, ddy = xy[2 * (size_t)v + 1] - y;
                            const uint64_t d2 = (uint64_t)(ddx * ddx + ddy * ddy);
                            if ((int64_t)d2 > R2) continue;
                            uint32_t s = isqrt64(d2);
                            if ((uint64_t)s * s < d2) ++s;  // ceil(sqrt(d2))
                            nb.emplace_back(v, std::max<uint32_t>(1, s));
                        }
                    }
                std::sort(nb.begin(), nb.end());
                rows_deg[k].push_back((int32_t)nb.size());
                for (auto &e : nb) {
                    rows_col[k].push_back(e.first);
                    rows_w[k].push_back(e.second);
                }
            }
        });
    for (auto &t : th) t.join();
    ddp_hgraph *g = new ddp_hgraph();
    g->N = N;
    g->off.resize(N);
    int64_t E = 0;
    for (unsigned k = 0; k < nt; ++k) E += (int64_t)rows_col[k].size();
    if (E > INT32_MAX) {
        delete g;
        return DDP_ERR_ARG;
    }
    g->col.reserve((size_t)E);
    g->w.reserve((size_t)E);
    int64_t pos = 0;
    int32_t u = 0;
    for (unsigned k = 0; k < nt; ++k) {
        for (int32_t d : rows_deg[k]) {
            g->off[u++] = (int32_t)pos;
            pos += d;
        }
        g->col.insert(g->col.end(), rows_col[k].begin(), rows_col[k].end());
        g->w.insert(g->w.end(), rows_w[k].begin(), rows_w[k].end());
    }
    g->E = (int32_t)E;
    g->xy = std::move(xy);
    *out = g;
    return DDP_OK;
}

extern "C" int ddp_gen_rmat(int32_t scale, int64_t E, double a, double b, double c, uint32_t wmin, uint32_t wmax,
                            uint64_t seed, ddp_hgraph **out)
{
    if (!out || scale < 1 || scale > 30 || E < 0 || E > INT32_MAX || wmin > wmax || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0)
        return DDP_ERR_ARG;
    const int32_t N = 1 << scale;
    std::vector<int32_t> src((size_t)E), dst((size_t)E);
    std::vector<uint32_t> w((size_t)E);
    const int chunks = 16;  // fixed so the output does not depend on the host's core count
    std::vector<std::thread> th;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (unsigned k = 0; k < nt; ++k)
        th.emplace_back([&, k] {
            for (int ch = (int)k; ch < chunks; ch += (int)nt) {
                Rng r(mix(seed * 1315423911ull + (uint64_t)ch));
                const int64_t lo = E * ch / chunks, hi = E * (ch + 1) / chunks;
                for (int64_t i = lo; i < hi; ++i) {
                    int32_t u = 0, v = 0;
                    for (int32_t l = 0; l < scale; ++l) {
                        const double x = r.unit();
                        const int bu = x >= a + b;                       // lower half: c or d
                        const int bv = (x >= a && x < a + b) || x >= a + b + c;  // right half: b or d
                        u = (u << 1) | bu;
                        v = (v << 1) | bv;
                    }
                    src[i] = u;
                    dst[i] = v;
                    w[i] = wmin + (uint32_t)r.below((uint64_t)wmax - wmin + 1);
                }
            }
        });
    for (auto &t : th) t.join();
    *out = csr_from_edges(N, src, dst, w);
    return DDP_OK;
}

extern "C" int ddp_gen_updates(const ddp_hgraph *g, int32_t nb, int32_t n_ins, int32_t n_del, uint32_t wmin,
                               uint32_t wmax, uint64_t seed, ddp_updates **out)
{
    if (!g || !out || nb < 0 || n_ins < 0 || n_del < 0 || wmin > wmax) return DDP_ERR_ARG;
    Rng r(seed);
    // live edge instances of the current graph version
    std::vector<std::pair<int32_t, int32_t>> live;
    live.reserve((size_t)g->E + (size_t)nb * n_ins);
    for (int32_t u = 0; u < g->N; ++u) {
        const int32_t b = g->off[u], e = (u + 1 < g->N) ? g->off[u + 1] : g->E;
        for (int32_t j = b; j < e; ++j)
            if (g->col[j] >= 0) live.emplace_back(u, g->col[j]);
    }
    ddp_updates *U = new ddp_updates();
    U->start.assign(1, 0);
    for (int32_t bt = 0; bt < nb; ++bt) {
        for (int32_t k = 0; k < n_del && !live.empty(); ++k) {
            const size_t idx = (size_t)r.below(live.size());
            U->lines.push_back(ddp_update_t{0, live[idx].first, live[idx].second, 0});
            live[idx] = live.back();
            live.pop_back();
        }
        std::vector<std::pair<int32_t, int32_t>> added;
        for (int32_t k = 0; k < n_ins; ++k) {
            const int32_t u = (int32_t)r.below((uint64_t)g->N);
            int32_t v = (int32_t)r.below((uint64_t)g->N - 1);
            if (v >= u) ++v;
            const uint32_t wt = wmin + (uint32_t)r.below((uint64_t)wmax - wmin + 1);
            U->lines.push_back(ddp_update_t{1, u, v, wt});
            added.emplace_back(u, v);
        }
        live.insert(live.end(), added.begin(), added.end());
        U->start.push_back((int32_t)U->lines.size());
    }
    *out = U;
    return DDP_OK;
}

extern "C" int ddp_gen_geo_pairs(const ddp_hgraph *g, int32_t nq, double dmin, double dmax, uint64_t seed,
                                 int32_t *src, int32_t *dst)
{
    if (!g || g->xy.empty() || nq < 0 || !src || !dst || dmin <= 0 || dmax < dmin) return DDP_ERR_ARG;
    const int32_t N = g->N;
    const auto &xy = g->xy;
    int32_t scale = 0;
    for (int32_t v : xy) scale = std::max(scale, v);
    // connected components (edges are symmetric): union-find
    std::vector<int32_t> par(N);
    std::iota(par.begin(), par.end(), 0);
    auto find = [&](int32_t x) {
        while (par[x] != x) {
            par[x] = par[par[x]];
            x = par[x];
        }
        return x;
    };
    for (int32_t u = 0; u < N; ++u) {
        const int32_t b = g->off[u], e = (u + 1 < N) ? g->off[u + 1] : g->E;
        for (int32_t j = b; j < e; ++j)
            if (g->col[j] >= 0) {
                const int32_t a = find(u), c = find(g->col[j]);
                if (a != c) par[a] = c;
            }
    }
    // cell index for nearest-node lookups
    const int32_t cs = std::max(1, (int32_t)std::sqrt((double)scale * scale / std::max(1, N)) * 2);
    const int32_t cells = scale / cs + 1;
    std::vector<int32_t> cstart((size_t)cells * cells + 1, 0), cpts(N);
    auto cell_of = [&](int32_t i) { return (xy[2 * (size_t)i + 1] / cs) * cells + xy[2 * (size_t)i] / cs; };
    for (int32_t i = 0; i < N; ++i) cstart[cell_of(i) + 1]++;
    for (size_t c = 0; c < (size_t)cells * cells; ++c) cstart[c + 1] += cstart[c];
    {
        std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
        for (int32_t i = 0; i < N; ++i) cpts[fill[cell_of(i)]++] = i;
    }
    auto nearest = [&](int64_t px, int64_t py) {
        const int32_t cx = (int32_t)(px / cs), cy = (int32_t)(py / cs);
        int32_t best = -1;
        int64_t bd = INT64_MAX;
        for (int32_t ring = 0; ring <= cells; ++ring) {
            for (int32_t ay = cy - ring; ay <= cy + ring; ++ay)
                for (int32_t ax = cx - ring; ax <= cx + ring; ++ax) {
                    if (std::max(std::abs(ax - cx), std::abs(ay - cy)) != ring) continue;
                    if (ax < 0 || ay < 0 || ax >= cells || ay >= cells) continue;
                    const size_t c = (size_t)ay * cells + ax;
                    for (int32_t p = cstart[c]; p < cstart[c + 1]; ++p) {
                        const int32_t v = cpts[p];
                        const int64_t dx = xy[2 * (size_t)v] - px, dy = xy[2 * (size_t)v + 1] - py;
                        const int64_t d = dx * dx + dy * dy;
                        if (d < bd || (d == bd && v < best)) {
                            bd = d;
                            best = v;
                        }
                    }
                }
            // every unvisited cell is at least ring*cs away
            if (best >= 0 && (int64_t)ring * cs * ring * cs > bd) break;
        }
        return best;
    };
    Rng r(seed);
    const double lr = std::log(dmax / dmin);
    for (int32_t q = 0; q < nq; ++q) {
        for (int tries = 0;; ++tries) {
            if (tries > 100000) return DDP_ERR_STATE;
            const int32_t s = (int32_t)r.below((uint64_t)N);
            const double d = dmin * std::exp(lr * r.unit());
            const double th = 2.0 * M_PI * r.unit();
            const double px = xy[2 * (size_t)s] + d * std::cos(th), py = xy[2 * (size_t)s + 1] + d * std::sin(th);
            if (px < 0 || py < 0 || px > scale || py > scale) continue;
            const int32_t t = nearest((int64_t)std::llround(px), (int64_t)std::llround(py));
            if (t < 0 || t == s || find(s) != find(t)) continue;
            src[q] = s;
            dst[q] = t;
            break;
        }
    }
    return DDP_OK;
}

extern "C" int ddp_gen_reachable_pairs(const ddp_hgraph *g, int32_t nq, uint64_t seed, int32_t *src, int32_t *dst)
{
    if (!g || nq < 0 || !src || !dst || g->N < 2) return DDP_ERR_ARG;
    Rng r(seed);
    std::vector<int32_t> mark(g->N, 0);
    int32_t stamp = 0;
    for (int32_t q = 0; q < nq; ++q) {
        for (int tries = 0;; ++tries) {
            if (tries > 100000) return DDP_ERR_STATE;
            const int32_t s = (int32_t)r.below((uint64_t)g->N);
            const int32_t t = (int32_t)r.below((uint64_t)g->N);
            if (s == t) continue;
            if (!reachable(g, s, t, mark, ++stamp)) continue;
            src[q] = s;
            dst[q] = t;
            break;
        }
    }
    return DDP_OK;
}

// Pairs through a hub: h = the node of largest out-degree; one forward search
// from h and one over reversed edges give In (nodes that reach h) and Out
// (nodes h reaches); s uniform in In and t uniform in Out (s != t) are
// connected through h.  Two graph traversals in total, where drawing uniform
// pairs and testing each one costs a traversal per rejected pair (ruinous on
// R-MAT graphs, whose many isolated nodes make most uniform pairs
// unreachable).
extern "C" int ddp_gen_hub_pairs(const ddp_hgraph *g, int32_t nq, uint64_t seed, int32_t *src, int32_t *dst)
{
    if (!g || nq < 0 || !src || !dst || g->N < 2) return DDP_ERR_ARG;
    const int32_t N = g->N;
    auto row_end = [&](int32_t u) { return (u + 1 < N) ? g->off[u + 1] : g->E; };
    int32_t hub = 0, best = -1;
    for (int32_t u = 0; u < N; ++u)
        if (row_end(u) - g->off[u] > best) {
            best = row_end(u) - g->off[u];
            hub = u;
        }
    // forward reach of the hub
    std::vector<char> fw(N, 0), bw(N, 0);
    std::vector<int32_t> q;
    q.reserve(N);
    q.push_back(hub);
    fw[hub] = 1;
    for (size_t i = 0; i < q.size(); ++i)
        for (int32_t j = g->off[q[i]]; j < row_end(q[i]); ++j) {
            const int32_t c = g->col[j];
            if (c >= 0 && !fw[c]) {
                fw[c] = 1;
                q.push_back(c);
            }
        }
    // reversed edges, then the nodes that reach the hub
    std::vector<int32_t> roff(N + 1, 0), rsrc;
    for (int32_t j = 0; j < g->E; ++j)
        if (g->col[j] >= 0) roff[g->col[j] + 1]++;
    for (int32_t i = 0; i < N; ++i) roff[i + 1] += roff[i];
    rsrc.resize(roff[N]);
    {
        std::vector<int32_t> fill(roff.begin(), roff.end() - 1);
        for (int32_t u = 0; u < N; ++u)
            for (int32_t j = g->off[u]; j < row_end(u); ++j)
                if (g->col[j] >= 0) rsrc[fill[g->col[j]]++] = u;
    }
    q.clear();
    q.push_back(hub);
    bw[hub] = 1;
    for (size_t i = 0; i < q.size(); ++i)
        for (int32_t j = roff[q[i]]; j < roff[q[i] + 1]; ++j) {
            const int32_t p = rsrc[j];
            if (!bw[p]) {
                bw[p] = 1;
                q.push_back(p);
            }
        }
    std::vector<int32_t> ins, outs;
    for (int32_t v = 0; v < N; ++v) {
        if (bw[v]) ins.push_back(v);
        if (fw[v]) outs.push_back(v);
    }
    if (ins.size() + outs.size() < 3 || (ins.size() == 1 && outs.size() == 1)) return DDP_ERR_STATE;
    Rng r(seed);
    for (int32_t k = 0; k < nq; ++k) {
        for (int tries = 0;; ++tries) {
            if (tries > 1000) return DDP_ERR_STATE;
            const int32_t s = ins[r.below(ins.size())], t = outs[r.below(outs.size())];
            if (s == t) continue;
            src[k] = s;
            dst[k] = t;
            break;
        }
    }
    return DDP_OK;
}
