// Launch-shape autotuner (see include/ddp_tune.h): times candidate shapes on
// a sample of the caller's own queries and caches the winner per graph.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/ddp_tune.h"

namespace {

std::string cache_path()
{
    if (const char *p = std::getenv("DDP_TUNE_CACHE")) return p;
    const char *home = std::getenv("HOME");
    if (!home) return "";
    std::string dir = std::string(home) + "/.cache";
    mkdir(dir.c_str(), 0755);
    dir += "/ddp_amd";
    mkdir(dir.c_str(), 0755);
    return dir + "/tune.txt";
}

std::string key(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, const char *what, int32_t nq)
{
    int32_t N = 0, E = 0, md = 0;
    uint32_t mw = 0;
    ddp_graph_info(g, &N, &E, &mw, &md);
    ddp_device_info_t di;
    std::memset(&di, 0, sizeof di);
    ddp_device_info(ctx, &di);
    std::ostringstream k;
    k << what << ":" << di.arch << ":" << N << ":" << E << ":" << md << ":" << mw << ":" << h.kind << ":" << nq;
    return k.str();
}

bool cache_get(const std::string &k, std::vector<long> &vals)
{
    const std::string p = cache_path();
    if (p.empty()) return false;
    std::ifstream f(p);
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string kk;
        ss >> kk;
        if (kk != k) continue;
        vals.clear();
        long v;
        while (ss >> v) vals.push_back(v);
        found = true;  // last entry wins
    }
    return found;
}

void cache_put(const std::string &k, const std::vector<long> &vals)
{
    const std::string p = cache_path();
    if (p.empty()) return;
    std::ofstream f(p, std::ios::app);
    if (!f) return;
    f << k;
    for (long v : vals) f << " " << v;
    f << "\n";
}

// Every stride-th query: long and short searches stay represented.
void sample_queries(int32_t nq, const int32_t *src, const int32_t *dst, int32_t sample, std::vector<int32_t> &s,
                    std::vector<int32_t> &t)
{
    const int32_t n = std::max(1, std::min(nq, sample));
    const double stride = (double)nq / n;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t k = std::min(nq - 1, (int32_t)(i * stride));
        s.push_back(src[k]);
        t.push_back(dst[k]);
    }
}

// Best of three timed launches of a batch with options o on the sample (the
// first launch warms up); -1 if the options do not fit or a query fails.
// Costs of the last launch go to costs (n entries) when given.
float time_batch(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, const ddp_batch_opts_t &o,
                 const std::vector<int32_t> &s, const std::vector<int32_t> &t, std::vector<int32_t> *costs)
{
    const int32_t n = (int32_t)s.size();
    ddp_batch *b = nullptr;
    if (ddp_batch_create(ctx, g, h, n, &o, &b) != DDP_OK) return -1.f;
    float ms = 1e30f;
    bool ok = ddp_batch_upload(b, n, s.data(), t.data()) == DDP_OK;
    for (int rep = 0; ok && rep < 3; ++rep) {
        ddp_batch_stats_t st;
        ok = ddp_batch_launch(b) == DDP_OK && ddp_batch_wait(b, &st) == DDP_OK && st.errors == 0;
        if (ok && rep > 0) ms = std::min(ms, st.kernel_ms);
    }
    if (ok && costs) {
        costs->assign(n, 0);
        ok = ddp_batch_download(b, costs->data(), nullptr, nullptr, nullptr, nullptr) == DDP_OK;
    }
    ddp_batch_destroy(b);
    return ok ? n / (ms * 1e-3f) : -1.f;
}

}  // namespace

extern "C" int ddp_tune_batch(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq, const int32_t *src,
                              const int32_t *dst, int32_t sample, ddp_batch_opts_t *best, float *qps)
{
    if (!ctx || !g || !best || nq <= 0 || !src || !dst) return DDP_ERR_ARG;
    // the key names every option the tuner varies, so entries of older tuners never match
    const std::string k = key(ctx, g, h, "batch:block,engine,heap_cap,lds_cap", std::min(nq, sample));
    std::vector<long> cached;
    if (cache_get(k, cached) && cached.size() >= 5) {
        std::memset(best, 0, sizeof *best);
        best->block = (int32_t)cached[0];
        best->engine = (int32_t)cached[1];
        best->heap_cap = (int32_t)cached[2];
        best->lds_cap = (int32_t)cached[3];
        if (qps) *qps = (float)cached[4];
        return DDP_OK;
    }
    std::vector<int32_t> s, t;
    sample_queries(nq, src, dst, sample, s, t);
    int32_t N = 0;
    ddp_graph_info(g, &N, nullptr, nullptr, nullptr);
    // (block, engine, heap_cap, lds_cap); 0 = the library's default.  Heap
    // depth: the heap engine's LDS levels per lane and its HBM entries per
    // lane; the pile engines' queue memory per lane (their default holds an
    // exhaustive search, 3N entries per slot and direction; half of it can
    // drop pushes into far-pile rebuilds on long searches).
    struct Cand {
        int32_t B, eng, cap, lcap;
    };
    std::vector<Cand> cand = {{256, DDP_BATCH_PILES, 0, 0},  {512, DDP_BATCH_PILES, 0, 0},
                              {1024, DDP_BATCH_PILES, 0, 0}, {1024, DDP_BATCH_PILES, std::max(2048, (3 * N / 2) / 1024), 0},
                              {512, DDP_BATCH_HEAPS, 0, 8},  {512, DDP_BATCH_HEAPS, 0, 16},
                              {512, DDP_BATCH_HEAPS, 4096, 16}};
    const bool closed_form = h.kind != DDP_H_ARRAY;
    if (closed_form) {  // needs transposed rows; create refuses the candidate without them
        cand.push_back({512, DDP_BATCH_BIDI, 0, 0});
        cand.push_back({1024, DDP_BATCH_BIDI, 0, 0});
    }
    float best_qps = -1.f;
    Cand win{256, DDP_BATCH_PILES, 0, 0};
    for (const Cand &c : cand) {
        ddp_batch_opts_t o;
        std::memset(&o, 0, sizeof o);
        o.block = c.B;
        o.engine = c.eng;
        o.heap_cap = c.cap;
        o.lds_cap = c.lcap;
        const float q = time_batch(ctx, g, h, o, s, t, nullptr);
        if (q > best_qps) {
            best_qps = q;
            win = c;
        }
    }
    if (best_qps < 0) return DDP_ERR_STATE;
    std::memset(best, 0, sizeof *best);
    best->block = win.B;
    best->engine = win.eng;
    best->heap_cap = win.cap;
    best->lds_cap = win.lcap;
    if (qps) *qps = best_qps;
    cache_put(k, {win.B, win.eng, win.cap, win.lcap, (long)best_qps});
    return DDP_OK;
}

extern "C" int ddp_tune_astar(ddp_ctx *ctx, const ddp_graph *g, ddp_heur_t h, int32_t nq, const int32_t *src,
                              const int32_t *dst, int32_t sample, ddp_astar_opts_t *best, float *mean_ms)
{
    if (!ctx || !g || !best || nq <= 0 || !src || !dst) return DDP_ERR_ARG;
    const std::string k = key(ctx, g, h, "astar:K,block,mode,heap_cap,lds_cap", std::min(nq, sample));
    std::vector<long> cached;
    if (cache_get(k, cached) && cached.size() >= 6) {
        std::memset(best, 0, sizeof *best);
        best->K = (int32_t)cached[0];
        best->block = (int32_t)cached[1];
        best->mode = (int32_t)cached[2];
        best->heap_cap = (int32_t)cached[3];
        best->lds_cap = (int32_t)cached[4];
        if (mean_ms) *mean_ms = cached[5] * 1e-3f;
        return DDP_OK;
    }
    std::vector<int32_t> s, t;
    sample_queries(nq, src, dst, sample, s, t);
    double best_ms = 1e300;
    ddp_astar_opts_t win;
    std::memset(&win, 0, sizeof win);
    // workgroup mode with 256..1024 queues, pile mode with 512/1024 lanes, the
    // search from both ends with 256..1024 lanes (closed-form heuristics; on a
    // graph without transposed rows ddp_astar refuses it before any launch and
    // the candidate drops out like any that fails), grid mode with 1024..65536
    // queues (a full search of a large power-law graph wants the most lanes:
    // K = 65,536 took 30 ms where 8,192 took 66 on R-MAT scale 23)
    struct Cand {
        int32_t K, B, mode;
    };
    std::vector<Cand> cands;
    for (int32_t K : {256, 512, 1024}) cands.push_back({K, 0, DDP_ASTAR_WORKGROUP});
    for (int32_t K : {512, 1024}) cands.push_back({K, 0, DDP_ASTAR_PILES});
    if (h.kind != DDP_H_ARRAY)
        for (int32_t K : {256, 512, 1024}) cands.push_back({K, 0, DDP_ASTAR_BIDI});
    for (int32_t K : {1024, 2048, 4096, 8192})
        for (int32_t B : {64, 256}) cands.push_back({K, B, DDP_ASTAR_GRID});
    for (int32_t K : {16384, 32768})
        for (int32_t B : {256, 1024}) cands.push_back({K, B, DDP_ASTAR_GRID});
    cands.push_back({65536, 1024, DDP_ASTAR_GRID});
    auto time_opts = [&](const ddp_astar_opts_t &o) {
        double total = 0;
        for (size_t i = 0; i < s.size(); ++i) {
            int32_t cost = 0;
            ddp_astar_stats_t st;
            if (ddp_astar(ctx, g, h, s[i], t[i], &o, &cost, nullptr, nullptr, 0, nullptr, &st) != DDP_OK) return 1e300;
            total += st.total_ms;
        }
        return total;
    };
    for (const Cand &c : cands) {
        ddp_astar_opts_t o;
        std::memset(&o, 0, sizeof o);
        o.K = c.K;
        o.block = c.B;
        o.mode = c.mode;
        const double total = time_opts(o);
        if (total < best_ms) {
            best_ms = total;
            win = o;
        }
    }
    if (best_ms >= 1e300) return DDP_ERR_STATE;
    // Heap depth of the winning shape: LDS levels per queue (grid and
    // workgroup modes; the rest spills to HBM) and HBM entries per queue.
    if (win.mode != DDP_ASTAR_PILES && win.mode != DDP_ASTAR_BIDI) {
        for (int32_t lc : {8, 16, 32}) {
            ddp_astar_opts_t o = win;
            o.lds_cap = lc;
            const double total = time_opts(o);
            if (total < best_ms) {
                best_ms = total;
                win = o;
            }
        }
    }
    *best = win;
    const double mean = best_ms / s.size();
    if (mean_ms) *mean_ms = (float)mean;
    cache_put(k, {win.K, win.block, win.mode, win.heap_cap, win.lds_cap, (long)(mean * 1e3)});
    return DDP_OK;
}
