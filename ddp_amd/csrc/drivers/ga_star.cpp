// ga_star: the GA* driver contract of GA_star/GA_v2.cu:298-577.
//   stdin   "K" then "s t"
//   files   graph.txt (CSR), Hx.txt (N ints, -1 = cannot reach t)
//   stdout  "completed input", "cost: C", then the path from t back to s
//           (GA_v2.cu:561-575) or "not found".
// The search runs in libddp.so's single-query GA* kernel on the GPU.
#include "common.h"

int main()
{
    int K = 0, s = 0, t = 0;
    if (std::scanf("%d", &K) != 1 || std::scanf("%d %d", &s, &t) != 2) drv::die("stdin", "expected \"K\" then \"s t\"");
    drv::HostGraph hg;
    drv::load_graph("graph.txt", hg);
    std::vector<int32_t> hx = drv::load_hx("Hx.txt", hg.N);
    std::printf("completed input\n");
    if (s < 0 || s >= hg.N || t < 0 || t >= hg.N) drv::die("stdin", "s or t outside [0, N)");
    ddp_ctx *ctx = drv::open_ctx();
    ddp_graph *g = nullptr;
    drv::check(ctx, ddp_graph_upload(ctx, hg.N, hg.E, hg.off, hg.col, hg.w, 0, &g), "ddp_graph_upload");
    drv::check(ctx, ddp_graph_set_h(g, hx.data()), "ddp_graph_set_h");
    ddp_astar_opts_t o = drv::opts_with_K(K);
    ddp_heur_t h{DDP_H_ARRAY, 1};
    int32_t cost = 0, len = 0;
    std::vector<int32_t> path(hg.N);
    ddp_astar_stats_t st;
    drv::check(ctx, ddp_astar(ctx, g, h, s, t, &o, &cost, nullptr, path.data(), hg.N, &len, &st), "ddp_astar");
    std::printf("cost: %d\n", cost);
    if (cost != DDP_INF_COST && len > 0) {
        for (int32_t i = len - 1; i > 0; --i) std::printf("%d ", path[i]);
        std::printf("%d\n", path[0]);
    } else {
        std::printf("not found\n");
    }
    if (std::getenv("DDP_STATS"))
        std::fprintf(stderr, "[INFO] K=%d iterations=%d expansions=%lld kernel_ms=%.3f\n", st.K, st.iterations,
                     (long long)st.expansions, st.kernel_ms);
    ddp_graph_destroy(g);
    ddp_ctx_destroy(ctx);
    return 0;
}
