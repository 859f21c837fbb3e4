// a_star_cpu: the process contract of GA_star/A_star_cpu.cpp:100-204, answered
// on the GPU.
//   stdin   "s t"
//   files   graph.txt (CSR), Hx.txt (N ints, -1 = cannot reach t)
//   stdout  "completed input", then "cost C" and the path from t back to s,
//           every node followed by a space (A_star_cpu.cpp:190-199), or
//           "unreachable !!" (A_star_cpu.cpp:201-203).
// The search is libddp.so's single-query GA* (ddp_astar); nothing here runs on
// the host but parsing and printing.  Costs equal the reference's for an
// admissible Hx.txt; among several shortest paths the one printed may differ.
//
// --row-end-quirk reproduces the reference's loop bound for the last node
// (A_star_cpu.cpp:169-172): node N-1's row ends at E-1, so edge E-1 is never
// relaxed when it belongs to that row.  Off by default, as on the reference's
// GPU path (GA_star/GA_v2.cu:110-112).
// DDP_K, DDP_MODE (ddp_astar_opts_t.mode), DDP_BLOCK and DDP_HEAP_CAP override
// the search options.
#include "common.h"

int main(int argc, char **argv)
{
    bool quirk = false;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--row-end-quirk") == 0) quirk = true;
        else drv::die("usage", "a_star_cpu [--row-end-quirk] < \"s t\"");
    }
    int s = 0, t = 0;
    if (std::scanf("%d %d", &s, &t) != 2) drv::die("stdin", "expected \"s t\"");
    drv::HostGraph hg;
    drv::load_graph("graph.txt", hg);
    std::vector<int32_t> hx = drv::load_hx("Hx.txt", hg.N);
    std::printf("completed input\n");
    if (s < 0 || s >= hg.N || t < 0 || t >= hg.N) drv::die("stdin", "s or t outside [0, N)");
    std::vector<int32_t> col(hg.col, hg.col + hg.E);
    if (quirk && hg.E > 0 && hg.off[hg.N - 1] <= hg.E - 1) col[hg.E - 1] = -1;  // a deleted edge to the kernels
    ddp_ctx *ctx = drv::open_ctx();
    ddp_graph *g = nullptr;
    drv::check(ctx, ddp_graph_upload(ctx, hg.N, hg.E, hg.off, col.data(), hg.w, 0, &g), "ddp_graph_upload");
    drv::check(ctx, ddp_graph_set_h(g, hx.data()), "ddp_graph_set_h");
    ddp_astar_opts_t o = drv::opts_with_K(0);
    if (const char *k = std::getenv("DDP_K")) o.K = std::atoi(k);
    if (const char *m = std::getenv("DDP_MODE")) o.mode = std::atoi(m);
    ddp_heur_t h{DDP_H_ARRAY, 1};
    int32_t cost = 0, len = 0;
    std::vector<int32_t> path(hg.N);
    drv::check(ctx, ddp_astar(ctx, g, h, s, t, &o, &cost, nullptr, path.data(), hg.N, &len, nullptr), "ddp_astar");
    if (cost != DDP_INF_COST && len > 0) {
        std::printf("cost %d\n", cost);
        for (int32_t i = len - 1; i >= 0; --i) std::printf("%d ", path[i]);
        std::printf("\n");
    } else {
        std::printf("unreachable !!\n");
    }
    ddp_graph_destroy(g);
    ddp_ctx_destroy(ctx);
    return 0;
}
