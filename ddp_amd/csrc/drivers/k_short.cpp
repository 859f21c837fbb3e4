// k_short: the alternative-path driver contract of k-short/GA_merge.cu:375-945.
//   stdin   "K" (queues), then "s t", then the number of further paths
//           (GA_merge.cu:378-388)
//   files   graph.txt, Hx.txt (-1 read as 0, as the DA* drivers read it)
//   stdout  "[OUT] Cost: C" and "[OUT] Path(in reverse): t ... s" (or
//           "not found") for the first search, then once per further path:
//           the edge into t on the current best path (parent(t) -> t) is
//           deleted -- its first live instance, GA_merge.cu:717-733 with
//           removeDelEdges -- and the search is repaired (DA*) instead of
//           recomputed; it stops early once t is unreachable, as
//           GA_merge.cu:730-731 does.  "[INFO] total run time N: ms" closes.
#include "common.h"

static bool print_result(ddp_dastar *d, int32_t cost, int32_t N, std::vector<int32_t> &path)
{
    std::printf("[OUT] Cost: %d\n", cost);
    std::printf("[OUT] Path(in reverse): ");
    int32_t len = 0;
    path.assign(N, 0);
    if (cost == DDP_INF_COST || ddp_dastar_path(d, path.data(), N, &len) != DDP_OK || len <= 0) {
        std::printf("not found\n");
        path.clear();
        return false;
    }
    path.resize(len);
    for (int32_t i = len - 1; i > 0; --i) std::printf("%d ", path[i]);
    std::printf("%d\n", path[0]);
    return true;
}

int main(int argc, char **argv)
{
    if (argc > 1) drv::die("usage", "k_short < \"K\\n s t\\n num_paths\"");
    int K = 0, s = 0, t = 0, num_paths = 0;
    if (std::scanf("%d", &K) != 1 || std::scanf("%d %d", &s, &t) != 2 || std::scanf("%d", &num_paths) != 1)
        drv::die("stdin", "expected \"K\", \"s t\", then the number of paths");
    drv::HostGraph hg;
    drv::load_graph("graph.txt", hg);
    std::vector<int32_t> hx = drv::load_hx("Hx.txt", hg.N);
    std::printf("[INFO] completed taking input\n");
    if (s < 0 || s >= hg.N || t < 0 || t >= hg.N) drv::die("stdin", "s or t outside [0, N)");
    ddp_ctx *ctx = drv::open_ctx();
    ddp_astar_opts_t o = drv::opts_with_K(K);
    const double t0 = drv::now_ms();
    ddp_dastar *d = nullptr;
    int32_t cost = 0;
    drv::check(ctx, ddp_dastar_open(ctx, hg.N, hg.E, hg.off, hg.col, hg.w, hx.data(), s, t, &o, &d, &cost),
               "ddp_dastar_open");
    std::vector<int32_t> path;
    bool found = print_result(d, cost, hg.N, path);
    int p = 0;
    for (; p < num_paths && found && path.size() >= 2; ++p) {
        const ddp_update_t del{0, path[path.size() - 2], t, 0};
        ddp_dastar_stats_t st;
        drv::check(ctx, ddp_dastar_apply(d, &del, 1, DDP_DASTAR_REPAIR, &cost, &st), "ddp_dastar_apply");
        std::printf("[INFO](%d) deleted %d -> %d\n", p, del.u, del.v);
        found = print_result(d, cost, hg.N, path);
    }
    std::printf("[INFO] total run time %d: %f\n", p, drv::now_ms() - t0);
    ddp_dastar_close(d);
    ddp_ctx_destroy(ctx);
    return 0;
}
