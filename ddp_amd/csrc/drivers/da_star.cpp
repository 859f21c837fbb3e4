// da_star: the DA* driver contract of DA_star/DA_dyn_insert_del_nocycle.cu:1138-2119
// (and, with --recompute, of the per-batch recompute driver DA_star/GA_merge.cu).
//   stdin   "K" then "s t"
//   files   graph.txt, Hx.txt (-1 read as 0, NC:1202-1210), Updates.txt;
//           graph_op.txt is not needed (the transposed rows are built here)
//   stdout  "[OUT] Cost: C" and "[OUT] Path(in reverse): t ... s" (or
//           "not found") for the initial search and after every batch
//           (NC:1469-1481, 1995-2009), "[INFO]" progress lines, and
//           "[INFO] RUNTIME: ms" (time after input, as NC:1385 starts it).
#include "common.h"

static void print_result(ddp_dastar *d, int32_t cost, int32_t N)
{
    std::printf("[OUT] Cost: %d\n", cost);
    std::printf("[OUT] Path(in reverse): ");
    if (cost == DDP_INF_COST) {
        std::printf("not found\n");
        return;
    }
    std::vector<int32_t> p(N);
    int32_t len = 0;
    if (ddp_dastar_path(d, p.data(), N, &len) != DDP_OK || len <= 0) {
        std::printf("not found\n");
        return;
    }
    for (int32_t i = len - 1; i > 0; --i) std::printf("%d ", p[i]);
    std::printf("%d\n", p[0]);
}

int main(int argc, char **argv)
{
    bool recompute = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--recompute")) recompute = true;
        else drv::die("usage", "da_star [--recompute] < \"K\\n s t\"");
    }
    int K = 0, s = 0, t = 0;
    if (std::scanf("%d", &K) != 1 || std::scanf("%d %d", &s, &t) != 2) drv::die("stdin", "expected \"K\" then \"s t\"");
    drv::HostGraph hg;
    drv::load_graph("graph.txt", hg);
    std::vector<int32_t> hx = drv::load_hx("Hx.txt", hg.N);
    ddp_updates *upd = nullptr;
    char err[512] = {0};
    if (ddp_read_updates_txt("Updates.txt", &upd, err, sizeof err) != DDP_OK) drv::die("Updates.txt", err, 4);
    std::printf("[INFO] completed taking input\n");
    if (s < 0 || s >= hg.N || t < 0 || t >= hg.N) drv::die("stdin", "s or t outside [0, N)");
    ddp_ctx *ctx = drv::open_ctx();
    ddp_astar_opts_t o = drv::opts_with_K(K);
    const double t0 = drv::now_ms();
    ddp_dastar *d = nullptr;
    int32_t cost = 0;
    drv::check(ctx, ddp_dastar_open(ctx, hg.N, hg.E, hg.off, hg.col, hg.w, hx.data(), s, t, &o, &d, &cost),
               "ddp_dastar_open");
    print_result(d, cost, hg.N);
    const int nb = ddp_updates_count(upd);
    for (int b = 0; b < nb; ++b) {
        const ddp_update_t *lines = nullptr;
        int32_t n = 0;
        ddp_updates_batch(upd, b, &lines, &n);
        ddp_dastar_stats_t st;
        drv::check(ctx, ddp_dastar_apply(d, lines, n, recompute ? DDP_DASTAR_RECOMPUTE : DDP_DASTAR_REPAIR, &cost, &st),
                   "ddp_dastar_apply");
        std::printf("[INFO](%d) insertion:%d\n", b, st.inserts);
        std::printf("[INFO](%d) deletion:%d, eff del:%d\n", b, st.deletes_effective, st.roots);
        print_result(d, cost, hg.N);
    }
    std::printf("[INFO] update count: %d\n", nb);
    std::printf("[INFO] RUNTIME: %f\n", drv::now_ms() - t0);
    ddp_dastar_close(d);
    ddp_updates_free(upd);
    ddp_ctx_destroy(ctx);
    return 0;
}
