// bfs_hx: the heuristic generator contract of GA_star/bfs_hx.cu:39-177.
//   stdin   "t"
//   files   graph.txt (the reference feeds the TRANSPOSED graph here)
//   output  Hx.txt: hop count per node, -1 if unreached ("%d " each)
//   stdout  "completed input", "max-level: L", "time: ms",
//           "unreachable U , fraction", "farthest node: v  val: d"
// Options (argv): --scale S writes hops*S; --transpose reads the forward
// graph and walks its transposed rows (so graph_op.txt is not needed).
#include "common.h"

int main(int argc, char **argv)
{
    int scale = 1;
    bool transpose = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--scale") && i + 1 < argc) scale = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--transpose")) transpose = true;
        else drv::die("usage", "bfs_hx [--scale S] [--transpose] < t");
    }
    int t = 0;
    if (std::scanf("%d", &t) != 1) drv::die("stdin", "expected \"t\"");
    drv::HostGraph hg;
    drv::load_graph("graph.txt", hg);
    std::printf("completed input\n");
    if (t < 0 || t >= hg.N) drv::die("stdin", "t outside [0, N)");
    ddp_ctx *ctx = drv::open_ctx();
    ddp_graph *g = nullptr;
    drv::check(ctx, ddp_graph_upload(ctx, hg.N, hg.E, hg.off, hg.col, hg.w, transpose ? DDP_GRAPH_REVERSE : 0, &g),
               "ddp_graph_upload");
    std::vector<int32_t> h(hg.N);
    int32_t levels = 0;
    const double t0 = drv::now_ms();
    drv::check(ctx, ddp_bfs_hops(ctx, g, t, transpose ? 1 : 0, 1, 0, h.data(), &levels), "ddp_bfs_hops");
    const double ms = drv::now_ms() - t0;
    int max_level = 0, unreachable = 0, far_node = 0, far_val = 0;
    for (int32_t i = 0; i < hg.N; ++i) {
        if (h[i] > max_level) max_level = h[i];
        if (h[i] == -1) ++unreachable;
        if (h[i] >= far_val) {
            far_val = h[i];
            far_node = i;
        }
    }
    std::printf("max-level: %d\n", max_level);
    std::printf("time: %f\n", ms);
    if (scale != 1)
        for (auto &x : h)
            if (x >= 0) x *= scale;
    if (ddp_write_hx_txt("Hx.txt", hg.N, h.data()) != DDP_OK) drv::die("Hx.txt", "cannot write", 5);
    std::printf("unreachable %d , %f\n", unreachable, unreachable / (hg.N * 1.0));
    std::printf("farthest node: %d  val: %d\n", far_node, far_val);
    ddp_graph_destroy(g);
    ddp_ctx_destroy(ctx);
    return 0;
}
