// Shared helpers of the command-line drivers (reference process contract:
// stdin arguments, graph.txt / Hx.txt / Updates.txt in the working directory,
// results on stdout).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/ddp_capi.h"
#include "../../../include/ddp_dastar.h"
#include "../../../include/ddp_io.h"

namespace drv {

[[noreturn]] inline void die(const char *what, const char *msg, int code = 1)
{
    std::fprintf(stderr, "%s: %s\n", what, msg);
    std::exit(code);
}

inline ddp_ctx *open_ctx()
{
    ddp_ctx *ctx = nullptr;
    if (ddp_ctx_create(0, &ctx) != DDP_OK) die("ddp_ctx_create", ddp_create_error(), 2);
    return ctx;
}

inline void check(ddp_ctx *ctx, int rc, const char *what)
{
    if (rc != DDP_OK) die(what, ctx ? ddp_last_error(ctx) : "error", 3);
}

struct HostGraph {
    ddp_hgraph *h = nullptr;
    int32_t N = 0, E = 0;
    const int32_t *off = nullptr, *col = nullptr;
    const uint32_t *w = nullptr;
    ~HostGraph()
    {
        if (h) ddp_hgraph_free(h);
    }
};

// DDP_GRAPH_CACHE=1 reads graph.txt through the binary cache graph.txt.ddpbin
// (ddp_read_graph_cached), written on the first run and reused while the text
// file is unchanged.
inline void load_graph(const char *path, HostGraph &g)
{
    char err[512] = {0};
    const char *c = std::getenv("DDP_GRAPH_CACHE");
    const int rc = (c && std::atoi(c)) ? ddp_read_graph_cached(path, nullptr, &g.h, err, sizeof err)
                                       : ddp_read_graph_txt(path, &g.h, err, sizeof err);
    if (rc != DDP_OK) die(path, err, 4);
    ddp_hgraph_view(g.h, &g.N, &g.E, &g.off, &g.col, &g.w);
}

inline std::vector<int32_t> load_hx(const char *path, int32_t N)
{
    std::vector<int32_t> h(N);
    char err[512] = {0};
    if (ddp_read_hx_txt(path, N, h.data(), err, sizeof err) != DDP_OK) die(path, err, 4);
    return h;
}

inline double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline ddp_astar_opts_t opts_with_K(int32_t K)
{
    ddp_astar_opts_t o;
    std::memset(&o, 0, sizeof o);
    o.K = K;
    if (const char *b = std::getenv("DDP_BLOCK")) o.block = std::atoi(b);
    if (const char *c = std::getenv("DDP_HEAP_CAP")) o.heap_cap = std::atoi(c);
    return o;
}

}  // namespace drv
