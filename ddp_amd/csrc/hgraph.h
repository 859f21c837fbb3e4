// Host CSR and update-batch containers behind the opaque handles of ddp_io.h.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/ddp_io.h"

struct ddp_hgraph {
    int32_t N = 0, E = 0;
    std::vector<int32_t> off;  // N entries (reference layout, no sentinel)
    std::vector<int32_t> col;
    std::vector<uint32_t> w;
    std::vector<int32_t> xy;   // 2*N coordinates for geometric graphs
};

struct ddp_updates {
    std::vector<ddp_update_t> lines;
    std::vector<int32_t> start{0};  // batch b = lines[start[b], start[b+1])
};
