// Routing of the server's low-load grid mode (server.hip): pure host-side
// functions of what the host knows -- the requests in flight, moving averages
// of finished queries' device times, and the gaps between submissions.  No
// HIP dependency, so the policy is also built and checked on the CPU.
#pragma once

#include <algorithm>
#include <cstdint>

// Requests without a result yet: queries and updates on the slot path, and
// requests handled by kernels of their own (grid-mode queries, updates applied
// directly); remain: the longest a slot-path query is expected to run on, from
// its submission and the slots' moving average (ms).
struct RouteInFlight {
    int64_t slot_q = 0, slot_u = 0, direct = 0;
    double remain = 0;
};

struct RouteState {
    double grid_ms = 0, slot_ms = 0;         // moving averages of finished queries' device times (0: none yet)
    double gap_ms = 0, last_submit_ms = -1;  // long-run mean gap between submissions
    int64_t gaps = 0;
    bool grid_regime = false;                // arrivals leave grid mode room (with hysteresis)
};

// A submission at host time now_ms.  The gap mean is the running mean of the
// first 50 gaps, then a moving average of weight 1/50.
inline void route_note_submit(RouteState &r, double now_ms)
{
    if (r.last_submit_ms >= 0) {
        const double wgt = std::max(0.02, 1.0 / (double)++r.gaps);
        r.gap_ms += wgt * ((now_ms - r.last_submit_ms) - r.gap_ms);
    }
    r.last_submit_ms = now_ms;
}

// A query finished without error in device_ms, in grid mode or in a slot.
inline void route_note_query(RouteState &r, bool grid, double device_ms)
{
    double &avg = grid ? r.grid_ms : r.slot_ms;
    avg = avg > 0 ? 0.75 * avg + 0.25 * device_ms : device_ms;
}

// Grid mode for the next query?
//  * never while an update is in flight on the slot path: the grid search
//    reads the graph without waiting for versions, and that update might not
//    be applied before it runs;
//  * always when nothing is in flight: alone on the GPU a query finishes
//    sooner in grid mode (a caller sets grid_k only where that holds);
//  * before any query has finished in a slot, while nothing runs on the
//    slot path and at most one grid query is in flight: with no slot time to
//    weigh the queue against, a query may wait behind one grid query (where
//    grid mode pays, one query alone finishes many times sooner there than in
//    a slot).  Sending it to a slot instead would cost a full slot time, and
//    the queries after it would follow, since a grid query waits for the
//    running slots;
//  * otherwise only with measurements to go by, while arrivals leave grid
//    mode room, and if the slot queries still running plus the grid queries
//    in flight plus this one are expected to finish before one slot query
//    would.  "Room" has hysteresis: grid mode is entered while a grid query
//    takes under 70% of the mean gap between submissions and left once it
//    takes more than the whole gap, so Poisson noise in the long-run mean
//    seldom crosses either threshold: each flip costs the queries that
//    straddle it (a slot query waits for the grid queue, a grid query for
//    the running slots).
inline bool route_grid(RouteState &r, const RouteInFlight &f)
{
    if (f.slot_u) return false;
    if (f.slot_q == 0 && f.direct == 0) return true;
    if (r.slot_ms <= 0) return f.slot_q == 0 && f.direct <= 1;
    if (r.grid_ms <= 0 || r.gap_ms <= 0) return false;
    const double rho = r.grid_ms / r.gap_ms;
    r.grid_regime = r.grid_regime ? rho <= 1.0 : rho < 0.7;
    return r.grid_regime && f.remain + (double)(f.direct + 1) * r.grid_ms < r.slot_ms;
}

// An update goes to a kernel of its own only while nothing is in flight on
// the slot path: a running slot kernel would otherwise have to be stopped and
// drained for it.
inline bool route_direct_update(const RouteInFlight &f) { return f.slot_q == 0 && f.slot_u == 0; }
