// Routing of the server's low-load grid mode (server.hip): pure host-side
// functions of what the host knows -- the requests in flight, a moving average
// of finished grid-mode queries' device times, and the gaps between
// submissions.  No
// HIP dependency, so the policy is also built and checked on the CPU.
#pragma once

#include <algorithm>
#include <cstdint>

// Requests without a result yet: queries and updates on the slot path, and
// requests handled by kernels of their own (grid-mode queries, updates applied
// directly); slot_age: how long ago the most recently submitted slot-path
// query in flight was submitted (ms; 0 when slot_q is 0); slots_open: a
// serve launch is running and has not been told to stop.
struct RouteInFlight {
    int64_t slot_q = 0, slot_u = 0, direct = 0;
    double slot_age = 0;
    bool slots_open = false;
};

// Gaps between submissions measured before the arrival rate counts: the
// first few gaps of a Poisson stream are too noisy to set the regime by.
constexpr int64_t kRouteMinGaps = 8;

struct RouteState {
    double grid_ms = 0;                      // moving average of grid-mode queries' device times (0: none yet)
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

// A grid-mode query finished without error in device_ms.
inline void route_note_grid(RouteState &r, double device_ms)
{
    r.grid_ms = r.grid_ms > 0 ? 0.75 * r.grid_ms + 0.25 * device_ms : device_ms;
}

// Grid mode for the next query?  Grid-mode queries run one after another on
// the server's stream, after the slot queries already running: routing one
// there stops the serve kernel, whose slots finish their requests and claim
// nothing published after it.  So a query sent to a slot while grid queries
// are queued starts after them as well.  With G a grid query's time and S a
// slot query's (many G where grid_k is set):
//  * never while an update is in flight on the slot path: the grid search
//    reads the graph without waiting for versions, and that update might not
//    be applied before it runs;
//  * always when nothing is in flight: alone on the GPU a query finishes
//    sooner in grid mode (a caller sets grid_k only where that holds);
//  * before G and kRouteMinGaps gaps are measured, only with nothing on the
//    slot path and at most one grid query ahead;
//  * otherwise only while arrivals leave grid mode room, and then
//    - with grid queries queued: in a slot the query would wait for the same
//      queue and then take S > G;
//    - with slot queries running and no grid query queued: it waits for them,
//      and the youngest ends about S after its submission, so it finishes at
//      about S - slot_age + G, against S in a slot: grid once slot_age > G.
//    "Room" has hysteresis: grid mode is entered while G is under 70% of the
//    mean gap between submissions and left once it exceeds the whole gap (a
//    single queue of grid queries then keeps up with arrivals), and Poisson
//    noise in the long-run mean seldom crosses either threshold: each flip
//    costs the queries that straddle it (a slot query waits for the grid
//    queue, a grid query for the running slots).
inline bool route_grid(RouteState &r, const RouteInFlight &f)
{
    if (f.slot_u) return false;
    if (f.slot_q == 0 && f.direct == 0) return true;
    if (r.grid_ms <= 0 || r.gaps < kRouteMinGaps) return f.slot_q == 0 && f.direct <= 1;
    const double rho = r.grid_ms / r.gap_ms;
    r.grid_regime = r.grid_regime ? rho <= 1.0 : rho < 0.7;
    if (!r.grid_regime) return false;
    if (f.direct) return true;
    return f.slot_age > r.grid_ms;
}

// An update goes to a kernel of its own unless an update is pending on the
// slot path (versions are applied in order) or an open serve launch runs slot
// queries (it would have to be stopped and drained for it).  Once the launch
// has been stopped for grid work, a slot update would wait for that queue and
// a relaunch, and while it is pending no query could go to grid mode.
inline bool route_direct_update(const RouteInFlight &f)
{
    return f.slot_u == 0 && (f.slot_q == 0 || !f.slots_open);
}
