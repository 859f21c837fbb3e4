// Routing of the server's low-load grid mode (server.hip): pure host-side
// functions of what the host knows -- the requests in flight, moving averages
// of finished queries' device times, and the gaps between submissions.  No
// HIP dependency, so the policy is also built and checked on the CPU.
#pragma once

#include <algorithm>
#include <cstdint>

// Requests without a result yet: queries and updates on the slot path, and
// requests handled by kernels of their own (grid-mode queries, updates applied
// directly); slot_age: how long ago the most recently submitted slot-path
// query in flight was submitted (ms; 0 when slot_q is 0).
struct RouteInFlight {
    int64_t slot_q = 0, slot_u = 0, direct = 0;
    double slot_age = 0;
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

// Grid mode for the next query?  A query routed there waits for every
// running slot query (the serve kernel is stopped and its slots finish first)
// and for the grid queries ahead of it, then takes one grid time G; in a slot
// it would take one slot time S from now.
//  * never while an update is in flight on the slot path: the grid search
//    reads the graph without waiting for versions, and that update might not
//    be applied before it runs;
//  * always when nothing is in flight: alone on the GPU a query finishes
//    sooner in grid mode (a caller sets grid_k only where that holds, and
//    where it pays one query alone finishes many times sooner there);
//  * before G and the arrival rate are measured, only behind at most one grid
//    query and with nothing on the slot path;
//  * otherwise while arrivals leave grid mode room and the grid queue plus
//    this query, (direct + 1) G, ends before a slot query would.  With slot
//    queries running, the youngest of them ends about S after its submission,
//    so the test is (direct + 1) G < S - (S - slot_age) = slot_age, which
//    needs no S (capped at S once that query has run longer than S); with none
//    running it is (direct + 1) G < S, or, before any slot query has finished,
//    at most one grid query ahead.  A query sent to a slot for want of S would
//    cost a full slot time and hold up every grid query after it.
//    "Room" has hysteresis: grid mode is entered while a grid query takes
//    under 70% of the mean gap between submissions and left once it takes
//    more than the whole gap, so Poisson noise in the long-run mean seldom
//    crosses either threshold: each flip costs the queries that straddle it
//    (a slot query waits for the grid queue, a grid query for the running
//    slots).
inline bool route_grid(RouteState &r, const RouteInFlight &f)
{
    if (f.slot_u) return false;
    if (f.slot_q == 0 && f.direct == 0) return true;
    if (r.grid_ms <= 0 || r.gap_ms <= 0) return f.slot_q == 0 && f.direct <= 1;
    const double rho = r.grid_ms / r.gap_ms;
    r.grid_regime = r.grid_regime ? rho <= 1.0 : rho < 0.7;
    if (!r.grid_regime) return false;
    const double queue = (double)(f.direct + 1) * r.grid_ms;
    if (f.slot_q) return queue < (r.slot_ms > 0 ? std::min(f.slot_age, r.slot_ms) : f.slot_age);
    return r.slot_ms > 0 ? queue < r.slot_ms : f.direct <= 1;
}

// An update goes to a kernel of its own only while nothing is in flight on
// the slot path: a running slot kernel would otherwise have to be stopped and
// drained for it.
inline bool route_direct_update(const RouteInFlight &f) { return f.slot_q == 0 && f.slot_u == 0; }
