// The server's grid-mode routing policy (ddp_amd/csrc/serve_route.h) on the
// host: every rule of route_grid / route_direct_update, the hysteresis of the
// grid regime and the moving averages, on hand-built situations.
#include <cmath>
#include <cstdio>

#include "serve_route.h"

static int failures = 0;

static void expect(bool ok, const char *what)
{
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

static RouteInFlight flight(int64_t slot_q, int64_t slot_u, int64_t direct, double slot_age = 0)
{
    RouteInFlight f;
    f.slot_q = slot_q;
    f.slot_u = slot_u;
    f.direct = direct;
    f.slot_age = slot_age;
    return f;
}

// grid 33 ms, a submission every gap ms (the running mean is exact)
static RouteState measured(double gap)
{
    RouteState r;
    for (int i = 0; i < 40; ++i) route_note_submit(r, i * gap);
    route_note_grid(r, 33.0);
    return r;
}

int main()
{
    {  // nothing in flight: grid mode, measurements or not
        RouteState r;
        expect(route_grid(r, flight(0, 0, 0)), "alone, no measurements: grid");
        RouteState m = measured(10.0);  // even when arrivals would outpace grid mode
        expect(route_grid(m, flight(0, 0, 0)), "alone, measured: grid");
    }
    {  // a slot-path update in flight: never grid (the search does not wait for versions)
        RouteState r = measured(100.0);
        expect(!route_grid(r, flight(0, 1, 0)), "slot update in flight: slots");
        expect(!route_grid(r, flight(3, 1, 2)), "slot update among others: slots");
        expect(!route_grid(r, flight(0, 1, 0, 5000.0)), "slot update, old slot query: slots");
    }
    {  // no grid time or no arrival rate yet: behind at most one grid query, nothing on the slot path
        RouteState r;
        expect(route_grid(r, flight(0, 0, 1)), "one grid query in flight, no measurements: grid");
        expect(!route_grid(r, flight(0, 0, 2)), "two grid queries in flight, no measurements: slots");
        expect(!route_grid(r, flight(1, 0, 0, 500.0)), "a slot query running, no measurements: slots");
        expect(!route_grid(r, flight(1, 0, 1, 500.0)), "slot query and grid query, no measurements: slots");
        route_note_grid(r, 33.0);
        expect(route_grid(r, flight(0, 0, 1)), "no gaps yet: behind one grid query");
        expect(!route_grid(r, flight(0, 0, 2)), "no gaps yet: not two");
        RouteState g;
        for (int i = 0; i < 5; ++i) route_note_submit(g, i * 100.0);
        expect(route_grid(g, flight(0, 0, 1)), "no grid time yet: behind one grid query");
        expect(!route_grid(g, flight(0, 0, 2)), "no grid time yet: not two");
        RouteState few;  // a grid time but fewer than kRouteMinGaps gaps: the rate does not count yet
        for (int i = 0; i < 8; ++i) route_note_submit(few, i * 10.0);  // 7 gaps, rho 3.3
        route_note_grid(few, 33.0);
        expect(route_grid(few, flight(0, 0, 1)), "7 gaps: still behind one grid query");
        expect(!route_grid(few, flight(0, 0, 2)), "7 gaps: not two");
        route_note_submit(few, 80.0);  // the 8th gap: rho 3.3 counts
        expect(!route_grid(few, flight(0, 0, 1)), "8 gaps at rho 3.3: slots");
        expect(!few.grid_regime, "8 gaps at rho 3.3: no regime");
    }
    {  // in the grid regime: grid whenever grid queries are queued (a slot query would wait for them)
        RouteState r = measured(100.0);  // rho = 0.33
        expect(route_grid(r, flight(0, 0, 1)), "one grid query queued: grid");
        expect(r.grid_regime, "rho 0.33 enters the grid regime");
        expect(route_grid(r, flight(0, 0, 40)), "forty queued: a slot query would wait for them too, grid");
        expect(route_grid(r, flight(3, 0, 2, 1.0)), "grid queries queued, a young slot query: grid");
    }
    {  // in the grid regime, slot queries running and no grid query queued: grid once the youngest is older than G
        RouteState r = measured(100.0);
        expect(route_grid(r, flight(1, 0, 0, 50.0)), "youngest slot query 50 ms old: 50 > 33, grid");
        expect(!route_grid(r, flight(1, 0, 0, 20.0)), "youngest slot query 20 ms old: 20 <= 33, slots");
        expect(!route_grid(r, flight(4, 0, 0, 33.0)), "exactly G old: slots");
        expect(route_grid(r, flight(4, 0, 0, 900.0)), "long-running slot queries: grid");
    }
    {  // out of the regime: slots unless nothing is in flight
        RouteState h = measured(10.0);  // rho = 3.3
        expect(!route_grid(h, flight(0, 0, 1)), "arrivals outpace grid mode: slots");
        expect(!route_grid(h, flight(1, 0, 0, 500.0)), "arrivals outpace grid mode, old slot query: slots");
        expect(!h.grid_regime, "regime off at rho 3.3");
    }
    {  // hysteresis: enter below 0.7, leave above 1.0
        RouteState r = measured(40.0);  // rho = 33 / 40 = 0.825
        expect(!route_grid(r, flight(0, 0, 1)), "rho 0.825 does not enter");
        expect(!r.grid_regime, "regime stays off at 0.825");
        r.grid_regime = true;
        expect(route_grid(r, flight(0, 0, 1)), "rho 0.825 stays in once in");
        RouteState q = measured(30.0);  // rho = 1.1
        q.grid_regime = true;
        expect(!route_grid(q, flight(0, 0, 1)), "rho 1.1 leaves");
        expect(!q.grid_regime, "regime off after rho 1.1");
    }
    {  // updates go to their own kernel unless a slot update is pending or an open launch runs slot queries
        RouteInFlight open_q = flight(1, 0, 0), stopped_q = flight(3, 0, 2), pend = flight(0, 1, 0);
        open_q.slots_open = true;
        pend.slots_open = false;
        expect(route_direct_update(flight(0, 0, 0)), "idle: direct update");
        RouteInFlight idle_open = flight(0, 0, 0);
        idle_open.slots_open = true;
        expect(route_direct_update(idle_open), "an open launch with no slot query: direct update");
        expect(route_direct_update(flight(0, 0, 5)), "grid queries queued: direct update");
        expect(!route_direct_update(open_q), "an open launch runs a slot query: slot update");
        expect(route_direct_update(stopped_q), "slot queries in a stopped launch: direct update");
        expect(!route_direct_update(pend), "slot update pending: slot update (versions in order)");
        RouteInFlight pend_open = flight(2, 1, 0);
        pend_open.slots_open = true;
        expect(!route_direct_update(pend_open), "slot update pending in an open launch: slot update");
    }
    {  // moving averages
        RouteState r;
        route_note_grid(r, 40.0);
        expect(r.grid_ms == 40.0, "first grid query sets its average");
        route_note_grid(r, 20.0);
        expect(std::fabs(r.grid_ms - 35.0) < 1e-9, "then 0.75 old + 0.25 new");
        RouteState g;
        route_note_submit(g, 5.0);
        expect(g.gap_ms == 0.0 && g.gaps == 0, "one submission: no gap yet");
        route_note_submit(g, 15.0);
        route_note_submit(g, 45.0);
        expect(std::fabs(g.gap_ms - 20.0) < 1e-9 && g.gaps == 2, "running mean of the first gaps");
        RouteState h;
        for (int i = 0; i <= 50; ++i) route_note_submit(h, i * 10.0);  // 50 gaps of 10
        route_note_submit(h, 500.0 + 110.0);                            // then one of 110
        expect(std::fabs(h.gap_ms - (10.0 + 0.02 * 100.0)) < 1e-9, "weight 1/50 after 50 gaps");
    }
    std::printf("failures %d\n", failures);
    return failures != 0;
}
