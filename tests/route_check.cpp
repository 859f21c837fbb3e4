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

// grid 33 ms, slot 700 ms, a submission every gap ms (the running mean is exact)
static RouteState measured(double gap)
{
    RouteState r;
    for (int i = 0; i < 40; ++i) route_note_submit(r, i * gap);
    route_note_query(r, true, 33.0);
    route_note_query(r, false, 700.0);
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
    }
    {  // no grid time or no arrival rate yet: behind at most one grid query, nothing on the slot path
        RouteState r;
        expect(route_grid(r, flight(0, 0, 1)), "one grid query in flight, no measurements: grid");
        expect(!route_grid(r, flight(0, 0, 2)), "two grid queries in flight, no measurements: slots");
        expect(!route_grid(r, flight(1, 0, 0, 500.0)), "a slot query running, no measurements: slots");
        expect(!route_grid(r, flight(1, 0, 1, 500.0)), "slot query and grid query, no measurements: slots");
        route_note_query(r, false, 700.0);
        expect(route_grid(r, flight(0, 0, 1)), "slot average only: still behind one grid query");
        route_note_query(r, true, 33.0);
        expect(route_grid(r, flight(0, 0, 1)), "no gaps yet: behind one grid query");
        expect(!route_grid(r, flight(0, 0, 2)), "no gaps yet: not two");
    }
    {  // grid time and arrival rate measured, no slot query finished yet
        RouteState r;
        for (int i = 0; i < 40; ++i) route_note_submit(r, i * 100.0);
        route_note_query(r, true, 33.0);  // rho = 0.33
        expect(route_grid(r, flight(0, 0, 1)), "no slot average, one grid query ahead: grid");
        expect(!route_grid(r, flight(0, 0, 2)), "no slot average, two grid queries ahead: slots");
        expect(route_grid(r, flight(1, 0, 0, 50.0)), "youngest slot query 50 ms old: 33 < 50, grid");
        expect(!route_grid(r, flight(1, 0, 0, 20.0)), "youngest slot query 20 ms old: 33 >= 20, slots");
        expect(route_grid(r, flight(3, 0, 2, 200.0)), "two grid queries ahead, 200 ms old: 99 < 200, grid");
        RouteState h;
        for (int i = 0; i < 40; ++i) route_note_submit(h, i * 10.0);
        route_note_query(h, true, 33.0);  // rho = 3.3
        expect(!route_grid(h, flight(0, 0, 1)), "arrivals outpace grid mode: slots");
        expect(!route_grid(h, flight(1, 0, 0, 500.0)), "arrivals outpace grid mode, old slot query: slots");
    }
    {  // low load: queue behind the grid queries while they finish before a slot query would
        RouteState r = measured(100.0);  // rho = 0.33
        expect(route_grid(r, flight(0, 0, 2)), "two grid queries queued: grid");
        expect(r.grid_regime, "rho 0.33 enters the grid regime");
        expect(route_grid(r, flight(0, 0, 19)), "20 x 33 = 660 < 700: grid");
        expect(!route_grid(r, flight(0, 0, 21)), "22 x 33 = 726 >= 700: slots");
        expect(route_grid(r, flight(1, 0, 0, 100.0)), "a slot query 100 ms old, ending in 600 ms: 633 < 700, grid");
        expect(!route_grid(r, flight(1, 0, 0, 20.0)), "a slot query 20 ms old, ending in 680 ms: 713 >= 700, slots");
        expect(route_grid(r, flight(1, 0, 2, 100.0)), "two grid queries and a 100 ms old slot query: 99 < 100, grid");
        expect(!route_grid(r, flight(1, 0, 3, 100.0)), "three grid queries and a 100 ms old slot query: 132 >= 100, slots");
        expect(route_grid(r, flight(1, 0, 20, 2000.0)), "a slot query past the average: 21 x 33 = 693 < 700, grid");
        expect(!route_grid(r, flight(1, 0, 21, 2000.0)), "a slot query past the average: 726 >= 700, slots");
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
    {  // updates go to their own kernel only with nothing on the slot path
        expect(route_direct_update(flight(0, 0, 0)), "idle: direct update");
        expect(route_direct_update(flight(0, 0, 5)), "grid queries queued: direct update");
        expect(!route_direct_update(flight(1, 0, 0)), "slot query running: slot update");
        expect(!route_direct_update(flight(0, 1, 0)), "slot update pending: slot update");
    }
    {  // moving averages
        RouteState r;
        route_note_query(r, false, 800.0);
        expect(r.slot_ms == 800.0 && r.grid_ms == 0.0, "first slot query sets its average");
        route_note_query(r, false, 400.0);
        expect(std::fabs(r.slot_ms - 700.0) < 1e-9, "then 0.75 old + 0.25 new");
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
