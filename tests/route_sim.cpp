// The server's grid-mode routing policy (ddp_amd/csrc/serve_route.h) driven
// by Poisson arrivals against a queueing model of the device it routes for:
//  * grid-mode queries and directly applied updates run one after another on
//    the server's stream;
//  * routing one there stops the open serve launch: its running slot
//    requests finish, and it claims nothing submitted after that point;
//  * slot requests submitted while no launch is open wait for a relaunch,
//    which the host issues once the stopped launch has ended and which the
//    stream runs after every kernel queued before it;
//  * an open launch claims a slot request at once (at these loads a slot is
//    always free).
// The host sees a request as in flight until its finish time, and learns a
// grid query's device time when it finishes.  Checked: at low load every
// query after the first few goes to grid mode and none waits a slot time; at
// high load grid mode is left, so the slots' throughput is kept.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "serve_route.h"

namespace {

enum Kind { GRID, SLOT, DIRECT_UPDATE, SLOT_UPDATE };

struct Req {
    double submit = 0, finish = -1;  // finish < 0: waiting for a relaunch
    double dev = 0;                  // device time (grid queries)
    Kind kind = SLOT;
};

struct Result {
    double grid_share = 0;  // of the queries after the warm-up
    double p99_ms = 0;      // their latency
};

constexpr double kUpdateShare = 0.05;
constexpr int kWarmup = 20;

Result simulate(double gap_ms, int n, unsigned seed)
{
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(1.0 / gap_ms);
    std::uniform_real_distribution<double> grid_t(25.0, 45.0), slot_t(450.0, 1000.0), u01(0.0, 1.0);
    RouteState route;
    std::vector<Req> req(n);
    std::vector<bool> polled(n, false);
    std::vector<int> pending;  // slot requests waiting for a relaunch
    double stream_free = 0;    // end of the kernels queued on the stream (an open launch excluded)
    bool open = false, stopping = false;
    double launch_start = 0, launch_end = 0, launch_last = 0;  // launch_last: last finish of its claims

    auto claim = [&](int i) {
        const double dur = req[i].kind == SLOT_UPDATE ? 1.0 : slot_t(rng);
        req[i].finish = std::max(req[i].submit, launch_start) + dur;
        launch_last = std::max(launch_last, req[i].finish);
    };
    auto stop_open = [&](double now) {
        if (!open) return;
        open = false;
        stopping = true;
        launch_end = std::max(std::max(now, launch_start), launch_last);
        stream_free = std::max(stream_free, launch_end);
    };
    auto relaunch = [&](double at) {
        open = true;
        stopping = false;
        launch_start = std::max(at, stream_free);
        launch_last = launch_start;
        for (int i : pending) claim(i);
        pending.clear();
    };

    double t = 0;
    for (int i = 0; i < n; ++i) {
        t += gap(rng);
        // what the host has seen by t: finished requests (grid times noted in
        // finish order) and the end of a stopped launch, which it relaunches
        std::vector<int> done;
        for (int k = 0; k < i; ++k)
            if (!polled[k] && req[k].finish >= 0 && req[k].finish <= t) done.push_back(k);
        std::sort(done.begin(), done.end(), [&](int a, int b) { return req[a].finish < req[b].finish; });
        for (int k : done) {
            polled[k] = true;
            if (req[k].kind == GRID) route_note_grid(route, req[k].dev);
        }
        if (stopping && launch_end <= t) {
            stopping = false;
            if (!pending.empty()) relaunch(launch_end);
        }
        RouteInFlight f;
        double youngest = 0;
        for (int k = 0; k < i; ++k) {
            if (polled[k]) continue;
            if (req[k].kind == GRID || req[k].kind == DIRECT_UPDATE) {
                ++f.direct;
            } else if (req[k].kind == SLOT_UPDATE) {
                ++f.slot_u;
            } else {
                ++f.slot_q;
                youngest = std::max(youngest, req[k].submit);
            }
        }
        f.slot_age = f.slot_q ? t - youngest : 0;
        f.slots_open = open;
        route_note_submit(route, t);
        req[i].submit = t;
        const bool update = u01(rng) < kUpdateShare;
        const bool direct = update ? route_direct_update(f) : route_grid(route, f);
        if (direct) {
            req[i].kind = update ? DIRECT_UPDATE : GRID;
            stop_open(t);
            const double start = std::max(t, stream_free);
            req[i].dev = update ? 0.2 : grid_t(rng);
            req[i].finish = start + req[i].dev;
            stream_free = req[i].finish;
        } else {
            req[i].kind = update ? SLOT_UPDATE : SLOT;
            if (open) {
                claim(i);
            } else {
                pending.push_back(i);
                if (!stopping) relaunch(t);
            }
        }
    }
    // the last relaunch, for requests still pending
    if (stopping && !pending.empty()) relaunch(launch_end);

    int queries = 0, grid = 0;
    std::vector<double> lat;
    for (int i = kWarmup; i < n; ++i) {
        if (req[i].kind == DIRECT_UPDATE || req[i].kind == SLOT_UPDATE) continue;
        ++queries;
        grid += req[i].kind == GRID;
        lat.push_back(req[i].finish - req[i].submit);
    }
    std::sort(lat.begin(), lat.end());
    Result r;
    r.grid_share = queries ? (double)grid / queries : 0;
    r.p99_ms = lat.empty() ? 0 : lat[(size_t)(0.99 * (double)(lat.size() - 1))];
    return r;
}

int failures = 0;

void expect(bool ok, const char *what, double gap_ms, const Result &r)
{
    std::printf("%-58s gap %6.1f ms: grid share %.3f, p99 %7.1f ms%s\n", what, gap_ms, r.grid_share, r.p99_ms,
                ok ? "" : "  FAIL");
    if (!ok) ++failures;
}

}  // namespace

int main()
{
    // grid queries take 25-45 ms (mean 35), slot queries 450-1000 ms
    for (double gap : {700.0, 233.0, 117.0, 78.0}) {  // rho = 0.05, 0.15, 0.3, 0.45
        for (unsigned seed = 1; seed <= 10; ++seed) {
            const Result r = simulate(gap, 1000, seed);
            expect(r.grid_share >= 0.99 && r.p99_ms < 450.0, "low load: grid mode, no query waits a slot time", gap, r);
        }
    }
    for (unsigned seed = 1; seed <= 4; ++seed) {  // rho = 3: grid mode alone would fall behind
        const Result r = simulate(12.0, 3000, seed);
        expect(r.grid_share <= 0.05, "high load: slots", 12.0, r);
    }
    std::printf("failures %d\n", failures);
    return failures != 0;
}
