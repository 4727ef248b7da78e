// CPU test of the request server's admission rule (csrc/vti_serve_policy.h),
// driven by a discrete-event simulation of the server's workers (test
// infrastructure; tests/test_serve_policy.py builds and runs it).
//
// Requests (a random mix of 2D gradients and 3D modeling with random memory
// estimates and service times) arrive at random times; `slots` workers take
// the oldest queued request whenever may_start() allows, as the server's
// workers do under its lock, and hold it for its service time.  Checked:
//   - requests start in arrival order;
//   - with exclusive3d, no run overlaps a 3D run;
//   - the estimates in flight stay within the budget left by the resident
//     models, except for a request that runs with nothing else in flight;
//   - every request completes (no wait that nothing ends).
// `./serve_policy shared` runs with exclusive3d off and reports how many 3D
// runs overlapped others (the test expects some, so the check can fail).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "vti_serve_policy.h"

namespace {

struct Req {
    double arrive, service;
    bool is3d;
    long long bytes;
    double start = -1, done = -1;
};

int run(bool exclusive, unsigned seed, int slots, int *overlaps3d)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const long long budget = 100, resident = 20;
    std::vector<Req> reqs;
    double t = 0;
    for (int k = 0; k < 400; ++k) {
        t += -std::log(1.0 - u(rng)) * 0.05;
        Req r;
        r.arrive = t;
        r.is3d = u(rng) < 0.3;
        r.service = r.is3d ? 0.1 + 0.3 * u(rng) : 0.01 + 0.05 * u(rng);
        r.bytes = 5 + (long long)(70 * u(rng));  // some need the whole budget to themselves
        reqs.push_back(r);
    }
    vti::ServeState st;
    st.budget = budget;
    st.resident = resident;
    st.exclusive3d = exclusive;
    std::vector<int> running;  // request indices in flight
    size_t next_arrival = 0, next_start = 0;
    double now = 0;
    int violations = 0;
    for (int guard = 0; guard < 100000; ++guard) {
        // start every request the rule admits, oldest first
        while (next_start < next_arrival && (int)running.size() < slots) {
            Req &r = reqs[next_start];
            if (!vti::may_start(st, r.is3d, r.bytes)) break;
            if (st.inflight + r.bytes > st.budget - st.resident && st.inflight != 0) {
                std::fprintf(stderr, "budget exceeded at %.3f\n", now);
                ++violations;
            }
            r.start = now;
            st.inflight += r.bytes;
            st.running += 1;
            if (exclusive && r.is3d) st.excl_running = true;
            running.push_back((int)next_start);
            ++next_start;
        }
        if (next_start == reqs.size() && running.empty()) break;
        // next event: an arrival or the earliest completion
        double t_done = 1e300;
        int who = -1;
        for (int i : running)
            if (reqs[i].start + reqs[i].service < t_done) {
                t_done = reqs[i].start + reqs[i].service;
                who = i;
            }
        const double t_arr = next_arrival < reqs.size() ? reqs[next_arrival].arrive : 1e300;
        if (t_arr <= t_done) {
            now = t_arr;
            ++next_arrival;
        } else if (who >= 0) {
            now = t_done;
            Req &r = reqs[who];
            r.done = now;
            st.inflight -= r.bytes;
            st.running -= 1;
            if (exclusive && r.is3d) st.excl_running = false;
            running.erase(std::find(running.begin(), running.end(), who));
        } else {
            std::fprintf(stderr, "stuck: queued requests and nothing running at %.3f\n", now);
            return violations + 1;
        }
    }
    int unfinished = 0;
    for (const Req &r : reqs) unfinished += r.done < 0;
    if (unfinished) {
        std::fprintf(stderr, "%d requests never finished\n", unfinished);
        violations += unfinished;
    }
    for (size_t k = 1; k < reqs.size(); ++k)
        if (reqs[k].start < reqs[k - 1].start) {
            std::fprintf(stderr, "request %zu started before request %zu\n", k, k - 1);
            ++violations;
        }
    *overlaps3d = 0;
    for (size_t a = 0; a < reqs.size(); ++a) {
        if (!reqs[a].is3d) continue;
        for (size_t b = 0; b < reqs.size(); ++b)
            if (b != a && reqs[b].start < reqs[a].done && reqs[a].start < reqs[b].done) {
                ++*overlaps3d;
                break;
            }
    }
    if (exclusive && *overlaps3d) {
        std::fprintf(stderr, "%d 3D runs overlapped other runs\n", *overlaps3d);
        violations += *overlaps3d;
    }
    return violations;
}

}  // namespace

int main(int argc, char **argv)
{
    const bool shared = argc > 1 && std::strcmp(argv[1], "shared") == 0;
    int violations = 0, overlaps = 0;
    for (unsigned seed = 1; seed <= 20; ++seed)
        for (int slots : {1, 2, 4}) {
            int o = 0;
            violations += run(!shared, seed, slots, &o);
            overlaps += o;
        }
    std::printf("violations %d overlapping 3D runs %d\n", violations, overlaps);
    return violations ? 1 : 0;
}
