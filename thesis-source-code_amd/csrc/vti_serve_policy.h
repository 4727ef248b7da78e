// Admission rule of the request server (csrc/vti_serve.cpp), as a function of
// the server's state only, so that it can be checked on the CPU.
#pragma once

namespace vti {

struct ServeState {
    long long budget = 0;      // device bytes the server may use
    long long resident = 0;    // held by resident models
    long long inflight = 0;    // estimates of the runs in flight
    int running = 0;           // runs in flight
    bool excl_running = false; // one of them is a 3D request running alone
    bool exclusive3d = true;   // 3D requests run alone
};

// Whether the oldest queued request (a 3D modeling request when is3d; its
// device-memory estimate `bytes`) may start now.  Requests start in queue
// order: when this is false the workers wait for a run to finish.
//  - Nothing starts beside a 3D request running alone.
//  - A 3D request that runs alone waits until nothing runs.
//  - Otherwise the request's estimate must fit beside those in flight within
//    the budget left by the resident models, or nothing may be in flight.
inline bool may_start(const ServeState &s, bool is3d, long long bytes)
{
    if (s.excl_running) return false;
    if (s.exclusive3d && is3d && s.running > 0) return false;
    return s.inflight == 0 || s.inflight + bytes <= s.budget - s.resident;
}

}  // namespace vti
