// serve.cpp -- the continuous-batching request loop (cfs_serve of
// include/cfs.h), written on the public batch entry points only.
//
// Rounds: admit what has arrived and fits, advance every resident problem
// steps_per_poll RK steps, wait, retire what finished.  Admission is in
// arrival order, first fit, with one reservation: when the oldest waiting
// request does not fit, a later one is admitted only if it leaves at least the
// oldest one's footprint free, so a large request is not starved by a stream
// of small ones.
#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "../../include/cfs.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

int64_t footprint_of(const cfs_request& r) { return cfs_batch_footprint(r.cfg, r.ib ? r.ib->n_ig : 0); }

}  // namespace

extern "C" int32_t cfs_serve(cfs_batch* b, cfs_request* reqs, int32_t n, int32_t steps_per_poll) {
    if (!b || n < 0 || (n > 0 && !reqs) || steps_per_poll < 1) return CFS_E_ARG;
    int64_t arena = 0;
    int32_t nslots = 0;
    int rc = cfs_batch_capacity(b, &arena, &nslots);
    if (rc) return rc;
    std::vector<int64_t> fp((size_t)n);
    for (int k = 0; k < n; ++k) {
        cfs_request& r = reqs[k];
        if (!r.cfg || !r.U || !(r.t_end > 0.0) || !(r.arrival_s >= 0.0)) return CFS_E_ARG;
        fp[k] = footprint_of(r);
        if (fp[k] < 0) return CFS_E_ARG;
        if (fp[k] > arena) return CFS_E_FULL;
        r.admitted_s = r.finished_s = r.t = 0.0;
        r.steps = 0;
        r.state = CFS_SLOT_FREE;
    }
    std::vector<int> order((size_t)n);
    for (int k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return reqs[a].arrival_s < reqs[c].arrival_s; });
    std::vector<int> slot_req((size_t)nslots, -1);
    std::vector<cfs_batch_info> info((size_t)nslots);
    std::vector<int> waiting;
    size_t next = 0;
    int resident = 0, finished = 0;
    const Clock::time_point t0 = Clock::now();
    while (finished < n) {
        const double now = seconds_since(t0);
        while (next < order.size() && reqs[order[next]].arrival_s <= now) waiting.push_back(order[next++]);
        // admission
        int64_t reserve = 0;
        for (size_t w = 0; w < waiting.size();) {
            const int k = waiting[w];
            int64_t free_bytes = 0;
            if ((rc = cfs_batch_free_bytes(b, &free_bytes))) return rc;
            int32_t slot = -1;
            rc = CFS_E_FULL;
            if (fp[k] + reserve <= free_bytes)
                rc = cfs_batch_admit(b, reqs[k].cfg, reqs[k].U, reqs[k].ib, reqs[k].cfl, reqs[k].t_end, &slot);
            if (rc == CFS_OK) {
                reqs[k].admitted_s = seconds_since(t0);
                slot_req[slot] = k;
                ++resident;
                waiting.erase(waiting.begin() + (long)w);
            } else if (rc == CFS_E_FULL) {
                if (w == 0) reserve = fp[k];   // the oldest waiting request keeps its room
                ++w;
            } else {
                return rc;
            }
        }
        if (resident == 0) {
            if (next < order.size()) {
                const double wait = reqs[order[next]].arrival_s - seconds_since(t0);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            }
            continue;
        }
        // one round of continuous batching
        if ((rc = cfs_batch_step(b, steps_per_poll))) return rc;
        int32_t m = 0;
        if ((rc = cfs_batch_poll(b, info.data(), nslots, &m))) return rc;
        const double done_at = seconds_since(t0);
        for (int q = 0; q < m && q < nslots; ++q) {
            const cfs_batch_info& in = info[q];
            if (in.state != CFS_SLOT_DONE && in.state != CFS_SLOT_FAILED) continue;
            const int k = slot_req[in.slot];
            if (k < 0) continue;
            cfs_request& r = reqs[k];
            r.finished_s = done_at;
            r.t = in.t;
            r.steps = in.steps;
            r.state = in.state;
            if ((rc = cfs_batch_retire(b, in.slot, r.U_out))) return rc;
            slot_req[in.slot] = -1;
            --resident;
            ++finished;
        }
    }
    return CFS_OK;
}
