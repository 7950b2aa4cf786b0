// serve.cpp -- the continuous-batching request loop (cfs_serve of
// include/cfs.h), written on the public batch and single-solve entry points.
//
// Rounds: admit what has arrived and fits, advance every resident problem
// steps_per_poll RK steps, wait, retire what finished.  Admission is in
// arrival order, first fit, with one reservation: when the oldest waiting
// request does not fit, a later one is admitted only if it leaves at least the
// oldest one's footprint free, so a large request is not starved by a stream
// of small ones.
//
// Solo lane: a request that fills the GPU on its own (at least as many tiles
// as the GPU holds stage-kernel workgroups at once) is not batched.  Two
// worker threads solve such requests in arrival order, each through the
// single-solve entry points on a stream of its own, while the rounds go on
// (two, so that the last partial wave of workgroups of one solve's stage
// overlaps the other's work).  Batching gains nothing for such a request, the single-solve kernel
// starts each tile with its arguments already in registers (the batched one
// first reads them from device memory), and the request no longer waits for
// every other resident problem at each of its steps, which matters because
// the largest requests take the most steps.  The arithmetic is the same
// kernel's either way, so its result does not depend on the lane.  The solo
// lane is admitted by footprint too: a worker takes the next request only
// when its footprint fits a budget of nine tenths of the device memory free
// at the start beside what the other worker holds, and keeps its solver
// handle (no allocation on the request path) while the next request has the
// same configuration.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cfs.h"
#include "serve_lane.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kSoloWorkers = 2;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

int64_t footprint_of(const cfs_request& r) { return cfs_batch_footprint(r.cfg, r.ib ? r.ib->n_ig : 0); }

// A worker's solver handle, kept from one request to the next when the next
// one has the same configuration (a new handle otherwise, or when the kept
// one carries an immersed boundary the next request does not have).  `bytes`
// is the device memory it counts against the solo lane's budget.
struct Handle {
    cfs_solver* h = nullptr;
    cfs_config cfg{};
    bool ib = false;
    int64_t bytes = 0;
    void drop() {
        cfs_destroy(h);
        h = nullptr;
        ib = false;
        bytes = 0;
    }
};

bool same_config(const cfs_config& a, const cfs_config& b) {
    if (a.nx != b.nx || a.ny != b.ny || a.lx != b.lx || a.ly != b.ly || a.x0 != b.x0 || a.y0 != b.y0 ||
        a.gamma != b.gamma || a.rk != b.rk || a.precision != b.precision || a.viscous != b.viscous ||
        a.positivity != b.positivity || a.hll_bounds != b.hll_bounds || a.mu != b.mu || a.prandtl != b.prandtl ||
        a.ma_ref != b.ma_ref)
        return false;
    for (int k = 0; k < 4; ++k) {
        if (a.bc[k] != b.bc[k]) return false;
        for (int q = 0; q < 4; ++q)
            if (a.bc_state[k][q] != b.bc_state[k][q]) return false;
    }
    return true;
}

bool reusable(const Handle& hd, const cfs_request& r) { return hd.h && same_config(hd.cfg, *r.cfg) && (!hd.ib || r.ib); }

// One request solved alone to its end time on the worker's handle; CFS_OK
// with r.state DONE, or FAILED on a non-finite CFL bound (as a batch slot
// would fail), else the error of the entry point that failed.
int32_t solve_alone(Handle& hd, cfs_request& r, int64_t footprint) {
    int32_t rc = CFS_OK;
    if (!reusable(hd, r)) {
        hd.drop();
        rc = cfs_create(r.cfg, nullptr, &hd.h);
        if (rc) return rc;
        hd.cfg = *r.cfg;
        hd.bytes = footprint;
    }
    if (r.ib) {
        rc = cfs_set_immersed(hd.h, r.ib->interp, r.ib->cell_type, r.ib->n_ig, r.ib->ig_cells, r.ib->donors,
                              r.ib->image_xy, r.ib->coef, r.ib->transp);
        hd.ib = true;
    }
    if (!rc) rc = cfs_set_state(hd.h, r.U);
    double t = 0.0;
    int32_t steps = 0;
    r.state = CFS_SLOT_DONE;
    while (!rc && t < r.t_end) {
        int32_t done = 0;
        const int32_t st = cfs_advance(hd.h, 1 << 30, 0.0, r.cfl, r.t_end, &t, &done);
        steps += done;
        if (st == CFS_E_NONFINITE) {
            r.state = CFS_SLOT_FAILED;
            break;
        }
        rc = st;
        if (done == 0) break;
    }
    if (!rc && r.U_out) rc = cfs_get_state(hd.h, r.U_out);
    r.t = t;
    r.steps = steps;
    if (rc) hd.drop();
    return rc;
}

// The solo lane: requests in arrival order, each taken by a worker once its
// footprint fits the lane's device-memory budget beside the handles the other
// workers hold.  A worker keeps its handle after a request only while the
// next queued request could reuse it; otherwise it frees it, so that waiting
// requests are never blocked by idle handles.
struct SoloLane {
    std::mutex m;
    std::condition_variable cv;
    std::deque<int> queue;
    bool closing = false;
    int32_t error = CFS_OK;      // first failure of an entry point (under m)
    std::atomic<int> finished{0};
    int64_t budget = 0;          // device bytes the lane's handles may hold at once
    const std::vector<int64_t>* fp = nullptr;
    std::vector<Handle> handles;

    int64_t held_by_others(int w) const {
        int64_t s = 0;
        for (int k = 0; k < (int)handles.size(); ++k)
            if (k != w) s += handles[k].bytes;
        return s;
    }

    void run(int w, const cfs_batch* b, cfs_request* reqs, Clock::time_point t0) {
        const int32_t dev_rc = batch_make_device_current(b);
        Handle& hd = handles[w];
        for (;;) {
            int k;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] {
                    return closing || (!queue.empty() && held_by_others(w) + (*fp)[queue.front()] <= budget);
                });
                if (closing && queue.empty()) break;
                k = queue.front();
                queue.pop_front();
                // the handle is replaced unless the request can reuse it
                if (!reusable(hd, reqs[k])) hd.drop();
            }
            cv.notify_all();
            cfs_request& r = reqs[k];
            r.admitted_s = seconds_since(t0);
            const int32_t rc = dev_rc ? dev_rc : solve_alone(hd, r, (*fp)[k]);
            r.finished_s = seconds_since(t0);
            {
                std::lock_guard<std::mutex> lk(m);
                if (rc && !error) error = rc;
                if (queue.empty() || !reusable(hd, reqs[queue.front()])) hd.drop();
            }
            finished.fetch_add(1);
            cv.notify_all();
        }
        std::lock_guard<std::mutex> lk(m);
        hd.drop();
    }
    void push(int k) {
        {
            std::lock_guard<std::mutex> lk(m);
            queue.push_back(k);
        }
        cv.notify_all();
    }
    void close() {   // the workers end after the solves in hand (nothing is queued on success)
        {
            std::lock_guard<std::mutex> lk(m);
            closing = true;
            queue.clear();
        }
        cv.notify_all();
    }
    int32_t failure() {
        std::lock_guard<std::mutex> lk(m);
        return error;
    }
};

}  // namespace

extern "C" int32_t cfs_serve(cfs_batch* b, cfs_request* reqs, int32_t n, int32_t steps_per_poll) {
    if (!b || n < 0 || (n > 0 && !reqs) || steps_per_poll < 1) return CFS_E_ARG;
    int64_t arena = 0;
    int32_t nslots = 0;
    int rc = cfs_batch_capacity(b, &arena, &nslots);
    if (rc) return rc;
    const int64_t resident_wg = batch_resident_workgroups(b);
    // the solo lane's budget: nine tenths of the device memory free now (the
    // batch's arena is already allocated)
    int64_t free_dev = 0;
    if ((rc = batch_device_free_bytes(b, &free_dev))) return rc;
    const int64_t solo_budget = free_dev / 10 * 9;
    std::vector<int64_t> fp((size_t)n);
    std::vector<char> solo((size_t)n, 0);
    for (int k = 0; k < n; ++k) {
        cfs_request& r = reqs[k];
        if (!r.cfg || !r.U || !(r.t_end > 0.0) || !(r.arrival_s >= 0.0)) return CFS_E_ARG;
        if ((rc = batch_check_problem(b, r.cfg, r.ib, r.cfl, r.t_end))) return rc;
        fp[k] = footprint_of(r);
        if (fp[k] < 0) return CFS_E_ARG;
        solo[k] = batch_problem_tiles(b, r.cfg) >= resident_wg;
        if (!solo[k] && fp[k] > arena) return CFS_E_FULL;
        if (solo[k] && fp[k] > solo_budget) return CFS_E_FULL;
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
    SoloLane lane;
    lane.budget = solo_budget;
    lane.fp = &fp;
    lane.handles.resize(kSoloWorkers);
    std::vector<std::thread> workers;
    auto stop = [&](int32_t code) {
        lane.close();
        for (auto& w : workers) w.join();
        return code;
    };
    try {
        for (int w = 0; w < kSoloWorkers; ++w) workers.emplace_back([&, w] { lane.run(w, b, reqs, t0); });
    } catch (...) {
        return stop(CFS_E_OOM);
    }
    while (finished + lane.finished.load() < n) {
        const double now = seconds_since(t0);
        while (next < order.size() && reqs[order[next]].arrival_s <= now) {
            const int k = order[next++];
            if (solo[k]) {
                lane.push(k);
            } else {
                waiting.push_back(k);
            }
        }
        if ((rc = lane.failure())) return stop(rc);
        // admission
        int64_t reserve = 0;
        for (size_t w = 0; w < waiting.size();) {
            const int k = waiting[w];
            int64_t free_bytes = 0;
            if ((rc = cfs_batch_free_bytes(b, &free_bytes))) return stop(rc);
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
                return stop(rc);
            }
        }
        if (resident == 0) {
            // nothing to batch: wait for the next arrival or the solo lane
            double wait = 1e-3;
            if (next < order.size()) wait = std::min(wait, reqs[order[next]].arrival_s - seconds_since(t0));
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            continue;
        }
        // one round of continuous batching
        if ((rc = cfs_batch_step(b, steps_per_poll))) return stop(rc);
        int32_t m = 0;
        if ((rc = cfs_batch_poll(b, info.data(), nslots, &m))) return stop(rc);
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
            if ((rc = cfs_batch_retire(b, in.slot, r.U_out))) return stop(rc);
            slot_req[in.slot] = -1;
            --resident;
            ++finished;
        }
    }
    return stop(lane.failure());
}
