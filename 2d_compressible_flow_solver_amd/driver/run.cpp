// run.cpp -- Run::initialize / Run::solver of the driver.
#include "run.h"

#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>

#include "output.h"

namespace cfsdrv {

Run::~Run() {
    if (log_) std::fclose(log_);
    cfs_destroy(h_);
}

namespace {
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

int Run::log_record(int step) {
    if (!log_) return CFS_OK;
    double sums[4];
    // returns once the stages launched so far have ended; a non-finite state
    // is logged (finite: false) and then ends the run
    const int rc = cfs_checksums(h_, sums);
    if (rc && rc != CFS_E_NONFINITE) return rc;
    const double w = now_s();
    const double dw = w - wall_last_;
    const int ds = stages_ - stages_last_;
    const double cells = (double)cfg_.nx * cfg_.ny;
    // algorithmic bytes per cell-stage: the RK snapshots read and the state
    // written (RK3 mean 96 B, RK2 80 B), +1 B of cell types with a body
    const double bytes = (cfg_.rk == 3 ? 96.0 : 80.0) + (cell_type_.empty() ? 0.0 : 1.0);
    const double mcups = dw > 0.0 ? cells * ds / dw / 1e6 : 0.0;
    bool finite = true;
    for (double v : sums) finite = finite && std::isfinite(v);
    std::fprintf(log_,
                 "{\"step\": %d, \"stage\": %d, \"t\": %.17g, \"wall_s\": %.6f, \"interval_s\": %.6f, "
                 "\"interval_stages\": %d, \"mcell_updates_per_s\": %.3f, \"algorithmic_gb_per_s\": %.3f, "
                 "\"checksums\": [%.17g, %.17g, %.17g, %.17g], \"finite\": %s}\n",
                 step, stages_, t_, w - wall0_, dw, ds, mcups, mcups * bytes / 1e3, sums[0], sums[1], sums[2],
                 sums[3], finite ? "true" : "false");
    std::fflush(log_);
    wall_last_ = w;
    stages_last_ = stages_;
    return rc;
}

int Run::initialize() {
    // mesh (Run.cpp:14-20) and plugin choices (Run.cpp:41-52)
    cfg_ = cfs_config{};
    cfg_.nx = opt_.nx;
    cfg_.ny = opt_.ny;
    cfg_.lx = opt_.lx;
    cfg_.ly = opt_.ly;
    cfg_.gamma = opt_.gamma;
    cfg_.precision = opt_.precision;
    opt_.interpolation.configure(cfg_);
    opt_.riemann.configure(cfg_);
    opt_.timeadvance.configure(cfg_);
    opt_.viscous.configure(cfg_);
    if (!opt_.initialcondition) return CFS_E_ARG;
    Boundaries bcs;   // Initialcondition::setBounbdaryCondition (Run.cpp:66)
    opt_.initialcondition->setBoundaryCondition(cfg_.gamma, bcs);
    for (auto& b : bcs) b->configure(cfg_);
    int rc = cfs_create(&cfg_, nullptr, &h_);
    if (rc) return rc;
    // CartesianMesh::setImmersedBoundary (Run.cpp:53, CartesianMesh.cpp:357-374)
    if (opt_.immersed) {
        const Circle& c = opt_.circle;
        cell_type_.assign((size_t)cfg_.nx * cfg_.ny, 0);
        int32_t n_ig = 0;
        // first call with cap 0 counts the ghosts (CFS_E_ARG with the count set)
        rc = cfs_circle_setup(&cfg_, c.cx, c.cy, c.radius, c.points, c.layers, c.interp, 0, cell_type_.data(), &n_ig,
                              nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc != CFS_OK && !(rc == CFS_E_ARG && n_ig > 0)) return rc;
        std::vector<int32_t> ig(n_ig), donors(4 * (size_t)n_ig), transp(4 * (size_t)n_ig);
        std::vector<double> ip(2 * (size_t)n_ig), coef(10 * (size_t)n_ig);
        rc = cfs_circle_setup(&cfg_, c.cx, c.cy, c.radius, c.points, c.layers, c.interp, n_ig, cell_type_.data(),
                              &n_ig, ig.data(), donors.data(), ip.data(), coef.data(), transp.data());
        if (rc) return rc;
        rc = cfs_set_immersed(h_, c.interp, cell_type_.data(), n_ig, ig.data(), donors.data(), ip.data(), coef.data(),
                              transp.data());
        if (rc) return rc;
    }
    return CFS_OK;
}

int Run::snapshot(int index) {
    if (opt_.output_dir.empty()) return CFS_OK;
    char name[64];
    std::snprintf(name, sizeof(name), "/snapshot_%06d", index);
    const std::string dir = opt_.output_dir + name;
    ::mkdir(opt_.output_dir.c_str(), 0755);
    ::mkdir(dir.c_str(), 0755);
    return Output::write(h_, cfg_, cell_type_, dir, stages_, t_);
}

int Run::solver() {
    // initial condition on fluid cells (Run.cpp:99-105); other cells 0
    const size_t n = (size_t)cfg_.nx * cfg_.ny;
    std::vector<double> U(4 * n, 0.0);
    const double dx = cfg_.lx / cfg_.nx, dy = cfg_.ly / cfg_.ny;
    for (int i = 0; i < cfg_.nx; ++i)
        for (int j = 0; j < cfg_.ny; ++j) {
            const size_t c = (size_t)i * cfg_.ny + j;
            if (!cell_type_.empty() && cell_type_[c] != CFS_CELL_FLUID) continue;
            double s[4];
            opt_.initialcondition->fillIn(cfg_.gamma, i, j, (i + 0.5) * dx + cfg_.x0, (j + 0.5) * dy + cfg_.y0, s);
            for (int q = 0; q < 4; ++q) U[q * n + c] = s[q];
        }
    // uploads and fills the immersed ghosts (Run.cpp:106-112)
    int rc = cfs_set_state(h_, U.data());
    if (rc) return rc;
    if (!opt_.log_path.empty() && !(log_ = std::fopen(opt_.log_path.c_str(), "w"))) return CFS_E_ARG;
    if (opt_.dt > 0.0) {
        // the reference's loop: one stage per iteration, fixed dt (Run.cpp:153-204)
        wall0_ = wall_last_ = now_s();
        for (int i = 1; i < opt_.stages; ++i) {
            rc = cfs_stage(h_, i, opt_.dt);
            if (rc) return rc;
            ++stages_;
            if (opt_.output_every > 0 && i % opt_.output_every == 0) {
                t_ = opt_.dt * stages_ / cfg_.rk;
                if ((rc = log_record(stages_ / cfg_.rk)) || (rc = snapshot(i))) return rc;
            }
        }
        t_ = opt_.dt * stages_ / cfg_.rk;
        return stages_ != stages_last_ ? log_record(stages_ / cfg_.rk) : CFS_OK;
    }
    // CFL steps on the device to t_end, snapshots every output_every steps;
    // the stage kernel's tile shape is tuned on the initial state first (the
    // results do not depend on it)
    if ((rc = cfs_autotune(h_, 10, nullptr, nullptr))) return rc;
    const int chunk = opt_.output_every > 0 ? opt_.output_every : 1000000;
    int k = 0;
    wall0_ = wall_last_ = now_s();
    while (opt_.t_end > 0.0 && t_ < opt_.t_end) {
        int32_t done = 0;
        rc = cfs_advance(h_, chunk, 0.0, opt_.cfl, opt_.t_end, &t_, &done);
        stages_ += done * cfg_.rk;
        if (rc) return rc;
        ++k;
        if ((rc = log_record(stages_ / cfg_.rk))) return rc;
        if (opt_.output_every > 0 && (rc = snapshot(k))) return rc;
        if (done < chunk) break;
    }
    return CFS_OK;
}

}  // namespace cfsdrv
