// run.h -- the driver's Run: the reference's Run::initialize / Run::solver
// (Run.cpp:10-204) with the per-stage body replaced by the device hot path
// behind include/cfs.h.
#pragma once

#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cfs.h"
#include "param.h"
#include "plugins.h"

namespace cfsdrv {

struct RunOptions {
    int nx = N, ny = N;
    double lx = LengthX, ly = LengthY;
    double gamma = Gamma;
    int precision = CFS_PREC_FP64;
    Weno5th interpolation;
    RiemannSolverHLL riemann;
    TimeAdvance timeadvance;
    ViscousTerms viscous;
    std::unique_ptr<InitialCondition> initialcondition;
    bool immersed = true;
    Circle circle;
    int stages = Stages;         // fixed-dt mode: stage loop length (Run.cpp:153)
    double dt = cfsdrv::dt;      // > 0: fixed step; <= 0: CFL steps
    double cfl = 0.5;
    double t_end = 0.0;          // CFL mode: end time
    int output_every = OutputEvery;   // stages (fixed dt) or steps (CFL); 0: none
    std::string output_dir;      // empty: no snapshots
    std::string log_path;        // JSON-lines run log (the reference's per-stage timing print,
                                 // Run.cpp:164-232); empty: none
};

class Run {
public:
    explicit Run(RunOptions opt) : opt_(std::move(opt)) {}
    ~Run();
    // mesh + plugin choices -> device handle; immersed boundary set-up
    int initialize();
    // initial condition on fluid cells, then the stage loop (fixed dt) or CFL
    // steps to t_end, with snapshots
    int solver();
    const cfs_config& config() const { return cfg_; }
    double time() const { return t_; }
    int stages_done() const { return stages_; }

private:
    int snapshot(int index);
    // one run-log record: progress since the previous record, checksums of
    // the current state (the checksum call also waits for the stages to end)
    int log_record(int step);
    RunOptions opt_;
    cfs_config cfg_{};
    cfs_solver* h_ = nullptr;
    std::vector<uint8_t> cell_type_;
    double t_ = 0.0;
    int stages_ = 0;
    std::FILE* log_ = nullptr;
    double wall0_ = 0.0, wall_last_ = 0.0;   // seconds: start of the stage loop, previous record
    int stages_last_ = 0;
};

}  // namespace cfsdrv
