// main.cpp -- command line of the driver (the reference's main.cpp runs
// Run::initialize and Run::solver once, with the choices compiled in).
//
//   cfs_run [--case cylinder|riemann|lax-liu|kh] [--nx N] [--ny N] [--lx L] [--ly L]
//           [--rk 2|3] [--fp32] [--stages S] [--dt DT | --cfl C --t-end T]
//           [--every K] [--out DIR] [--positivity] [--velocity-bounds]
//           [--mu MU] [--radius R] [--cx X] [--cy Y] [--ib-reference] [--log FILE]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "run.h"

using namespace cfsdrv;

static void usage() {
    std::fprintf(stderr,
                 "usage: cfs_run [--case cylinder|riemann|lax-liu|kh] [--nx N] [--ny N] [--lx L] [--ly L]\n"
                 "               [--rk 2|3] [--fp32] [--stages S] [--dt DT | --cfl C --t-end T] [--every K]\n"
                 "               [--out DIR] [--positivity] [--velocity-bounds] [--mu MU]\n"
                 "               [--radius R] [--cx X] [--cy Y] [--ib-reference] [--log FILE]\n");
}

int main(int argc, char** argv) {
    RunOptions o;
    std::string cas = "cylinder";
    bool nx_set = false, ny_set = false;
    for (int a = 1; a < argc; ++a) {
        const std::string k = argv[a];
        auto val = [&]() -> const char* {
            if (a + 1 >= argc) { usage(); std::exit(2); }
            return argv[++a];
        };
        if (k == "--case") cas = val();
        else if (k == "--nx") { o.nx = std::atoi(val()); nx_set = true; }
        else if (k == "--ny") { o.ny = std::atoi(val()); ny_set = true; }
        else if (k == "--lx") o.lx = std::atof(val());
        else if (k == "--ly") o.ly = std::atof(val());
        else if (k == "--rk") o.timeadvance.rk = std::atoi(val());
        else if (k == "--fp32") o.precision = CFS_PREC_FP32_BULK;
        else if (k == "--stages") o.stages = std::atoi(val());
        else if (k == "--dt") o.dt = std::atof(val());
        else if (k == "--cfl") { o.cfl = std::atof(val()); o.dt = 0.0; }
        else if (k == "--t-end") o.t_end = std::atof(val());
        else if (k == "--every") o.output_every = std::atoi(val());
        else if (k == "--out") o.output_dir = val();
        else if (k == "--log") o.log_path = val();
        else if (k == "--positivity") o.interpolation.positivity = true;
        else if (k == "--velocity-bounds") o.riemann.velocity_bounds = true;
        else if (k == "--mu") { o.viscous.on = true; o.viscous.mu = std::atof(val()); }
        else if (k == "--radius") o.circle.radius = std::atof(val());
        else if (k == "--cx") o.circle.cx = std::atof(val());
        else if (k == "--cy") o.circle.cy = std::atof(val());
        else if (k == "--ib-reference") o.circle.interp = CFS_IB_REFERENCE;
        else if (k == "--help" || k == "-h") { usage(); return 0; }
        else { usage(); return 2; }
    }
    if (cas == "cylinder") {
        o.initialcondition = std::make_unique<CylinderFlow>();
    } else if (cas == "riemann") {   // the reference's Reimman2D on its 2 x 2 square
        o.initialcondition = std::make_unique<Riemann2D>();
        o.immersed = false;
    } else if (cas == "lax-liu") {   // Lax-Liu configuration 3 on the unit square, split (0.5, 0.5)
        auto r = std::make_unique<Riemann2D>();
        r->sx = r->sy = 0.5;
        const double sw[4] = {0.138, 0.138 * 1.206, 0.138 * 1.206, 0.029};
        const double nw[4] = {0.5323, 0.5323 * 1.206, 0.0, 0.3};
        const double se[4] = {0.5323, 0.0, 0.5323 * 1.206, 0.3};
        const double ne[4] = {1.5, 0.0, 0.0, 1.5};
        r->set(sw, nw, se, ne);
        o.initialcondition = std::move(r);
        o.immersed = false;
        if (!nx_set) o.nx = 400;
        if (!ny_set) o.ny = 400;
        o.lx = o.ly = 1.0;
    } else if (cas == "kh") {
        o.initialcondition = std::make_unique<KelvinHelmholtz>();
        o.immersed = false;
        o.lx = o.ly = 1.0;
        if (!o.viscous.on) { o.viscous.on = true; o.viscous.mu = 1e-5; }
    } else {
        usage();
        return 2;
    }
    Run run(std::move(o));
    int rc = run.initialize();
    if (rc == CFS_OK) rc = run.solver();
    if (rc != CFS_OK) {
        std::fprintf(stderr, "cfs_run: %s\n", cfs_status_string(rc));
        return 1;
    }
    std::printf("{\"nx\": %d, \"ny\": %d, \"stages\": %d, \"t\": %.17g}\n", run.config().nx, run.config().ny,
                run.stages_done(), run.time());
    return 0;
}
