// plugins.h -- the driver's strategy classes.  They keep the roles of the
// reference's plugins (Interpolation / RiemmanSolver / Explicit / Boundary /
// Initialcondition / ImmersedBoundary / ForthOrderCS, held as globals in
// Run.cpp:5-8) but do no per-cell work: each one only writes its choice into
// the cfs_config of the device handle, whose stage kernel does the work.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/cfs.h"

namespace cfsdrv {

// EulerPhase::computeEnergyfromOther (EulerPhase.cpp:124-128):
// E = p/(gamma-1) + 0.5*(mx*mx + my*my)/rho
inline void conserved(double gamma, double rho, double mx, double my, double p, double out[4]) {
    const double e = p / (gamma - 1);
    out[0] = rho;
    out[1] = mx;
    out[2] = my;
    out[3] = e + 0.5 * (mx * mx + my * my) / rho;
}

// ---- Interpolation (Weno5th.cpp) ----
struct Weno5th {
    bool positivity = false;   // inadmissible face state -> adjacent cell state (no reference counterpart)
    void configure(cfs_config& c) const { c.positivity = positivity ? 1 : 0; }
};

// ---- Riemann solver (RiemmanSolverHLL.cpp) ----
struct RiemannSolverHLL {
    bool velocity_bounds = false;   // false: the reference's momentum bounds (RiemmanSolverHLL.cpp:41-42)
    void configure(cfs_config& c) const { c.hll_bounds = velocity_bounds ? 1 : 0; }
};

// ---- time advance (TVD2rdRK.cpp / TVD3rdRK.cpp) ----
struct TimeAdvance {
    int rk = 2;   // Run.cpp:47 uses TVD2rdRK
    void configure(cfs_config& c) const { c.rk = rk; }
};

// ---- viscous terms (FourthOrderCS.cpp; not called by the reference driver) ----
struct ViscousTerms {
    bool on = false;
    double mu = 1.0e-3, prandtl = 0.72, ma_ref = 1.0;   // mu_ref, Ma_ref of param.h:10-12
    void configure(cfs_config& c) const {
        c.viscous = on ? 1 : 0;
        c.mu = mu;
        c.prandtl = prandtl;
        c.ma_ref = ma_ref;
    }
};

// ---- physical boundaries (source/boundary) ----
struct Boundary {
    int side = CFS_SIDE_LEFT;
    virtual ~Boundary() = default;
    virtual void configure(cfs_config& c) const = 0;
};
struct NeumannBoundary : Boundary {   // NeummanBoundary.cpp:31-45: ghost = adjacent cell
    explicit NeumannBoundary(int s) { side = s; }
    void configure(cfs_config& c) const override { c.bc[side] = CFS_BC_NEUMANN; }
};
struct DirichletBoundary : Boundary {   // DirichletBoundary.cpp:32-46: ghost = fixed state
    double state[4];
    DirichletBoundary(int s, const double st[4]) {
        side = s;
        for (int q = 0; q < 4; ++q) state[q] = st[q];
    }
    void configure(cfs_config& c) const override {
        c.bc[side] = CFS_BC_DIRICHLET;
        for (int q = 0; q < 4; ++q) c.bc_state[side][q] = state[q];
    }
};
struct PeriodicBoundary : Boundary {   // no reference counterpart (periods = {0,0})
    explicit PeriodicBoundary(int s) { side = s; }
    void configure(cfs_config& c) const override { c.bc[side] = CFS_BC_PERIODIC; }
};
using Boundaries = std::vector<std::unique_ptr<Boundary>>;

// ---- initial conditions (source/initialcondition) ----
struct InitialCondition {
    virtual ~InitialCondition() = default;
    // conserved state of cell (i, j) with centre (x, y)
    virtual void fillIn(double gamma, int i, int j, double x, double y, double out[4]) const = 0;
    virtual void setBoundaryCondition(double gamma, Boundaries& b) const = 0;
};

// Reimman2D.cpp:5-39: four quadrants split at (sx, sy); the reference's
// states in exact rationals and its split 1.6 by default.
struct Riemann2D : InitialCondition {
    double sx = 1.6, sy = 1.6;
    // (rho, rho*u, rho*v, p) per quadrant: SW, NW, SE, NE
    double q[4][4];
    Riemann2D() {
        const double r1 = 77.0 / 558.0, r2 = 33.0 / 62.0;
        const double sw[4] = {r1, r1 * 4 / std::sqrt(11.0), r1 * 4 / std::sqrt(11.0), 9.0 / 310.0};
        const double nw[4] = {r2, r2 * 4 / std::sqrt(11.0), 0.0, 0.3};
        const double se[4] = {r2, 0.0, r2 * 4 / std::sqrt(11.0), 0.3};
        const double ne[4] = {1.5, 0.0, 0.0, 1.5};
        set(sw, nw, se, ne);
    }
    void set(const double sw[4], const double nw[4], const double se[4], const double ne[4]) {
        for (int k = 0; k < 4; ++k) {
            q[0][k] = sw[k];
            q[1][k] = nw[k];
            q[2][k] = se[k];
            q[3][k] = ne[k];
        }
    }
    void fillIn(double gamma, int, int, double x, double y, double out[4]) const override {
        const int k = (x < sx) ? (y < sy ? 0 : 1) : (y < sy ? 2 : 3);
        conserved(gamma, q[k][0], q[k][1], q[k][2], q[k][3], out);
    }
    void setBoundaryCondition(double, Boundaries& b) const override {   // Reimman2D.cpp:41-51
        for (int s = CFS_SIDE_LEFT; s <= CFS_SIDE_TOP; ++s) b.push_back(std::make_unique<NeumannBoundary>(s));
    }
};

// Cylinder.cpp:3-40: uniform freestream; Dirichlet inflow on the left,
// Neumann elsewhere.
struct CylinderFlow : InitialCondition {
    double rho = 1.0, u = 1.0, v = 0.0, p = 1.0;
    void fillIn(double gamma, int, int, double, double, double out[4]) const override {
        conserved(gamma, rho, rho * u, rho * v, p, out);
    }
    void setBoundaryCondition(double gamma, Boundaries& b) const override {
        double st[4];
        conserved(gamma, rho, rho * u, rho * v, p, st);
        b.push_back(std::make_unique<DirichletBoundary>(CFS_SIDE_LEFT, st));
        b.push_back(std::make_unique<NeumannBoundary>(CFS_SIDE_RIGHT));
        b.push_back(std::make_unique<NeumannBoundary>(CFS_SIDE_BOTTOM));
        b.push_back(std::make_unique<NeumannBoundary>(CFS_SIDE_TOP));
    }
};

// SplitMix64 finaliser (Steele, Lea, Flood 2014)
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Viscous shear layer on the periodic unit square (no reference counterpart):
// rho = 2, u = +0.5 in 0.25 < y < 0.75, rho = 1, u = -0.5 elsewhere, p = 2.5,
// v = amp sin(4 pi x) + noise * (2 U - 1), U from splitmix64(splitmix64(seed)
// ^ (i << 32 | j)) (top 53 bits).
struct KelvinHelmholtz : InitialCondition {
    uint64_t seed = 0x5EED;
    double amp = 0.01, noise = 1e-3;
    void fillIn(double gamma, int i, int j, double x, double y, double out[4]) const override {
        const bool band = y > 0.25 && y < 0.75;
        const double rho = band ? 2.0 : 1.0, u = band ? 0.5 : -0.5;
        const uint64_t h = splitmix64(splitmix64(seed) ^ (((uint64_t)i << 32) | (uint64_t)j));
        const double r = 2.0 * ((double)(h >> 11) * 0x1.0p-53) - 1.0;
        const double v = amp * std::sin(4.0 * M_PI * x) + noise * r;
        conserved(gamma, rho, rho * u, rho * v, 2.5, out);
    }
    void setBoundaryCondition(double, Boundaries& b) const override {
        for (int s = CFS_SIDE_LEFT; s <= CFS_SIDE_TOP; ++s) b.push_back(std::make_unique<PeriodicBoundary>(s));
    }
};

// ---- immersed boundary (Circle.cpp; CartesianMesh::setImmersedBoundary) ----
struct Circle {
    double radius = 0.1, cx = 0.5, cy = 1.0;
    int points = 400, layers = 3;
    int interp = CFS_IB_BILINEAR_POSITIVE;
};

}  // namespace cfsdrv
