// cfs_oracle.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// A strict-IEEE FP64 CPU restatement of the reference solver's per-stage hot
// path (xinxiangsama/2d_Compressible_Flow_Solver).  It exists only to check
// the HIP library: tests, the smoke check and bench.py's CPU-baseline leg may
// load it, nothing in the product may.  Build: oracle/Makefile
// (g++ -O2 -ffp-contract=off, no fast-math: every operation is one IEEE
// double operation in the source order of the reference, C++ left-to-right
// associativity).
//
// It is deliberately written the way the reference computes, not the way the
// GPU does: per-face loops over an explicit 6-cell stencil produced by a
// restatement of the reference's face walk, explicit ghost-cell arrays
// refreshed after every stage, per-cell RK snapshot copies, and an LDL^T
// re-factorisation at every immersed-boundary fill.
//
// Parity status of each function (how it is pinned) is listed in the tests
// (tests/test_oracle_*.py).  Parity unpinned: the exact bits of Eigen's LDLT
// inner products (ib_ldlt_factor / ib_ldlt_solve restate the published
// algorithm with sequential sums; Eigen may sum in SIMD order) -- the
// reference cannot be built here (no Eigen), so no reference vectors exist.
#include "../include/cfs.h"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <thread>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// WENO5 -- source/interpolation/Weno5th.cpp:4-28 (eps = 1e-2, ideal weights
// 0.1/0.6/0.3, v[2] is the upwind cell).  Same operations, same order.
double weno5(const double v[5]) {
    const double eps = 1e-2;
    const double c1312 = 13.0 / 12.0;
    const double s0 = (v[0] - 2 * v[1]) + v[2];
    const double t0 = (v[0] - 4 * v[1]) + 3 * v[2];
    const double beta0 = c1312 * (s0 * s0) + 0.25 * (t0 * t0);
    const double s1 = (v[1] - 2 * v[2]) + v[3];
    const double t1 = v[1] - v[3];
    const double beta1 = c1312 * (s1 * s1) + 0.25 * (t1 * t1);
    const double s2 = (v[2] - 2 * v[3]) + v[4];
    const double t2 = (3 * v[2] - 4 * v[3]) + v[4];
    const double beta2 = c1312 * (s2 * s2) + 0.25 * (t2 * t2);

    const double alpha0 = 0.1 / ((eps + beta0) * (eps + beta0));
    const double alpha1 = 0.6 / ((eps + beta1) * (eps + beta1));
    const double alpha2 = 0.3 / ((eps + beta2) * (eps + beta2));
    const double sum = (alpha0 + alpha1) + alpha2;
    const double w0 = alpha0 / sum, w1 = alpha1 / sum, w2 = alpha2 / sum;

    const double q0 = (v[0] / 3.0 - 7.0 * v[1] / 6.0) + 11.0 * v[2] / 6.0;
    const double q1 = (-v[1] / 6.0 + 5.0 * v[2] / 6.0) + v[3] / 3.0;
    const double q2 = (v[2] / 3.0 + 5.0 * v[3] / 6.0) - v[4] / 6.0;
    return (w0 * q0 + w1 * q1) + w2 * q2;
}

struct Phase {            // the fields of EulerPhase the hot path reads
    double rho, mu, mv, E, p;   // mu/mv are momenta (EulerPhase.h:52)
};

// EulerPhase::computePressurefromOther (EulerPhase.cpp:130-133)
double pressure(double g, double rho, double mu, double mv, double E) {
    return (g - 1) * (E - 0.5 * (mu * mu + mv * mv) / rho);
}
// EulerPhase::computeEnergyfromOther (EulerPhase.cpp:124-128)
double energy(double g, double rho, double mu, double mv, double p) {
    const double e = p / (g - 1);
    return e + 0.5 * (mu * mu + mv * mv) / rho;
}

// EulerMod::comFlux (EulerMod.cpp:33-54)
void euler_flux(double g, const Phase& s, bool xdir, double F[4]) {
    const double p = pressure(g, s.rho, s.mu, s.mv, s.E);
    if (xdir) {
        F[0] = s.mu;
        F[1] = s.mu * s.mu / s.rho + p;
        F[2] = s.mu * s.mv / s.rho;
        F[3] = (s.E + p) * (s.mu / s.rho);
    } else {
        F[0] = s.mv;
        F[1] = s.mu * s.mv / s.rho;
        F[2] = s.mv * s.mv / s.rho + p;
        F[3] = (s.E + p) * (s.mv / s.rho);
    }
}

double std_min(double a, double b) { return (b < a) ? b : a; }   // std::min
double std_max(double a, double b) { return (a < b) ? b : a; }   // std::max

// RiemmanSolverHLL::Solver (RiemmanSolverHLL.cpp:5-75), including its use of
// the momentum (not the velocity) in the outer wave-speed bounds and the
// left state in SR.
void hll(double g, const Phase& L, const Phase& R, const double FL[4],
         const double FR[4], bool xdir, double out[4], int bounds = 0) {
    const double uL[2] = {L.mu, L.mv}, uR[2] = {R.mu, R.mv};
    const double aL = std::sqrt(g * L.p / L.rho);
    const double aR = std::sqrt(g * R.p / R.rho);
    const int k = xdir ? 0 : 1;
    const double a_ave = 0.5 * (aL + aR);
    const double rho_ave = 0.5 * (L.rho + R.rho);
    const double p_star = 0.5 * (L.p + R.p) - 0.5 * (uR[k] / R.rho - uL[k] / L.rho) * rho_ave * a_ave;
    const double u_star = 0.5 * (uR[k] / R.rho + uL[k] / L.rho) - 0.5 * (R.p - L.p) / (rho_ave * a_ave);
    const double rs_L = L.rho + (uL[k] / L.rho - u_star) * rho_ave / a_ave;
    const double rs_R = R.rho - (uR[k] / R.rho - u_star) * rho_ave / a_ave;
    const double as_L = std::sqrt(g * p_star / rs_L);
    const double as_R = std::sqrt(g * p_star / rs_R);
    double SL = std_min(uL[k] - aL, u_star - as_L);
    double SR = std_max(uL[k] + aL, u_star + as_R);
    if (bounds == 1) {   // velocity bounds, right state in SR (not the reference)
        SL = std_min(uL[k] / L.rho - aL, u_star - as_L);
        SR = std_max(uR[k] / R.rho + aR, u_star + as_R);
    }
    if (SL >= 0) {
        for (int q = 0; q < 4; ++q) out[q] = FL[q];
    } else if (SR <= 0) {
        for (int q = 0; q < 4; ++q) out[q] = FR[q];
    } else {
        const double inv = 1.0 / (SR - SL);
        out[0] = (SR * FL[0] - SL * FR[0] + SL * SR * (R.rho - L.rho)) * inv;
        out[1] = (SR * FL[1] - SL * FR[1] + SL * SR * (uR[0] - uL[0])) * inv;
        out[2] = (SR * FL[2] - SL * FR[2] + SL * SR * (uR[1] - uL[1])) * inv;
        out[3] = (SR * FL[3] - SL * FR[3] + SL * SR * (R.E - L.E)) * inv;
    }
}

// ---------------------------------------------------------------------------
// Immersed boundary -- source/object/Circle.cpp, ImagePoint.cpp.
struct Vec2 { double x, y; };
double norm2(Vec2 v) { return std::sqrt(v.x * v.x + v.y * v.y); }          // Eigen norm()
Vec2 normalized(Vec2 v) {                                                  // Eigen normalized()
    const double z = v.x * v.x + v.y * v.y;
    if (z > 0.0) { const double s = std::sqrt(z); return {v.x / s, v.y / s}; }
    return v;
}

// Eigen::LDLT<Matrix4d, Lower>::compute restated (pivot on the largest
// remaining |diagonal|, first one on ties; only the lower triangle of A is
// read or written).  A is column-major 4x4 stored as A[row][col].
void ib_ldlt_factor(double A[4][4], int transp[4]) {
    const int n = 4;
    for (int k = 0; k < n; ++k) {
        int big = k;
        double bigv = std::fabs(A[k][k]);
        for (int i = k + 1; i < n; ++i) {
            if (std::fabs(A[i][i]) > bigv) { bigv = std::fabs(A[i][i]); big = i; }
        }
        transp[k] = big;
        if (big != k) {
            for (int c = 0; c < k; ++c) std::swap(A[k][c], A[big][c]);
            for (int r = big + 1; r < n; ++r) std::swap(A[r][k], A[r][big]);
            std::swap(A[k][k], A[big][big]);
            for (int i = k + 1; i < big; ++i) {
                const double tmp = A[i][k];
                A[i][k] = A[big][i];
                A[big][i] = tmp;
            }
        }
        double temp[4];
        if (k > 0) {
            for (int i = 0; i < k; ++i) temp[i] = A[i][i] * A[k][i];
            double dot = 0.0;
            for (int i = 0; i < k; ++i) dot = (i == 0) ? A[k][i] * temp[i] : dot + A[k][i] * temp[i];
            A[k][k] -= dot;
            for (int r = k + 1; r < n; ++r) {
                double s = 0.0;
                for (int i = 0; i < k; ++i) s = (i == 0) ? A[r][i] * temp[i] : s + A[r][i] * temp[i];
                A[r][k] -= s;
            }
        }
        const double akk = A[k][k];
        if (akk != 0.0) {
            for (int r = k + 1; r < n; ++r) A[r][k] /= akk;
        }
    }
}

// Eigen LDLT::solve restated: P b, unit-lower solve (column sweep), D^-1 with
// the pseudo-inverse cut at DBL_MIN, unit-upper (L^T) solve (row dots),
// P^T.
void ib_ldlt_solve(const double A[4][4], const int transp[4], double x[4]) {
    const int n = 4;
    for (int k = 0; k < n; ++k) std::swap(x[k], x[transp[k]]);
    for (int c = 0; c < n; ++c)
        for (int r = c + 1; r < n; ++r) x[r] -= x[c] * A[r][c];
    for (int i = 0; i < n; ++i) {
        if (std::fabs(A[i][i]) > std::numeric_limits<double>::min()) x[i] /= A[i][i];
        else x[i] = 0.0;
    }
    for (int i = n - 1; i >= 0; --i) {
        if (i == n - 1) continue;
        double s = 0.0;
        for (int j = i + 1; j < n; ++j) s = (j == i + 1) ? A[j][i] * x[j] : s + A[j][i] * x[j];
        x[i] -= s;
    }
    for (int k = n - 1; k >= 0; --k) std::swap(x[k], x[transp[k]]);
}

struct ImmersedGhost {
    int cell;          // i*ny + j
    Vec2 ip;           // image point
    int donor[4];      // i*ny + j, in distance order
    double w[4];       // donor weights (positivity-preserving interpolation)
};

// ---------------------------------------------------------------------------
struct Oracle {
    cfs_config cfg{};
    int nx = 0, ny = 0;
    double dx = 0, dy = 0, g = 1.4;
    std::vector<double> U;            // [4][nx][ny]
    std::vector<double> ghost[4];     // per side: [4][len]  (len = ny for L/R, nx for B/T)
    std::vector<uint8_t> type;        // CFS_CELL_*
    std::vector<double> snap0, snap1; // RK snapshots (TVD*RK m_Phase*)
    std::vector<double> FX, FY;       // face fluxes [(nx+1)*ny][4], [nx*(ny+1)][4]
    std::vector<double> RHS;          // last cell RHS [4][nx][ny] (Cell::m_flux)
    std::vector<ImmersedGhost> ig;
    int ib_interp = CFS_IB_REFERENCE;
    int counter = 1;
    int nthreads = 1;
    int ranks = 1, dims[2] = {1, 1};  // multi-rank mode (stage_ranks)

    size_t at(int q, int i, int j) const { return (size_t)q * nx * ny + (size_t)i * ny + j; }
    bool periodic_x() const { return cfg.bc[0] == CFS_BC_PERIODIC; }
    bool periodic_y() const { return cfg.bc[2] == CFS_BC_PERIODIC; }
};

template <class F>
void parallel_for(int n, int nthreads, F f) {
    if (nthreads <= 1 || n < 2 * nthreads) { f(0, n); return; }
    std::vector<std::thread> th;
    const int chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([=]() { f(a, b); });
    }
    for (auto& t : th) t.join();
}

// Where the face phase reads interior cell states: the global state, or -- in
// the multi-rank mode -- the reading rank's own block plus its exchanged
// 3-cell ghost frame (set per worker thread).
struct View {
    const double* U;
    size_t qs;
    int ld, i0, j0;
    double at(int q, int i, int j) const { return U[q * qs + (size_t)(i - i0) * ld + (j - j0)]; }
};
thread_local const View* tl_view = nullptr;
inline double ucell(const Oracle& o, int q, int i, int j) {
    return tl_view ? tl_view->at(q, i, j) : o.U[o.at(q, i, j)];
}

// CellInterface::collectCells (CellInterface.cpp:135-195) restated on
// indices: along one grid line with cells 0..N-1 and one ghost cell at -1 and
// at N, the face f lies between cells f-1 and f; the walk takes the left
// cell of the backward face and the right cell of the forward face three
// times and does not step past a boundary face (f = 0 or f = N).
void collect_stencil(int f, int N, bool periodic, int out[6]) {
    if (periodic) {
        for (int k = 0; k < 6; ++k) out[k] = ((f - 3 + k) % N + N) % N;
        return;
    }
    std::deque<int> cells;
    int back = f, fwd = f;
    for (int k = 0; k < 3; ++k) {
        cells.push_front(back - 1);
        cells.push_back(fwd);
        if (!(back == 0 || back == N)) back = back - 1;
        if (!(fwd == 0 || fwd == N)) fwd = fwd + 1;
    }
    for (int k = 0; k < 6; ++k) out[k] = cells[k];
}

// The multi-rank mode's stencil: periodic indices stay unwrapped (the frame
// holds the wrapped copies there); otherwise the reference's face walk.
void collect_stencil_rank(int f, int N, bool periodic, int out[6]) {
    if (periodic) {
        for (int k = 0; k < 6; ++k) out[k] = f - 3 + k;
        return;
    }
    collect_stencil(f, N, false, out);
}

Phase cell_phase_x(const Oracle& o, int i, int j) {
    // i in [-1, nx]: -1 and nx are the left/right ghost cells of row j
    Phase s{};
    double v[4];
    // (periodic axes never reach a ghost: their stencils are wrapped, or, in
    // the multi-rank mode, read from the rank's frame at unwrapped indices)
    if (i == -1 && !o.periodic_x()) for (int q = 0; q < 4; ++q) v[q] = o.ghost[0][(size_t)q * o.ny + j];
    else if (i == o.nx && !o.periodic_x()) for (int q = 0; q < 4; ++q) v[q] = o.ghost[1][(size_t)q * o.ny + j];
    else for (int q = 0; q < 4; ++q) v[q] = ucell(o, q, i, j);
    s.rho = v[0]; s.mu = v[1]; s.mv = v[2]; s.E = v[3];
    return s;
}
Phase cell_phase_y(const Oracle& o, int i, int j) {
    Phase s{};
    double v[4];
    if (j == -1 && !o.periodic_y()) for (int q = 0; q < 4; ++q) v[q] = o.ghost[2][(size_t)q * o.nx + i];
    else if (j == o.ny && !o.periodic_y()) for (int q = 0; q < 4; ++q) v[q] = o.ghost[3][(size_t)q * o.nx + i];
    else for (int q = 0; q < 4; ++q) v[q] = ucell(o, q, i, j);
    s.rho = v[0]; s.mu = v[1]; s.mv = v[2]; s.E = v[3];
    return s;
}

// Weno5th::comLeftPhase / comRightPhase + comInterpolation
// (Weno5th.cpp:30-81): left state from list[0..4], right state from
// list[5..1]; pressure recomputed from the reconstructed conserved state.
void face_flux(const Oracle& o, const Phase st[6], bool xdir, double out[4]) {
    Phase L{}, R{};
    double v[5];
    const double Phase::*field[4] = {&Phase::rho, &Phase::mu, &Phase::mv, &Phase::E};
    double left[4], right[4];
    for (int q = 0; q < 4; ++q) {
        for (int k = 0; k < 5; ++k) v[k] = st[k].*field[q];
        left[q] = weno5(v);
        for (int k = 0; k < 5; ++k) v[k] = st[5 - k].*field[q];
        right[q] = weno5(v);
    }
    L.rho = left[0]; L.mu = left[1]; L.mv = left[2]; L.E = left[3];
    R.rho = right[0]; R.mu = right[1]; R.mv = right[2]; R.E = right[3];
    L.p = pressure(o.g, L.rho, L.mu, L.mv, L.E);
    R.p = pressure(o.g, R.rho, R.mu, R.mv, R.E);
    if (o.cfg.positivity) {
        // inadmissible reconstruction -> state of the cell on that side of the face
        if (!(L.rho > 0.0 && L.p > 0.0)) {
            L = st[2];
            L.p = pressure(o.g, L.rho, L.mu, L.mv, L.E);
        }
        if (!(R.rho > 0.0 && R.p > 0.0)) {
            R = st[3];
            R.p = pressure(o.g, R.rho, R.mu, R.mv, R.E);
        }
    }
    double FL[4], FR[4];
    euler_flux(o.g, L, xdir, FL);   // CellInterface::computeFlux (CellInterface.cpp:68-113)
    euler_flux(o.g, R, xdir, FR);
    hll(o.g, L, R, FL, FR, xdir, out, o.cfg.hll_bounds);
}

// Boundary::EnforceBoundary for every boundary face
// (CartesianMesh::ApplyBoundaryCondition, CartesianMesh.cpp:280-305).
void apply_bc(Oracle& o) {
    for (int side = 0; side < 4; ++side) {
        const int kind = o.cfg.bc[side];
        if (kind == CFS_BC_PERIODIC) continue;
        const int len = (side < 2) ? o.ny : o.nx;
        for (int t = 0; t < len; ++t) {
            for (int q = 0; q < 4; ++q) {
                double val;
                if (kind == CFS_BC_DIRICHLET) {
                    val = o.cfg.bc_state[side][q];
                } else {   // Neumann: copy of the adjacent interior cell (EulerPhase::cpyPhase)
                    if (side == 0) val = o.U[o.at(q, 0, t)];
                    else if (side == 1) val = o.U[o.at(q, o.nx - 1, t)];
                    else if (side == 2) val = o.U[o.at(q, t, 0)];
                    else val = o.U[o.at(q, t, o.ny - 1)];
                }
                o.ghost[side][(size_t)q * len + t] = val;
            }
        }
    }
}

// Cell::comImmersedGhostPhase -> ImagePoint::InterpolatePhase
// (Cell.cpp:172-181, ImagePoint.cpp:20-82)
void ib_fill_one(Oracle& o, const ImmersedGhost& gcell) {
    const double dx = o.dx, dy = o.dy;
    if (o.ib_interp == CFS_IB_BILINEAR_POSITIVE) {
        double val[4];
        for (int q = 0; q < 4; ++q) {
            double f[4];
            for (int d = 0; d < 4; ++d) f[d] = o.U[(size_t)q * o.nx * o.ny + gcell.donor[d]];
            val[q] = gcell.w[0] * f[0] + gcell.w[1] * f[1] + gcell.w[2] * f[2] + gcell.w[3] * f[3];
        }
        const int i = gcell.cell / o.ny, j = gcell.cell % o.ny;
        o.U[o.at(0, i, j)] = val[0];
        o.U[o.at(1, i, j)] = -val[1];
        o.U[o.at(2, i, j)] = -val[2];
        o.U[o.at(3, i, j)] = val[3];
        return;
    }
    double A[4][4];
    double f[4][4];   // [var][donor]
    for (int d = 0; d < 4; ++d) {
        const int c = gcell.donor[d];
        const int i = c / o.ny, j = c % o.ny;
        const double px = (i + 0.5) * dx + o.cfg.x0;
        const double py = (j + 0.5) * dy + o.cfg.y0;
        A[d][0] = px * py; A[d][1] = px; A[d][2] = py; A[d][3] = 1.0;
        for (int q = 0; q < 4; ++q) f[q][d] = o.U[o.at(q, i, j)];
    }
    int tr[4];
    ib_ldlt_factor(A, tr);
    double val[4];
    for (int q = 0; q < 4; ++q) {
        double w[4] = {f[q][0], f[q][1], f[q][2], f[q][3]};
        ib_ldlt_solve(A, tr, w);
        val[q] = w[0] * gcell.ip.x * gcell.ip.y + w[1] * gcell.ip.x + w[2] * gcell.ip.y + w[3];
    }
    const int i = gcell.cell / o.ny, j = gcell.cell % o.ny;
    o.U[o.at(0, i, j)] = val[0];
    o.U[o.at(1, i, j)] = -val[1];
    o.U[o.at(2, i, j)] = -val[2];
    o.U[o.at(3, i, j)] = val[3];
}

void ib_fill(Oracle& o) {
    for (const ImmersedGhost& gcell : o.ig) ib_fill_one(o, gcell);
}

// TVD2rdRK::advance (TVD2rdRK.cpp:3-31), TVD3rdRK::advance (TVD3rdRK.cpp:3-48)
// with their coefficient tables (TVD2rdRK.h:17-20, TVD3rdRK.h:14-18).
int rk_substage(int rk, int counter) {
    if (rk == 2) return (counter % 2 == 0) ? 1 : 0;
    return (counter % 3 == 0) ? 2 : (counter % 3 == 1) ? 0 : 1;
}

// One cell, one variable: the stage formula with the snapshot bookkeeping of
// the reference's m_Phase / m_Phase0 / m_Phase1 maps.
void rk_update(int rk, int sub, double& u, double& s0, double& s1, double rhs, double dt) {
    static const double c2[2][3] = {{0.0, 1.0, 1.0}, {0.5, 0.5, 0.5}};
    static const double c3[3][3] = {{0.0, 1.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    const double* c = (rk == 2) ? c2[sub] : c3[sub];
    if (sub == 0) {
        s0 = u;
        u = c[1] * s0 + c[2] * dt * rhs;
    } else if (rk == 2) {
        u = c[0] * s0 + c[1] * u + c[2] * dt * rhs;
    } else if (sub == 1) {
        s1 = u;
        u = c[0] * s0 + c[1] * s1 + c[2] * dt * rhs;
    } else {
        u = c[0] * s0 + c[1] * s1 + c[2] * dt * rhs;
    }
}

// ---------------------------------------------------------------------------
// Viscous terms (no live reference counterpart: source/gradient/FourthOrderCS.cpp
// is never called, Run.cpp:177,188, and its physics is incomplete).  Textbook
// compressible Navier-Stokes with constant mu and a Prandtl number, T =
// gamma Ma^2 p / rho (EulerPhase.cpp:137), on the reference's stencil tables
// (FourthOrderCS.h:27-42) with the interpolation row for i-3/2 corrected to
// the Lagrange weights [5,15,-5,1]/16.

const double kThetaR[3][6] = {
    {71.0 / 1920, -141.0 / 128, 69.0 / 64, 1.0 / 192, -3.0 / 128, 3.0 / 640},
    {-3.0 / 640, 25.0 / 384, -75.0 / 64, 75.0 / 64, -25.0 / 384, 3.0 / 640},
    {-3.0 / 640, 3.0 / 128, -1.0 / 192, -69.0 / 64, 141.0 / 128, -71.0 / 1920}};
const double kThetaC[5] = {1.0 / 12, -8.0 / 12, 0.0, 8.0 / 12, -1.0 / 12};
const double kZetaR[3][4] = {
    {5.0 / 16, 15.0 / 16, -5.0 / 16, 1.0 / 16},
    {-1.0 / 16, 9.0 / 16, 9.0 / 16, -1.0 / 16},
    {1.0 / 16, -5.0 / 16, 15.0 / 16, 5.0 / 16}};

// Any cell index, including corners outside both axes: periodic axes wrap;
// otherwise the cell is a ghost of the first side found (x before y):
// Dirichlet = the fixed state, Neumann = the interior cell at the clamped index.
// (In the multi-rank mode a periodic index is not wrapped: the rank's frame
// holds the wrapped copy at the unwrapped index.)
void cell_any(const Oracle& o, int i, int j, double v[4]) {
    int side = -1;
    const bool wrap = tl_view == nullptr;
    if (i < 0) { if (o.periodic_x()) { if (wrap) i += o.nx; } else side = 0; }
    else if (i >= o.nx) { if (o.periodic_x()) { if (wrap) i -= o.nx; } else side = 1; }
    if (j < 0) { if (o.periodic_y()) { if (wrap) j += o.ny; } else if (side < 0) side = 2; }
    else if (j >= o.ny) { if (o.periodic_y()) { if (wrap) j -= o.ny; } else if (side < 0) side = 3; }
    if (side >= 0 && o.cfg.bc[side] == CFS_BC_DIRICHLET) {
        for (int q = 0; q < 4; ++q) v[q] = o.cfg.bc_state[side][q];
        return;
    }
    if (!o.periodic_x()) i = i < 0 ? 0 : (i >= o.nx ? o.nx - 1 : i);
    if (!o.periodic_y()) j = j < 0 ? 0 : (j >= o.ny ? o.ny - 1 : j);
    for (int q = 0; q < 4; ++q) v[q] = ucell(o, q, i, j);
}

struct Prim { double u, v, T; };

Prim primitive(const Oracle& o, const double s[4]) {
    const double ma = o.cfg.ma_ref;
    const double p = pressure(o.g, s[0], s[1], s[2], s[3]);
    return {s[1] / s[0], s[2] / s[0], o.g * (ma * ma) * p / s[0]};
}

Prim prim_any(const Oracle& o, int i, int j) {
    double s[4];
    cell_any(o, i, j, s);
    return primitive(o, s);
}

double dot_n(const double* c, const double* v, int n) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += c[k] * v[k];
    return s;
}

// 4th-order central cell gradient (FourthOrderCS.cpp:120-151 stencil)
// of u and v along x (xdir) or y at cell (i, j)
void cell_grad(const Oracle& o, int i, int j, bool xdir, double& du, double& dv) {
    double u[5], v[5];
    for (int k = 0; k < 5; ++k) {
        const Prim p = xdir ? prim_any(o, i - 2 + k, j) : prim_any(o, i, j - 2 + k);
        u[k] = p.u;
        v[k] = p.v;
    }
    const double h = xdir ? o.dx : o.dy;
    du = dot_n(kThetaC, u, 5) / h;
    dv = dot_n(kThetaC, v, 5) / h;
}

// Viscous flux through a face with stencil cells idx[0..5] along the line
// (i = idx, fixed j for x faces; fixed i, j = idx for y faces), written in x
// form: [0, tau_xx, tau_xy, u tau_xx + v tau_xy + k dT/dx] for x faces and
// [0, tau_xy, tau_yy, u tau_xy + v tau_yy + k dT/dy] for y faces, each
// evaluated at i-3/2, i-1/2, i+1/2 and combined with (-1, 26, -1)/24
// (FourthOrderCS.cpp:77-112).
void viscous_flux(const Oracle& o, const int idx[6], int line, bool xdir, double out[4]) {
    const double mu = o.cfg.mu, ma = o.cfg.ma_ref;
    const double kappa = mu / ((o.g - 1) * (ma * ma) * o.cfg.prandtl);
    double u[6], v[6], T[6], gu[4], gv[4];
    for (int k = 0; k < 6; ++k) {
        const Prim p = xdir ? prim_any(o, idx[k], line) : prim_any(o, line, idx[k]);
        u[k] = p.u; v[k] = p.v; T[k] = p.T;
    }
    for (int m = 0; m < 4; ++m) {   // cross derivatives at stencil cells 1..4
        if (xdir) cell_grad(o, idx[1 + m], line, false, gu[m], gv[m]);
        else cell_grad(o, line, idx[1 + m], true, gu[m], gv[m]);
    }
    const double h = xdir ? o.dx : o.dy;
    double F[3][3];
    for (int r = 0; r < 3; ++r) {
        const double dudn = dot_n(kThetaR[r], u, 6) / h;
        const double dvdn = dot_n(kThetaR[r], v, 6) / h;
        const double dTdn = dot_n(kThetaR[r], T, 6) / h;
        const double dudt = dot_n(kZetaR[r], gu, 4);
        const double dvdt = dot_n(kZetaR[r], gv, 4);
        const double uf = dot_n(kZetaR[r], u + 1, 4);
        const double vf = dot_n(kZetaR[r], v + 1, 4);
        if (xdir) {
            const double txx = mu * ((4.0 / 3.0) * dudn - (2.0 / 3.0) * dvdt);
            const double txy = mu * (dudt + dvdn);
            F[r][0] = txx;
            F[r][1] = txy;
            F[r][2] = (uf * txx + vf * txy) + kappa * dTdn;
        } else {
            const double tyy = mu * ((4.0 / 3.0) * dvdn - (2.0 / 3.0) * dudt);
            const double txy = mu * (dudn + dvdt);
            F[r][0] = txy;
            F[r][1] = tyy;
            F[r][2] = (uf * txy + vf * tyy) + kappa * dTdn;
        }
    }
    out[0] = 0.0;
    for (int q = 0; q < 3; ++q)
        out[q + 1] = ((-1.0 / 24.0) * F[0][q] + (26.0 / 24.0) * F[1][q]) + (-1.0 / 24.0) * F[2][q];
}

// ---------------------------------------------------------------------------
// Positivity-preserving flux limiter (positivity = 2; no reference
// counterpart; Hu, Adams & Shu, J. Comput. Phys. 242 (2013) 169-180), in the
// face-normal frame (n: normal momentum, t: tangential): F = F_LF +
// theta (F_hat - F_LF), theta the largest value in [0, 1] keeping both
// adjacent cells' shares U -+ k (F - F(U)) admissible, k = 2 (ln an + lt at)/an.
// theta = 1 at once when both shares taken with F_hat keep rho >= 1e-13 and
// rho p >= 1e-13 rho (a floor every limited share also respects).

double maxr(double x, double y) { return (x < y) ? y : x; }

// density times pressure without a division: rho p = (g-1) (E rho - (n^2 + t^2)/2)
double rho_p(double g, double r, double n, double t, double E) {
    return (g - 1) * (E * r - 0.5 * (n * n + t * t));
}

// one division: ir = 1/r, un = n*ir, ut = t*ir, p = (g-1)*(E - 0.5*(n*un + t*ut))
void euler_flux_n(double g, const double U[4], double F[4]) {
    const double ir = 1.0 / U[0];
    const double un = U[1] * ir, ut = U[2] * ir;
    const double p = (g - 1) * (U[3] - 0.5 * (U[1] * un + U[2] * ut));
    F[0] = U[1];
    F[1] = U[1] * un + p;
    F[2] = U[1] * ut;
    F[3] = (U[3] + p) * un;
}

// |u| + c (xdir) or |v| + c of any cell, ghosts included; ir = 1/rho,
// p = (g-1)*(E - 0.5*(mx*ux + my*uy)), c = sqrt((g*p)*ir)
double wave_speed(const Oracle& o, int i, int j, bool xdir) {
    double s[4];
    cell_any(o, i, j, s);
    const double ir = 1.0 / s[0];
    const double ux = s[1] * ir, uy = s[2] * ir;
    const double p = (o.g - 1) * (s[3] - 0.5 * (s[1] * ux + s[2] * uy));
    const double c = std::sqrt((o.g * p) * ir);
    const double u = xdir ? ux : uy;
    return (u < 0.0 ? -u : u) + c;
}

// the pressure test p1 < min(p0, 1e-13) is made on rho p, multiplied out
// (positive densities); the pressures themselves only when limiting
double share_theta(double g, const double B0[4], const double Dl[4]) {
    const double eps = 1e-13;
    const double r0 = B0[0];
    const double P0 = rho_p(g, B0[0], B0[1], B0[2], B0[3]);
    const double er = (r0 < eps) ? r0 : eps;
    double th = 1.0;
    const double r1 = B0[0] + Dl[0];
    if (r1 < er) th = (r0 - er) / (r0 - r1);
    double B1[4];
    for (int q = 0; q < 4; ++q) B1[q] = B0[q] + th * Dl[q];
    const double P1 = rho_p(g, B1[0], B1[1], B1[2], B1[3]);
    const bool low0 = P0 < eps * r0;
    const bool below = low0 ? (P1 * r0 < P0 * B1[0]) : (P1 < eps * B1[0]);
    if (below) {
        const double p0 = P0 / r0, p1 = P1 / B1[0];
        const double ep = (p0 < eps) ? p0 : eps;
        th = th * ((p0 - ep) / (p0 - p1));
    }
    return th;
}

// face between cells (iL, jL) and (iR, jR) (stencil entries 2 and 3), normal
// along x (xdir) or y; F in the x frame, limited in place
void limit_face(const Oracle& o, const Phase& cL, const Phase& cR, int iL, int jL, int iR, int jR, bool xdir,
                double dt, double F[4]) {
    const double g = o.g;
    const int nrm = xdir ? 1 : 2, tng = xdir ? 2 : 1;
    const double UL[4] = {cL.rho, xdir ? cL.mu : cL.mv, xdir ? cL.mv : cL.mu, cL.E};
    const double UR[4] = {cR.rho, xdir ? cR.mu : cR.mv, xdir ? cR.mv : cR.mu, cR.E};
    const double fh[4] = {F[0], F[nrm], F[tng], F[3]};
    double FL[4], FR[4];
    euler_flux_n(g, UL, FL);
    euler_flux_n(g, UR, FR);
    const int di = xdir ? 1 : 0, dj = xdir ? 0 : 1;   // neighbour offsets along the normal
    auto share_k = [&](int i, int j) {
        const double an = maxr(maxr(wave_speed(o, i - di, j - dj, xdir), wave_speed(o, i, j, xdir)),
                               wave_speed(o, i + di, j + dj, xdir));
        const double at = maxr(maxr(wave_speed(o, i - dj, j - di, !xdir), wave_speed(o, i, j, !xdir)),
                               wave_speed(o, i + dj, j + di, !xdir));
        const double ln = dt / (xdir ? o.dx : o.dy), lt = dt / (xdir ? o.dy : o.dx);
        return 2.0 * (ln * an + lt * at) / an;
    };
    const double kL = share_k(iL, jL), kR = share_k(iR, jR);
    // both shares with the face's own flux above the fixed floor (rho >= eps,
    // rho p >= eps rho, eps = 1e-13): theta = 1, the flux stays
    const double eps = 1e-13;
    double HL[4], HR[4];
    for (int q = 0; q < 4; ++q) {
        HL[q] = UL[q] - kL * (fh[q] - FL[q]);
        HR[q] = UR[q] + kR * (fh[q] - FR[q]);
    }
    const bool accL = HL[0] >= eps && rho_p(g, HL[0], HL[1], HL[2], HL[3]) >= eps * HL[0];
    const bool accR = HR[0] >= eps && rho_p(g, HR[0], HR[1], HR[2], HR[3]) >= eps * HR[0];
    if (accL && accR) return;
    const double a = maxr(wave_speed(o, iL, jL, xdir), wave_speed(o, iR, jR, xdir));
    double Flf[4], D[4];
    for (int q = 0; q < 4; ++q) {
        Flf[q] = 0.5 * (FL[q] + FR[q]) - (0.5 * a) * (UR[q] - UL[q]);
        D[q] = fh[q] - Flf[q];
    }
    double B0[4], Dl[4];
    double thL = 1.0, thR = 1.0;
    if (!accL) {
        for (int q = 0; q < 4; ++q) {
            B0[q] = UL[q] - kL * (Flf[q] - FL[q]);
            Dl[q] = -(kL * D[q]);
        }
        thL = share_theta(g, B0, Dl);
    }
    if (!accR) {
        for (int q = 0; q < 4; ++q) {
            B0[q] = UR[q] + kR * (Flf[q] - FR[q]);
            Dl[q] = kR * D[q];
        }
        thR = share_theta(g, B0, Dl);
    }
    double th = (thR < thL) ? thR : thL;
    th = (th < 0.0) ? 0.0 : th;
    if (th < 1.0) {
        double fn[4];
        for (int q = 0; q < 4; ++q) fn[q] = Flf[q] + th * D[q];
        F[0] = fn[0];
        F[nrm] = fn[1];
        F[tng] = fn[2];
        F[3] = fn[3];
    }
}

void stage(Oracle& o, int counter, double dt) {
    const int nx = o.nx, ny = o.ny;
    ib_fill(o);
    // x faces (horizontal interfaces), then y faces: WENO + flux + HLL
    parallel_for(ny, o.nthreads, [&](int j0, int j1) {
        int idx[6];
        Phase st[6];
        for (int j = j0; j < j1; ++j)
            for (int f = 0; f <= nx; ++f) {
                collect_stencil(f, nx, o.periodic_x(), idx);
                for (int k = 0; k < 6; ++k) st[k] = cell_phase_x(o, idx[k], j);
                double* F = &o.FX[((size_t)f * ny + j) * 4];
                face_flux(o, st, true, F);
                if (o.cfg.positivity == 2) limit_face(o, st[2], st[3], idx[2], j, idx[3], j, true, dt, F);
                if (o.cfg.viscous) {   // net flux = inviscid - viscous
                    double Fv[4];
                    viscous_flux(o, idx, j, true, Fv);
                    for (int q = 0; q < 4; ++q) F[q] = F[q] - Fv[q];
                }
            }
    });
    parallel_for(nx, o.nthreads, [&](int i0, int i1) {
        int idx[6];
        Phase st[6];
        for (int i = i0; i < i1; ++i)
            for (int gf = 0; gf <= ny; ++gf) {
                collect_stencil(gf, ny, o.periodic_y(), idx);
                for (int k = 0; k < 6; ++k) st[k] = cell_phase_y(o, i, idx[k]);
                double* F = &o.FY[((size_t)i * (ny + 1) + gf) * 4];
                face_flux(o, st, false, F);
                if (o.cfg.positivity == 2) limit_face(o, st[2], st[3], i, idx[2], i, idx[3], false, dt, F);
                if (o.cfg.viscous) {
                    double Fv[4];
                    viscous_flux(o, idx, i, false, Fv);
                    for (int q = 0; q < 4; ++q) F[q] = F[q] - Fv[q];
                }
            }
    });
    // RK update of fluid cells (Run.cpp:193-200)
    const int sub = rk_substage(o.cfg.rk, counter);
    parallel_for(nx, o.nthreads, [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i)
            for (int j = 0; j < ny; ++j) {
                if (o.type[(size_t)i * ny + j] != CFS_CELL_FLUID) continue;
                // Cell::comRHS (Cell.cpp:110-125); interfaces L,R,B,T
                const double* fl = &o.FX[((size_t)i * ny + j) * 4];
                const double* fr = &o.FX[((size_t)(i + 1) * ny + j) * 4];
                const double* gb = &o.FY[((size_t)i * (ny + 1) + j) * 4];
                const double* gt = &o.FY[((size_t)i * (ny + 1) + j + 1) * 4];
                for (int q = 0; q < 4; ++q) {
                    const double rhs = -(fr[q] - fl[q]) / o.dx - (gt[q] - gb[q]) / o.dy;
                    const size_t a = o.at(q, i, j);
                    o.RHS[a] = rhs;
                    rk_update(o.cfg.rk, sub, o.U[a], o.snap0[a], o.snap1[a], rhs, dt);
                }
            }
    });
    apply_bc(o);   // Run.cpp:203
}

// ---------------------------------------------------------------------------
// Multi-rank restatement of the reference's MPI CPU path
// (source/parallel/CartesianParallel.cpp): the grid is cut into dims[0] x
// dims[1] blocks (ZoneDcomposition, :3-47, with MPI_Dims_create's balanced
// factorisation; a remainder of cells is spread over the first blocks, where
// the reference's N / dims drops it), one worker thread per block (a rank).
// Each stage every rank
//   1. fills the immersed ghosts of its block (Run.cpp:156-162),
//   2. copies its block and the neighbours' edge strips into a frame of
//      3-cell-deep ghost cells (InsertCpuBoundaryGhostCells, :60-170, keeps
//      3 per process boundary; exchangedata, :172-427, refreshes them every
//      stage) -- all ranks wait for every copy before any cell changes,
//   3. computes the faces of its block from the frame, the process-boundary
//      faces on both sides (each rank walks its own ghost chain, as the
//      reference does), and updates its cells.
// The physical-boundary fill follows.  Every face sees the same stencil
// values as in stage(), so the result is bit-identical to it for every
// decomposition.

class Barrier {
public:
    explicit Barrier(int n) : n_(n) {}
    void wait() {
        std::unique_lock<std::mutex> lk(m_);
        const int gen = gen_;
        if (++count_ == n_) {
            count_ = 0;
            ++gen_;
            cv_.notify_all();
        } else {
            cv_.wait(lk, [&] { return gen != gen_; });
        }
    }

private:
    std::mutex m_;
    std::condition_variable cv_;
    int n_, count_ = 0, gen_ = 0;
};

// MPI_Dims_create(n, 2): the two factors of n closest to each other, larger first
void dims_create(int n, int dims[2]) {
    int a = 1;
    for (int f = 1; f * f <= n; ++f)
        if (n % f == 0) a = f;
    dims[0] = n / a;
    dims[1] = a;
}

void stage_ranks(Oracle& o, int counter, double dt) {
    const int nx = o.nx, ny = o.ny;
    const int d0 = o.dims[0], d1 = o.dims[1];
    const int sub = rk_substage(o.cfg.rk, counter);
    Barrier bar(d0 * d1);
    auto rank_main = [&](int cx, int cy) {
        const int bx0 = (int)((long long)cx * nx / d0), bx1 = (int)((long long)(cx + 1) * nx / d0);
        const int by0 = (int)((long long)cy * ny / d1), by1 = (int)((long long)(cy + 1) * ny / d1);
        const int bnx = bx1 - bx0, bny = by1 - by0;
        // 1. immersed ghosts of the block
        for (const ImmersedGhost& gc : o.ig) {
            const int i = gc.cell / ny, j = gc.cell % ny;
            if (i >= bx0 && i < bx1 && j >= by0 && j < by1) ib_fill_one(o, gc);
        }
        bar.wait();
        // 2. exchange: block + 3-deep frame (periodic axes wrap; past a
        // physical edge the frame is unused: those reads go to the boundary ghosts)
        const int fx = bnx + 6, fy = bny + 6;
        std::vector<double> frame(4 * (size_t)fx * fy, 0.0);
        for (int a = 0; a < fx; ++a)
            for (int b = 0; b < fy; ++b) {
                int i = bx0 - 3 + a, j = by0 - 3 + b;
                if (o.periodic_x()) i = (i % nx + nx) % nx;
                if (o.periodic_y()) j = (j % ny + ny) % ny;
                if (i < 0 || i >= nx || j < 0 || j >= ny) continue;
                for (int q = 0; q < 4; ++q) frame[(size_t)q * fx * fy + (size_t)a * fy + b] = o.U[o.at(q, i, j)];
            }
        bar.wait();   // every frame is filled before any cell of U changes
        const View v{frame.data(), (size_t)fx * fy, fy, bx0 - 3, by0 - 3};
        // the frame holds wrapped copies: read them at unwrapped indices
        tl_view = &v;
        // 3. faces of the block (x: f = bx0..bx1, y: g = by0..by1), then the update
        std::vector<double> FX((size_t)(bnx + 1) * bny * 4), FY((size_t)bnx * (bny + 1) * 4);
        int idx[6];
        Phase st[6];
        for (int j = by0; j < by1; ++j)
            for (int f = bx0; f <= bx1; ++f) {
                collect_stencil_rank(f, nx, o.periodic_x(), idx);
                for (int k = 0; k < 6; ++k) st[k] = cell_phase_x(o, idx[k], j);
                double* F = &FX[((size_t)(f - bx0) * bny + (j - by0)) * 4];
                face_flux(o, st, true, F);
                if (o.cfg.positivity == 2) limit_face(o, st[2], st[3], idx[2], j, idx[3], j, true, dt, F);
                if (o.cfg.viscous) {
                    double Fv[4];
                    viscous_flux(o, idx, j, true, Fv);
                    for (int q = 0; q < 4; ++q) F[q] = F[q] - Fv[q];
                }
            }
        for (int i = bx0; i < bx1; ++i)
            for (int gf = by0; gf <= by1; ++gf) {
                collect_stencil_rank(gf, ny, o.periodic_y(), idx);
                for (int k = 0; k < 6; ++k) st[k] = cell_phase_y(o, i, idx[k]);
                double* F = &FY[((size_t)(i - bx0) * (bny + 1) + (gf - by0)) * 4];
                face_flux(o, st, false, F);
                if (o.cfg.positivity == 2) limit_face(o, st[2], st[3], i, idx[2], i, idx[3], false, dt, F);
                if (o.cfg.viscous) {
                    double Fv[4];
                    viscous_flux(o, idx, i, false, Fv);
                    for (int q = 0; q < 4; ++q) F[q] = F[q] - Fv[q];
                }
            }
        tl_view = nullptr;
        for (int i = bx0; i < bx1; ++i)
            for (int j = by0; j < by1; ++j) {
                // global face arrays: each face written by one rank (the one
                // left of / below it, the last rank for the far edge)
                for (int q = 0; q < 4; ++q) {
                    o.FX[((size_t)i * ny + j) * 4 + q] = FX[((size_t)(i - bx0) * bny + (j - by0)) * 4 + q];
                    o.FY[((size_t)i * (ny + 1) + j) * 4 + q] = FY[((size_t)(i - bx0) * (bny + 1) + (j - by0)) * 4 + q];
                    if (i == nx - 1)
                        o.FX[((size_t)nx * ny + j) * 4 + q] = FX[((size_t)bnx * bny + (j - by0)) * 4 + q];
                    if (j == ny - 1)
                        o.FY[((size_t)i * (ny + 1) + ny) * 4 + q] = FY[((size_t)(i - bx0) * (bny + 1) + bny) * 4 + q];
                }
                if (o.type[(size_t)i * ny + j] != CFS_CELL_FLUID) continue;
                const double* fl = &FX[((size_t)(i - bx0) * bny + (j - by0)) * 4];
                const double* fr = &FX[((size_t)(i + 1 - bx0) * bny + (j - by0)) * 4];
                const double* gb = &FY[((size_t)(i - bx0) * (bny + 1) + (j - by0)) * 4];
                const double* gt = &FY[((size_t)(i - bx0) * (bny + 1) + (j + 1 - by0)) * 4];
                for (int q = 0; q < 4; ++q) {
                    const double rhs = -(fr[q] - fl[q]) / o.dx - (gt[q] - gb[q]) / o.dy;
                    const size_t a = o.at(q, i, j);
                    o.RHS[a] = rhs;
                    rk_update(o.cfg.rk, sub, o.U[a], o.snap0[a], o.snap1[a], rhs, dt);
                }
            }
    };
    std::vector<std::thread> th;
    for (int cx = 0; cx < d0; ++cx)
        for (int cy = 0; cy < d1; ++cy) th.emplace_back(rank_main, cx, cy);
    for (auto& t : th) t.join();
    apply_bc(o);   // Run.cpp:203
}

// CFL wave-speed bound (no reference counterpart): max over fluid cells of
// (max(|u|,|rho u|)+c)/dx + (max(|v|,|rho v|)+c)/dy with c = sqrt(gamma p / rho)
// (the momentum terms bound the reference HLL's outer wave speeds).
double cfl_smax(const Oracle& o) {
    double s = 0.0;
    for (int i = 0; i < o.nx; ++i)
        for (int j = 0; j < o.ny; ++j) {
            if (o.type[(size_t)i * o.ny + j] != CFS_CELL_FLUID) continue;
            const double rho = o.U[o.at(0, i, j)], mu = o.U[o.at(1, i, j)];
            const double mv = o.U[o.at(2, i, j)], E = o.U[o.at(3, i, j)];
            const double p = pressure(o.g, rho, mu, mv, E);
            const double c = std::sqrt(o.g * p / rho);
            const double u = mu / rho, v = mv / rho;
            double ax = std::fabs(u), ay = std::fabs(v);
            if (o.cfg.hll_bounds == 0) {
                if (ax < std::fabs(mu)) ax = std::fabs(mu);
                if (ay < std::fabs(mv)) ay = std::fabs(mv);
            }
            const double sx = (ax + c) / o.dx;
            const double sy = (ay + c) / o.dy;
            const double si = sx + sy;
            if (si > s || si != si) s = si;   // a NaN is kept once met
        }
    return s;
}

// Fixed-order checksum tree: the cell sequence (fluid cells; others count
// +0.0) is cut into chunks of 256; each chunk is summed by halving
// (a[t] += a[t+h], h = 128..1); the chunk sums form the next sequence until
// one value remains.
double tree_sum(std::vector<double> a) {
    if (a.empty()) return 0.0;
    while (a.size() > 1) {
        const size_t nchunk = (a.size() + 255) / 256;
        std::vector<double> next(nchunk);
        for (size_t c = 0; c < nchunk; ++c) {
            double buf[256];
            for (int t = 0; t < 256; ++t) {
                const size_t k = c * 256 + t;
                buf[t] = (k < a.size()) ? a[k] : 0.0;
            }
            for (int h = 128; h >= 1; h >>= 1)
                for (int t = 0; t < h; ++t) buf[t] = buf[t] + buf[t + h];
            next[c] = buf[0];
        }
        a.swap(next);
    }
    return a[0];
}

// ---------------------------------------------------------------------------
// Circle set-up: Circle::Initialize/DistinguishWallandFluid/SetGhostcell/
// SetImagepoint/FindNearest4FluidCellsIndexed (Circle.cpp:12-215).
struct Segment { Vec2 a, b, sloop, normal; double length; };

int circle_setup(Oracle& o, double cx, double cy, double r, int npts, int layers) {
    const int nx = o.nx, ny = o.ny;
    std::vector<Vec2> pts(npts);
    for (int i = 0; i < npts; ++i) {
        const double angle = 2.0 * M_PI * (double)(size_t)i / (double)(size_t)npts;
        pts[i] = {cx + r * std::cos(angle), cy + r * std::sin(angle)};
    }
    std::vector<Segment> seg(npts);
    for (int i = 0; i < npts; ++i) {
        seg[i].a = pts[i];
        seg[i].b = pts[(i + 1) % npts];
        const Vec2 v{seg[i].b.x - seg[i].a.x, seg[i].b.y - seg[i].a.y};
        seg[i].length = norm2(v);
        seg[i].sloop = normalized(v);
        seg[i].normal = normalized(Vec2{-v.y, v.x});
    }
    auto centre = [&](int i, int j) {
        return Vec2{(i + 0.5) * o.dx + o.cfg.x0, (j + 0.5) * o.dy + o.cfg.y0};
    };
    o.type.assign((size_t)nx * ny, CFS_CELL_FLUID);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const Vec2 p = centre(i, j);
            const double dist = norm2(Vec2{p.x - cx, p.y - cy});
            if (dist <= r) o.type[(size_t)i * ny + j] = CFS_CELL_SOLID;
        }
    const int dirs[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            if (o.type[(size_t)i * ny + j] != CFS_CELL_FLUID) continue;
            for (auto& d : dirs) {
                const int ni = i + d[0], nj = j + d[1];
                if (!(ni >= 0 && ni < nx && nj >= 0 && nj < ny)) continue;
                if (o.type[(size_t)ni * ny + nj] == CFS_CELL_FLUID) continue;
                for (int k = 1; k <= layers; ++k) {
                    const int gi = i + k * d[0], gj = j + k * d[1];
                    if (gi >= 0 && gi < nx && gj >= 0 && gj < ny && o.type[(size_t)gi * ny + gj] != CFS_CELL_FLUID)
                        o.type[(size_t)gi * ny + gj] = CFS_CELL_IMMERSED_GHOST;
                    else
                        break;
                }
            }
        }
    o.ig.clear();
    const Vec2 origin = centre(0, 0);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            if (o.type[(size_t)i * ny + j] != CFS_CELL_IMMERSED_GHOST) continue;
            const Vec2 p = centre(i, j);
            double best = std::numeric_limits<double>::max();
            Vec2 hit{0, 0};
            for (const Segment& s : seg) {
                const double t = s.sloop.x * (p.x - s.a.x) + s.sloop.y * (p.y - s.a.y);
                const Vec2 proj{t * s.sloop.x + s.a.x, t * s.sloop.y + s.a.y};
                const double dist = norm2(Vec2{proj.x - p.x, proj.y - p.y});
                if (dist < best) { best = dist; hit = proj; }
            }
            const Vec2 n = normalized(Vec2{hit.x - p.x, hit.y - p.y});
            ImmersedGhost gc{};
            gc.cell = i * ny + j;
            gc.ip = {p.x + 2 * (best * n.x), p.y + 2 * (best * n.y)};
            const int ic = (int)((gc.ip.x - origin.x) / o.dx);
            const int jc = (int)((gc.ip.y - origin.y) / o.dy);
            std::vector<std::pair<double, int>> cand;
            for (int a = ic - 5; a <= ic + 5; ++a)
                for (int b = jc - 5; b <= jc + 5; ++b) {
                    if (!(a >= 0 && a < nx && b >= 0 && b < ny)) continue;
                    if (o.type[(size_t)a * ny + b] != CFS_CELL_FLUID) continue;
                    const Vec2 q = centre(a, b);
                    const double ddx = q.x - gc.ip.x, ddy = q.y - gc.ip.y;
                    cand.emplace_back(ddx * ddx + ddy * ddy, a * ny + b);
                }
            std::sort(cand.begin(), cand.end(),
                      [](const std::pair<double, int>& u, const std::pair<double, int>& v) { return u.first < v.first; });
            if (cand.size() < 4) return CFS_E_ARG;
            for (int d = 0; d < 4; ++d) gc.donor[d] = cand[d].second;
            o.ig.push_back(gc);
        }
    return CFS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface for the tests (ctypes).
extern "C" {

struct oracle_solver { Oracle o; };

oracle_solver* oracle_create(const cfs_config* cfg, int nthreads) {
    if (!cfg || cfg->nx < 1 || cfg->ny < 1) return nullptr;
    auto* s = new oracle_solver();
    Oracle& o = s->o;
    o.cfg = *cfg;
    o.nx = cfg->nx; o.ny = cfg->ny;
    o.dx = cfg->lx / cfg->nx; o.dy = cfg->ly / cfg->ny;
    o.g = cfg->gamma;
    const size_t n = (size_t)o.nx * o.ny;
    o.U.assign(4 * n, 0.0);
    o.snap0.assign(4 * n, 0.0);
    o.snap1.assign(4 * n, 0.0);
    o.RHS.assign(4 * n, 0.0);
    o.type.assign(n, CFS_CELL_FLUID);
    o.ghost[0].assign(4 * (size_t)o.ny, 0.0);
    o.ghost[1].assign(4 * (size_t)o.ny, 0.0);
    o.ghost[2].assign(4 * (size_t)o.nx, 0.0);
    o.ghost[3].assign(4 * (size_t)o.nx, 0.0);
    o.FX.assign((size_t)(o.nx + 1) * o.ny * 4, 0.0);
    o.FY.assign((size_t)o.nx * (o.ny + 1) * 4, 0.0);
    o.nthreads = nthreads < 1 ? 1 : nthreads;
    return s;
}

void oracle_destroy(oracle_solver* s) { delete s; }

// Multi-rank mode with n ranks (1: off); dims_out receives the decomposition.
int oracle_set_ranks(oracle_solver* s, int n, int* dims_out) {
    if (!s || n < 1) return CFS_E_ARG;
    Oracle& o = s->o;
    int d[2];
    dims_create(n, d);
    if (d[0] > o.nx || d[1] > o.ny) return CFS_E_ARG;
    o.ranks = n;
    o.dims[0] = d[0];
    o.dims[1] = d[1];
    if (dims_out) { dims_out[0] = d[0]; dims_out[1] = d[1]; }
    return 0;
}

// Initial state: fills all cells, then (like Run::solver's prologue,
// Run.cpp:106-132) the immersed-ghost fill and the boundary fill.
int oracle_set_state(oracle_solver* s, const double* U) {
    Oracle& o = s->o;
    std::memcpy(o.U.data(), U, o.U.size() * sizeof(double));
    ib_fill(o);
    apply_bc(o);
    return 0;
}

int oracle_get_state(oracle_solver* s, double* U) {
    std::memcpy(U, s->o.U.data(), s->o.U.size() * sizeof(double));
    return 0;
}

int oracle_get_rhs(oracle_solver* s, double* R) {
    std::memcpy(R, s->o.RHS.data(), s->o.RHS.size() * sizeof(double));
    return 0;
}

// face fluxes of the last stage: FX [(nx+1)][ny][4], FY [nx][ny+1][4]
int oracle_get_face_flux(oracle_solver* s, double* fx, double* fy) {
    std::memcpy(fx, s->o.FX.data(), s->o.FX.size() * sizeof(double));
    std::memcpy(fy, s->o.FY.data(), s->o.FY.size() * sizeof(double));
    return 0;
}

// the viscous stencil tables: theta_r [3][6], theta_c [5], zeta_r [3][4]
int oracle_viscous_tables(double* thr, double* thc, double* zr) {
    std::memcpy(thr, kThetaR, sizeof(kThetaR));
    std::memcpy(thc, kThetaC, sizeof(kThetaC));
    std::memcpy(zr, kZetaR, sizeof(kZetaR));
    return 0;
}

int oracle_set_circle(oracle_solver* s, double cx, double cy, double r, int npts, int layers) {
    return circle_setup(s->o, cx, cy, r, npts, layers);
}

// Immersed-boundary tables given as data (cell types, ghost cells, donors,
// image points and, for the positivity-preserving interpolation, the donor
// weights); the reference interpolation re-factorises from the donors.
int oracle_set_ib_table(oracle_solver* s, int interp, const uint8_t* cell_type, int n, const int32_t* cells,
                        const int32_t* donors, const double* image_xy, const double* coef) {
    Oracle& o = s->o;
    std::memcpy(o.type.data(), cell_type, o.type.size());
    o.ig.assign(n, ImmersedGhost{});
    for (int k = 0; k < n; ++k) {
        o.ig[k].cell = cells[k];
        o.ig[k].ip = {image_xy[2 * k], image_xy[2 * k + 1]};
        for (int d = 0; d < 4; ++d) {
            o.ig[k].donor[d] = donors[4 * k + d];
            o.ig[k].w[d] = coef[10 * k + d];
        }
    }
    o.ib_interp = interp;
    return 0;
}

int oracle_num_ig(oracle_solver* s) { return (int)s->o.ig.size(); }

int oracle_get_ib(oracle_solver* s, uint8_t* cell_type, int32_t* ig_cells, int32_t* donors, double* image_xy) {
    const Oracle& o = s->o;
    std::memcpy(cell_type, o.type.data(), o.type.size());
    for (size_t k = 0; k < o.ig.size(); ++k) {
        ig_cells[k] = o.ig[k].cell;
        for (int d = 0; d < 4; ++d) donors[4 * k + d] = o.ig[k].donor[d];
        image_xy[2 * k] = o.ig[k].ip.x;
        image_xy[2 * k + 1] = o.ig[k].ip.y;
    }
    return 0;
}

int oracle_stage(oracle_solver* s, int counter, double dt) {
    if (counter != s->o.counter) return CFS_E_STATE;
    if (s->o.ranks > 1) stage_ranks(s->o, counter, dt);
    else stage(s->o, counter, dt);
    s->o.counter++;
    return 0;
}

int oracle_step(oracle_solver* s, double dt) {
    const int n = s->o.cfg.rk == 2 ? 2 : 3;
    for (int k = 0; k < n; ++k) {
        const int rc = oracle_stage(s, s->o.counter, dt);
        if (rc) return rc;
    }
    return 0;
}

double oracle_cfl_smax(oracle_solver* s) { return cfl_smax(s->o); }

int oracle_checksums(oracle_solver* s, double out[4]) {
    const Oracle& o = s->o;
    const size_t n = (size_t)o.nx * o.ny;
    for (int q = 0; q < 4; ++q) {
        std::vector<double> a(n);
        for (size_t c = 0; c < n; ++c) a[c] = (o.type[c] == CFS_CELL_FLUID) ? o.U[(size_t)q * n + c] : 0.0;
        out[q] = tree_sum(a);
    }
    return 0;
}

// scalar kernels, for known-answer tests
double oracle_weno5(const double v[5]) { return weno5(v); }
void oracle_hll(double g, const double L[4], const double R[4], int xdir, double out[4]) {
    Phase a{L[0], L[1], L[2], L[3], 0}, b{R[0], R[1], R[2], R[3], 0};
    a.p = pressure(g, a.rho, a.mu, a.mv, a.E);
    b.p = pressure(g, b.rho, b.mu, b.mv, b.E);
    double FL[4], FR[4];
    euler_flux(g, a, xdir != 0, FL);
    euler_flux(g, b, xdir != 0, FR);
    hll(g, a, b, FL, FR, xdir != 0, out);
}
void oracle_euler_flux(double g, const double S[4], int xdir, double out[4]) {
    Phase a{S[0], S[1], S[2], S[3], 0};
    euler_flux(g, a, xdir != 0, out);
}
void oracle_collect_stencil(int f, int N, int periodic, int out[6]) { collect_stencil(f, N, periodic != 0, out); }
void oracle_ldlt_factor(double* A16, int32_t* transp) {
    double A[4][4];
    for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) A[r][c] = A16[4 * r + c];
    int tr[4];
    ib_ldlt_factor(A, tr);
    for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) A16[4 * r + c] = A[r][c];
    for (int k = 0; k < 4; ++k) transp[k] = tr[k];
}
void oracle_ldlt_solve(const double* A16, const int32_t* transp, double* x) {
    double A[4][4];
    for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) A[r][c] = A16[4 * r + c];
    int tr[4] = {transp[0], transp[1], transp[2], transp[3]};
    ib_ldlt_solve(A, tr, x);
}
// RK combination of one value over consecutive sub-stages with a given rhs
// per sub-stage (closed-form tests of the stage formulas)
double oracle_rk_apply(int rk, double u, const double* rhs, int nsub, double dt) {
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < nsub; ++k) rk_update(rk, rk_substage(rk, k + 1), u, s0, s1, rhs[k], dt);
    return u;
}
double oracle_pressure(double g, const double S[4]) { return pressure(g, S[0], S[1], S[2], S[3]); }
double oracle_energy(double g, double rho, double mu, double mv, double p) { return energy(g, rho, mu, mv, p); }
double oracle_tree_sum(const double* a, int64_t n) { return tree_sum(std::vector<double>(a, a + n)); }

}  // extern "C"
