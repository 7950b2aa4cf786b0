// ib_setup.cpp -- host-side set-up of the circle immersed boundary (run once
// per geometry; the per-stage fill is the ib_fill_kernel on the GPU).
//
// Follows source/object/Circle.cpp and ImagePoint.cpp of the reference:
//   Lagrangian polygon        Circle::Initialize                 (Circle.cpp:12-41)
//   solid classification      Circle::DistinguishWallandFluid    (Circle.cpp:43-51)
//   immersed-ghost layers     Circle::SetGhostcell               (Circle.cpp:53-103)
//   image points              Circle::SetImagepoint              (Circle.cpp:105-138)
//   4 donor fluid cells       Circle::FindNearest4FluidCellsIndexed (Circle.cpp:160-215)
//   donor-matrix factorisation ImagePoint::InterpolatePhase      (ImagePoint.cpp:22-51)
// The reference factorises the 4x4 donor matrix M (rows [x*y, x, y, 1]) with
// Eigen's LDLT every stage although the geometry never changes; LDLT reads the
// lower triangle only, so the system solved is sym(tril(M)).  The factors are
// computed once here, with the same pivoting (largest remaining |diagonal|,
// first on ties) and the same operations, and handed to the device.
//
// Built with -ffp-contract=off: every value here must equal the reference's
// source semantics bit for bit (cos/sin/sqrt are the C library's).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

#include "../../include/cfs.h"

namespace {

struct P2 {
    double x, y;
};

inline double length_of(double x, double y) { return std::sqrt(x * x + y * y); }

// Eigen::MatrixBase::normalized(): v / sqrt(squaredNorm) when positive
inline P2 unit(double x, double y) {
    const double n2 = x * x + y * y;
    if (!(n2 > 0.0)) return {x, y};
    const double n = std::sqrt(n2);
    return {x / n, y / n};
}

class CircleGeometry {
public:
    CircleGeometry(double cx, double cy, double r, int npts) : cx_(cx), cy_(cy), r_(r) {
        start_.resize(npts);
        dir_.resize(npts);
        std::vector<P2> pt(npts);
        for (int k = 0; k < npts; ++k) {
            const double angle = 2.0 * M_PI * (double)(size_t)k / (double)(size_t)npts;
            pt[k] = {cx + r * std::cos(angle), cy + r * std::sin(angle)};
        }
        for (int k = 0; k < npts; ++k) {
            const P2 a = pt[k], b = pt[(k + 1) % npts];
            start_[k] = a;
            dir_[k] = unit(b.x - a.x, b.y - a.y);
        }
    }
    bool solid(P2 c) const { return length_of(c.x - cx_, c.y - cy_) <= r_; }

    // nearest segment line (first strict minimum), mirrored through it
    P2 image_point(P2 g) const {
        double best = std::numeric_limits<double>::max();
        P2 foot{0.0, 0.0};
        for (size_t k = 0; k < start_.size(); ++k) {
            const P2 a = start_[k], s = dir_[k];
            const double along = s.x * (g.x - a.x) + s.y * (g.y - a.y);
            const P2 f{along * s.x + a.x, along * s.y + a.y};
            const double d = length_of(f.x - g.x, f.y - g.y);
            if (d < best) {
                best = d;
                foot = f;
            }
        }
        const P2 n = unit(foot.x - g.x, foot.y - g.y);
        return {g.x + 2 * (best * n.x), g.y + 2 * (best * n.y)};
    }

private:
    double cx_, cy_, r_;
    std::vector<P2> start_, dir_;
};

// Pivoted LDL^T of the lower triangle of a 4x4 matrix, in place:
// a[r][c] for r >= c holds L below the diagonal and D on it afterwards.
void ldlt4(double a[4][4], int32_t piv[4]) {
    for (int k = 0; k < 4; ++k) {
        int p = k;
        for (int i = k + 1; i < 4; ++i)
            if (std::fabs(a[i][i]) > std::fabs(a[p][p])) p = i;
        piv[k] = p;
        if (p != k) {
            for (int c = 0; c < k; ++c) std::swap(a[k][c], a[p][c]);
            for (int r = p + 1; r < 4; ++r) std::swap(a[r][k], a[r][p]);
            std::swap(a[k][k], a[p][p]);
            for (int i = k + 1; i < p; ++i) {
                const double t = a[i][k];
                a[i][k] = a[p][i];
                a[p][i] = t;
            }
        }
        if (k > 0) {
            double w[4];
            for (int i = 0; i < k; ++i) w[i] = a[i][i] * a[k][i];
            double acc = a[k][0] * w[0];
            for (int i = 1; i < k; ++i) acc = acc + a[k][i] * w[i];
            a[k][k] -= acc;
            for (int r = k + 1; r < 4; ++r) {
                double t = a[r][0] * w[0];
                for (int i = 1; i < k; ++i) t = t + a[r][i] * w[i];
                a[r][k] -= t;
            }
        }
        const double d = a[k][k];
        if (d != 0.0)
            for (int r = k + 1; r < 4; ++r) a[r][k] /= d;
    }
}

// Donor weights of the positivity-preserving interpolation: the bilinear fit
// f = a*xi*eta + b*xi + c*eta + d through the 4 donors in coordinates local to
// the image point (xi = (x - x_ip)/dx, eta = (y - y_ip)/dy) has value d at the
// image point, d = g.f with A^T g = e4 (A rows [xi*eta, xi, eta, 1]).  Solved
// by Gaussian elimination with partial pivoting; used when no pivot is below
// 1e-6 (donors not degenerate, e.g. not collinear) and all g >= 0.  Otherwise
// the inverse-squared-distance weights w_k = 1/d_k^2, normalised (a donor at
// distance 0 takes all the weight).
void positive_weights(const P2 donor[4], P2 ip, double dx, double dy, const double d2[4], double g[4]) {
    double a[4][5];
    for (int k = 0; k < 4; ++k) {   // row r of A^T = column r of A
        const double xi = (donor[k].x - ip.x) / dx, eta = (donor[k].y - ip.y) / dy;
        a[0][k] = xi * eta;
        a[1][k] = xi;
        a[2][k] = eta;
        a[3][k] = 1.0;
    }
    a[0][4] = 0.0; a[1][4] = 0.0; a[2][4] = 0.0; a[3][4] = 1.0;
    bool ok = true;
    for (int c = 0; c < 4 && ok; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (!(std::fabs(a[p][c]) >= 1e-6)) { ok = false; break; }
        if (p != c)
            for (int k = 0; k < 5; ++k) std::swap(a[c][k], a[p][k]);
        for (int r = c + 1; r < 4; ++r) {
            const double f = a[r][c] / a[c][c];
            for (int k = c; k < 5; ++k) a[r][k] = a[r][k] - f * a[c][k];
        }
    }
    if (ok) {
        for (int r = 3; r >= 0; --r) {
            double t = a[r][4];
            for (int k = r + 1; k < 4; ++k) t = t - a[r][k] * g[k];
            g[r] = t / a[r][r];
        }
        for (int k = 0; k < 4; ++k)
            if (!(g[k] >= 0.0)) ok = false;
    }
    if (ok) return;
    for (int k = 0; k < 4; ++k) {
        if (d2[k] == 0.0) {
            for (int m = 0; m < 4; ++m) g[m] = (m == k) ? 1.0 : 0.0;
            return;
        }
    }
    double w[4], sum = 0.0;
    for (int k = 0; k < 4; ++k) {
        w[k] = 1.0 / d2[k];
        sum = sum + w[k];
    }
    for (int k = 0; k < 4; ++k) g[k] = w[k] / sum;
}

}  // namespace

extern "C" int32_t cfs_circle_setup(const cfs_config* cfg, double cx, double cy, double radius,
                                    int32_t n_lagrangian, int32_t n_layers, int32_t interp, int32_t cap,
                                    uint8_t* cell_type, int32_t* n_ig, int32_t* ig_cells, int32_t* donors,
                                    double* image_xy, double* coef, int32_t* transp) {
    if (!cfg || !cell_type || !n_ig || cfg->nx < 1 || cfg->ny < 1 || n_lagrangian < 3 || n_layers < 1 ||
        !(radius > 0.0) || cap < 0 || (interp != CFS_IB_REFERENCE && interp != CFS_IB_BILINEAR_POSITIVE))
        return CFS_E_ARG;
    if (cap > 0 && (!ig_cells || !donors || !image_xy || !coef || !transp)) return CFS_E_ARG;
    const int nx = cfg->nx, ny = cfg->ny;
    const double dx = cfg->lx / nx, dy = cfg->ly / ny;
    auto centre = [&](int i, int j) { return P2{(i + 0.5) * dx + cfg->x0, (j + 0.5) * dy + cfg->y0}; };
    auto T = [&](int i, int j) -> uint8_t& { return cell_type[(size_t)i * ny + j]; };

    const CircleGeometry circle(cx, cy, radius, n_lagrangian);
    // Only cells of the body's bounding box, widened by 3 cells, can be solid:
    // a cell centre farther than r + 2 dx from the centre along x or y is at
    // a distance > r whatever the rounding.  Classification, ghost marking and
    // the ghost list scan that box (plus one cell for the marking walks, which
    // start at fluid cells next to the body) instead of the whole grid.
    std::fill(cell_type, cell_type + (size_t)nx * ny, (uint8_t)CFS_CELL_FLUID);
    auto clampi = [](double v, int lo, int hi) {
        if (!(v > lo)) return lo;   // also NaN
        if (!(v < hi)) return hi;
        return (int)v;
    };
    const int i0 = clampi(std::floor((cx - radius - cfg->x0) / dx) - 3, 0, nx - 1);
    const int i1 = clampi(std::ceil((cx + radius - cfg->x0) / dx) + 3, 0, nx - 1);
    const int j0 = clampi(std::floor((cy - radius - cfg->y0) / dy) - 3, 0, ny - 1);
    const int j1 = clampi(std::ceil((cy + radius - cfg->y0) / dy) + 3, 0, ny - 1);
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) T(i, j) = circle.solid(centre(i, j)) ? CFS_CELL_SOLID : CFS_CELL_FLUID;

    // from each fluid cell, walk into the body along the 4 grid directions and
    // mark up to n_layers consecutive non-fluid cells as immersed ghosts
    static const int step[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    const int wi0 = std::max(i0 - 1, 0), wi1 = std::min(i1 + 1, nx - 1);
    const int wj0 = std::max(j0 - 1, 0), wj1 = std::min(j1 + 1, ny - 1);
    for (int i = wi0; i <= wi1; ++i)
        for (int j = wj0; j <= wj1; ++j) {
            if (T(i, j) != CFS_CELL_FLUID) continue;
            for (const auto& d : step) {
                for (int k = 1; k <= n_layers; ++k) {
                    const int a = i + k * d[0], b = j + k * d[1];
                    if (a < 0 || a >= nx || b < 0 || b >= ny || T(a, b) == CFS_CELL_FLUID) break;
                    T(a, b) = CFS_CELL_IMMERSED_GHOST;
                }
            }
        }

    const P2 origin = centre(0, 0);
    int count = 0;
    bool short_window = false;
    std::vector<std::pair<double, int>> cand;
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
            if (T(i, j) != CFS_CELL_IMMERSED_GHOST) continue;
            const int k = count++;
            if (k >= cap) continue;
            const P2 ip = circle.image_point(centre(i, j));
            ig_cells[k] = i * ny + j;
            image_xy[2 * k] = ip.x;
            image_xy[2 * k + 1] = ip.y;
            // candidates: fluid cells in a +-5 window around the cell holding
            // the image point, i outer / j inner, sorted by squared distance
            const int ic = (int)((ip.x - origin.x) / dx);
            const int jc = (int)((ip.y - origin.y) / dy);
            cand.clear();
            for (int a = ic - 5; a <= ic + 5; ++a)
                for (int b = jc - 5; b <= jc + 5; ++b) {
                    if (a < 0 || a >= nx || b < 0 || b >= ny || T(a, b) != CFS_CELL_FLUID) continue;
                    const P2 c = centre(a, b);
                    const double ex = c.x - ip.x, ey = c.y - ip.y;
                    cand.emplace_back(ex * ex + ey * ey, a * ny + b);
                }
            // std::sort with the reference's comparator: ties keep whatever
            // order libstdc++'s introsort leaves, as in the reference
            std::sort(cand.begin(), cand.end(),
                      [](const std::pair<double, int>& u, const std::pair<double, int>& v) { return u.first < v.first; });
            if (cand.size() < 4) {
                short_window = true;
                continue;
            }
            double m[4][4];
            P2 dc[4];
            double d2[4];
            for (int d = 0; d < 4; ++d) {
                donors[4 * k + d] = cand[d].second;
                const P2 c = centre(cand[d].second / ny, cand[d].second % ny);
                dc[d] = c;
                d2[d] = cand[d].first;
                m[d][0] = c.x * c.y;
                m[d][1] = c.x;
                m[d][2] = c.y;
                m[d][3] = 1.0;
            }
            double* l = coef + 10 * k;
            if (interp == CFS_IB_REFERENCE) {
                ldlt4(m, transp + 4 * k);
                l[0] = m[1][0]; l[1] = m[2][0]; l[2] = m[2][1];
                l[3] = m[3][0]; l[4] = m[3][1]; l[5] = m[3][2];
                for (int d = 0; d < 4; ++d) l[6 + d] = m[d][d];
            } else {
                positive_weights(dc, ip, dx, dy, d2, l);
                for (int d = 4; d < 10; ++d) l[d] = 0.0;
                for (int d = 0; d < 4; ++d) transp[4 * k + d] = d;
            }
        }
    *n_ig = count;
    if (count > cap || short_window) return CFS_E_ARG;
    return CFS_OK;
}
