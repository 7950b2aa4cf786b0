// ib_setup.hip -- set-up of the circle immersed boundary, run once per
// geometry (the per-stage fill is done by the stage kernels): on the host
// (cfs_circle_setup) and on the device (cfs_circle_setup_device), with the
// same tables bit for bit.
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
// first on ties) and the same operations, and handed to the stage kernels.
//
// Built with -ffp-contract=off: every value here must equal the reference's
// source semantics bit for bit (cos/sin are the C library's, computed on the
// host for both paths; square roots and divisions are IEEE on both sides).
// The arithmetic the two paths share is written once, __host__ __device__.
//
// The device path: a kernel classifies the cells of the body's bounding box,
// one marks the immersed-ghost layers (a non-fluid cell is a ghost when the
// first fluid cell met walking away from it along a grid direction lies within
// n_layers and inside the region the host walks start from -- the same cells
// the host's walks into the body mark), one workgroup compacts the ghost list
// in the host's order (i outer, j inner), and one thread per ghost computes
// its image point, donors and coefficients.  The reference sorts each ghost's
// donor candidates with std::sort, whose order among equal distances is
// libstdc++'s introsort's own; a ghost whose five nearest candidates include
// two at the same distance is therefore flagged and finished on the host with
// the same std::sort, so the tables never depend on the path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

#include "../../include/cfs.h"
#include "setup_lock.h"

namespace {

struct P2 {
    double x, y;
};

__host__ __device__ inline double length_of(double x, double y) { return sqrt(x * x + y * y); }

// Eigen::MatrixBase::normalized(): v / sqrt(squaredNorm) when positive
__host__ __device__ inline P2 unit(double x, double y) {
    const double n2 = x * x + y * y;
    if (!(n2 > 0.0)) return {x, y};
    const double n = sqrt(n2);
    return {x / n, y / n};
}

// the Lagrangian polygon: segment starts and unit directions
struct Polygon {
    std::vector<P2> start, dir;
    Polygon(double cx, double cy, double r, int npts) : start(npts), dir(npts) {
        std::vector<P2> pt(npts);
        for (int k = 0; k < npts; ++k) {
            const double angle = 2.0 * M_PI * (double)(size_t)k / (double)(size_t)npts;
            pt[k] = {cx + r * std::cos(angle), cy + r * std::sin(angle)};
        }
        for (int k = 0; k < npts; ++k) {
            const P2 a = pt[k], b = pt[(k + 1) % npts];
            start[k] = a;
            dir[k] = unit(b.x - a.x, b.y - a.y);
        }
    }
};

// nearest segment line (first strict minimum), mirrored through it
__host__ __device__ inline P2 image_point(const P2* start, const P2* dir, int n, P2 g) {
    double best = 1.7976931348623157e308;   // std::numeric_limits<double>::max()
    P2 foot{0.0, 0.0};
    for (int k = 0; k < n; ++k) {
        const P2 a = start[k], s = dir[k];
        const double along = s.x * (g.x - a.x) + s.y * (g.y - a.y);
        const P2 f{along * s.x + a.x, along * s.y + a.y};
        const double d = length_of(f.x - g.x, f.y - g.y);
        if (d < best) {
            best = d;
            foot = f;
        }
    }
    const P2 nn = unit(foot.x - g.x, foot.y - g.y);
    return {g.x + 2 * (best * nn.x), g.y + 2 * (best * nn.y)};
}

__host__ __device__ inline void swap_d(double& a, double& b) {
    const double t = a;
    a = b;
    b = t;
}

// Pivoted LDL^T of the lower triangle of a 4x4 matrix, in place:
// a[r][c] for r >= c holds L below the diagonal and D on it afterwards.
__host__ __device__ inline void ldlt4(double a[4][4], int32_t piv[4]) {
    for (int k = 0; k < 4; ++k) {
        int p = k;
        for (int i = k + 1; i < 4; ++i)
            if (fabs(a[i][i]) > fabs(a[p][p])) p = i;
        piv[k] = p;
        if (p != k) {
            for (int c = 0; c < k; ++c) swap_d(a[k][c], a[p][c]);
            for (int r = p + 1; r < 4; ++r) swap_d(a[r][k], a[r][p]);
            swap_d(a[k][k], a[p][p]);
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
__host__ __device__ inline void positive_weights(const P2 donor[4], P2 ip, double dx, double dy, const double d2[4],
                                                 double g[4]) {
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
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        if (!(fabs(a[p][c]) >= 1e-6)) { ok = false; break; }
        if (p != c)
            for (int k = 0; k < 5; ++k) swap_d(a[c][k], a[p][k]);
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

// grid geometry both paths use: cell centres exactly as the host computes them
struct Grid {
    int nx, ny;
    double dx, dy, x0, y0;
    __host__ __device__ P2 centre(int i, int j) const { return P2{(i + 0.5) * dx + x0, (j + 0.5) * dy + y0}; }
};

// the coefficients of one ghost from its 4 donors in distance order
__host__ __device__ inline void donor_tables(const Grid& G, int interp, const int32_t dn[4], const double d2[4],
                                             P2 ip, double* l, int32_t* tr) {
    double m[4][4];
    P2 dc[4];
    for (int d = 0; d < 4; ++d) {
        const P2 c = G.centre(dn[d] / G.ny, dn[d] % G.ny);
        dc[d] = c;
        m[d][0] = c.x * c.y;
        m[d][1] = c.x;
        m[d][2] = c.y;
        m[d][3] = 1.0;
    }
    if (interp == CFS_IB_REFERENCE) {
        ldlt4(m, tr);
        l[0] = m[1][0]; l[1] = m[2][0]; l[2] = m[2][1];
        l[3] = m[3][0]; l[4] = m[3][1]; l[5] = m[3][2];
        for (int d = 0; d < 4; ++d) l[6 + d] = m[d][d];
    } else {
        positive_weights(dc, ip, G.dx, G.dy, d2, l);
        for (int d = 4; d < 10; ++d) l[d] = 0.0;
        for (int d = 0; d < 4; ++d) tr[d] = d;
    }
}

// cell of the image point, the centre of the +-5 candidate window
__host__ __device__ inline void window_centre(const Grid& G, P2 ip, int& ic, int& jc) {
    const P2 origin = G.centre(0, 0);
    ic = (int)((ip.x - origin.x) / G.dx);
    jc = (int)((ip.y - origin.y) / G.dy);
}

// Host: one ghost's tables with the reference's candidate sort; false when
// its window holds fewer than 4 fluid cells.
bool host_ghost(const Grid& G, const uint8_t* T, int interp, P2 ip, int32_t* dn, double* l, int32_t* tr,
                std::vector<std::pair<double, int>>& cand) {
    int ic, jc;
    window_centre(G, ip, ic, jc);
    cand.clear();
    for (int a = ic - 5; a <= ic + 5; ++a)
        for (int b = jc - 5; b <= jc + 5; ++b) {
            if (a < 0 || a >= G.nx || b < 0 || b >= G.ny || T[(size_t)a * G.ny + b] != CFS_CELL_FLUID) continue;
            const P2 c = G.centre(a, b);
            const double ex = c.x - ip.x, ey = c.y - ip.y;
            cand.emplace_back(ex * ex + ey * ey, a * G.ny + b);
        }
    // std::sort with the reference's comparator: ties keep whatever order
    // libstdc++'s introsort leaves, as in the reference
    std::sort(cand.begin(), cand.end(),
              [](const std::pair<double, int>& u, const std::pair<double, int>& v) { return u.first < v.first; });
    if (cand.size() < 4) return false;
    double d2[4];
    for (int d = 0; d < 4; ++d) {
        dn[d] = cand[d].second;
        d2[d] = cand[d].first;
    }
    donor_tables(G, interp, dn, d2, ip, l, tr);
    return true;
}

// bounding box of the body widened by 3 cells: a cell centre farther than
// r + 2 dx from the centre along x or y is at a distance > r whatever the
// rounding, so only the box's cells can be solid
struct Box {
    int i0, i1, j0, j1;     // classification box
    int wi0, wi1, wj0, wj1; // where the marking walks start (the box plus one cell)
};

Box body_box(const cfs_config* cfg, double cx, double cy, double radius, double dx, double dy) {
    const int nx = cfg->nx, ny = cfg->ny;
    auto clampi = [](double v, int lo, int hi) {
        if (!(v > lo)) return lo;   // also NaN
        if (!(v < hi)) return hi;
        return (int)v;
    };
    Box B;
    B.i0 = clampi(std::floor((cx - radius - cfg->x0) / dx) - 3, 0, nx - 1);
    B.i1 = clampi(std::ceil((cx + radius - cfg->x0) / dx) + 3, 0, nx - 1);
    B.j0 = clampi(std::floor((cy - radius - cfg->y0) / dy) - 3, 0, ny - 1);
    B.j1 = clampi(std::ceil((cy + radius - cfg->y0) / dy) + 3, 0, ny - 1);
    B.wi0 = std::max(B.i0 - 1, 0);
    B.wi1 = std::min(B.i1 + 1, nx - 1);
    B.wj0 = std::max(B.j0 - 1, 0);
    B.wj1 = std::min(B.j1 + 1, ny - 1);
    return B;
}

bool setup_args_ok(const cfs_config* cfg, int32_t n_lagrangian, int32_t n_layers, double radius, int32_t interp,
                   int32_t cap, const uint8_t* cell_type, const int32_t* n_ig, const int32_t* ig_cells,
                   const int32_t* donors, const double* image_xy, const double* coef, const int32_t* transp) {
    if (!cfg || !cell_type || !n_ig || cfg->nx < 1 || cfg->ny < 1 || n_lagrangian < 3 || n_layers < 1 ||
        !(radius > 0.0) || cap < 0 || (interp != CFS_IB_REFERENCE && interp != CFS_IB_BILINEAR_POSITIVE))
        return false;
    if (cap > 0 && (!ig_cells || !donors || !image_xy || !coef || !transp)) return false;
    return true;
}

// ---- device kernels -----------------------------------------------------------

__global__ void ib_classify_kernel(uint8_t* T, Grid G, Box B, double cx, double cy, double r) {
    const int bw = B.j1 - B.j0 + 1;
    const int n = (B.i1 - B.i0 + 1) * bw;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int i = B.i0 + q / bw, j = B.j0 + q % bw;
        const P2 c = G.centre(i, j);
        T[(size_t)i * G.ny + j] = length_of(c.x - cx, c.y - cy) <= r ? CFS_CELL_SOLID : CFS_CELL_FLUID;
    }
}

// a non-fluid cell of the box is a ghost when, along one of the 4 grid
// directions, the first fluid cell lies within n_layers and inside the walk
// region: the host's walk from that fluid cell reaches it through non-fluid
// cells only.  Writes change solid to ghost, which other threads' tests
// (fluid or not) do not see.
__global__ void ib_mark_kernel(uint8_t* T, Grid G, Box B, int n_layers) {
    const int bw = B.j1 - B.j0 + 1;
    const int n = (B.i1 - B.i0 + 1) * bw;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int i = B.i0 + q / bw, j = B.j0 + q % bw;
        if (T[(size_t)i * G.ny + j] == CFS_CELL_FLUID) continue;
        bool ghost = false;
        for (int d = 0; d < 4 && !ghost; ++d) {
            const int di = d == 0 ? 1 : d == 1 ? -1 : 0, dj = d == 2 ? 1 : d == 3 ? -1 : 0;
            for (int m = 1; m <= n_layers; ++m) {
                const int a = i - m * di, b = j - m * dj;
                if (a < 0 || a >= G.nx || b < 0 || b >= G.ny) break;
                if (T[(size_t)a * G.ny + b] == CFS_CELL_FLUID) {
                    ghost = a >= B.wi0 && a <= B.wi1 && b >= B.wj0 && b <= B.wj1;
                    break;
                }
            }
        }
        if (ghost) T[(size_t)i * G.ny + j] = CFS_CELL_IMMERSED_GHOST;
    }
}

// one workgroup of 1024 threads: the ghost cells of the box in i-major order
__global__ __launch_bounds__(1024) void ib_compact_kernel(const uint8_t* T, Grid G, Box B, int32_t* list,
                                                          int cap, int32_t* count) {
    __shared__ int scan[1024];
    const int bw = B.j1 - B.j0 + 1;
    const int n = (B.i1 - B.i0 + 1) * bw;
    const int tid = threadIdx.x;
    int base = 0;
    for (int q0 = 0; q0 < n; q0 += 1024) {
        const int q = q0 + tid;
        int cell = -1;
        if (q < n) {
            const int i = B.i0 + q / bw, j = B.j0 + q % bw;
            if (T[(size_t)i * G.ny + j] == CFS_CELL_IMMERSED_GHOST) cell = i * G.ny + j;
        }
        scan[tid] = cell >= 0 ? 1 : 0;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {   // inclusive scan
            const int v = tid >= off ? scan[tid - off] : 0;
            __syncthreads();
            scan[tid] += v;
            __syncthreads();
        }
        const int pos = base + scan[tid] - 1;
        if (cell >= 0 && pos < cap) list[pos] = cell;
        base += scan[1023];
        __syncthreads();   // scan[] is rewritten by the next chunk
    }
    if (tid == 0) *count = base;
}

// one thread per ghost; flag: 0 tables written, 1 distance tie among the five
// nearest candidates (finish on the host), 2 fewer than 4 fluid candidates
__device__ inline void ib_ghost(const uint8_t* T, Grid G, const P2* start, const P2* dir, int npts,
                                const int32_t* list, int n_ig, int interp, double* image_xy, int32_t* donors,
                                double* coef, int32_t* transp, int32_t* flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_ig) return;
    const int cell = list[k];
    const P2 ip = image_point(start, dir, npts, G.centre(cell / G.ny, cell % G.ny));
    image_xy[2 * k] = ip.x;
    image_xy[2 * k + 1] = ip.y;
    int ic, jc;
    window_centre(G, ip, ic, jc);
    // the five nearest fluid candidates, in distance order; candidates are met
    // i outer / j inner as the host lists them, and an equal distance is
    // inserted after the ones already held
    double bd[5];
    int32_t bi[5];
    int nb = 0;
    for (int a = ic - 5; a <= ic + 5; ++a)
        for (int b = jc - 5; b <= jc + 5; ++b) {
            if (a < 0 || a >= G.nx || b < 0 || b >= G.ny || T[(size_t)a * G.ny + b] != CFS_CELL_FLUID) continue;
            const P2 c = G.centre(a, b);
            const double ex = c.x - ip.x, ey = c.y - ip.y;
            const double d = ex * ex + ey * ey;
            int p = nb < 5 ? nb : 5;
            while (p > 0 && d < bd[p - 1]) --p;
            if (p >= 5) continue;
            for (int s = (nb < 5 ? nb : 4); s > p; --s) {
                bd[s] = bd[s - 1];
                bi[s] = bi[s - 1];
            }
            bd[p] = d;
            bi[p] = a * G.ny + b;
            if (nb < 5) ++nb;
        }
    if (nb < 4) {
        flag[k] = 2;
        return;
    }
    // the order of the first four is std::sort's own only without ties among
    // them or with the fifth
    bool tie = false;
    for (int s = 0; s + 1 < nb; ++s) tie = tie || bd[s] == bd[s + 1];
    if (tie) {
        flag[k] = 1;
        return;
    }
    donor_tables(G, interp, bi, bd, ip, coef + 10 * k, transp + 4 * k);
    for (int d = 0; d < 4; ++d) donors[4 * k + d] = bi[d];
    flag[k] = 0;
}

// one device allocation for every buffer of a set-up, freed on return
// ib_ghost_kernel for a count held in device memory
__global__ void ib_ghost_kernel_counted(const uint8_t* T, Grid G, const P2* start, const P2* dir, int npts,
                                        const int32_t* list, const int32_t* count, int interp, double* image_xy,
                                        int32_t* donors, double* coef, int32_t* transp, int32_t* flag) {
    ib_ghost(T, G, start, dir, npts, list, *count, interp, image_xy, donors, coef, transp, flag);
}

struct DevArena {
    char* base = nullptr;
    ~DevArena() {
        if (base) (void)hipFree(base);
    }
};

inline size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

#define IB_HIP(x) HIP_OK(x)

}  // namespace

extern "C" int32_t cfs_circle_setup(const cfs_config* cfg, double cx, double cy, double radius,
                                    int32_t n_lagrangian, int32_t n_layers, int32_t interp, int32_t cap,
                                    uint8_t* cell_type, int32_t* n_ig, int32_t* ig_cells, int32_t* donors,
                                    double* image_xy, double* coef, int32_t* transp) {
    if (!setup_args_ok(cfg, n_lagrangian, n_layers, radius, interp, cap, cell_type, n_ig, ig_cells, donors,
                       image_xy, coef, transp))
        return CFS_E_ARG;
    const int nx = cfg->nx, ny = cfg->ny;
    const Grid G{nx, ny, cfg->lx / nx, cfg->ly / ny, cfg->x0, cfg->y0};
    auto T = [&](int i, int j) -> uint8_t& { return cell_type[(size_t)i * ny + j]; };

    const Polygon poly(cx, cy, radius, n_lagrangian);
    // classification, ghost marking and the ghost list scan the body's box
    // (plus one cell for the marking walks, which start at fluid cells next to
    // the body) instead of the whole grid
    std::fill(cell_type, cell_type + (size_t)nx * ny, (uint8_t)CFS_CELL_FLUID);
    const Box B = body_box(cfg, cx, cy, radius, G.dx, G.dy);
    for (int i = B.i0; i <= B.i1; ++i)
        for (int j = B.j0; j <= B.j1; ++j) {
            const P2 c = G.centre(i, j);
            T(i, j) = length_of(c.x - cx, c.y - cy) <= radius ? CFS_CELL_SOLID : CFS_CELL_FLUID;
        }

    // from each fluid cell, walk into the body along the 4 grid directions and
    // mark up to n_layers consecutive non-fluid cells as immersed ghosts
    static const int step[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    for (int i = B.wi0; i <= B.wi1; ++i)
        for (int j = B.wj0; j <= B.wj1; ++j) {
            if (T(i, j) != CFS_CELL_FLUID) continue;
            for (const auto& d : step) {
                for (int k = 1; k <= n_layers; ++k) {
                    const int a = i + k * d[0], b = j + k * d[1];
                    if (a < 0 || a >= nx || b < 0 || b >= ny || T(a, b) == CFS_CELL_FLUID) break;
                    T(a, b) = CFS_CELL_IMMERSED_GHOST;
                }
            }
        }

    int count = 0;
    bool short_window = false;
    std::vector<std::pair<double, int>> cand;
    for (int i = B.i0; i <= B.i1; ++i)
        for (int j = B.j0; j <= B.j1; ++j) {
            if (T(i, j) != CFS_CELL_IMMERSED_GHOST) continue;
            const int k = count++;
            if (k >= cap) continue;
            const P2 ip = image_point(poly.start.data(), poly.dir.data(), n_lagrangian, G.centre(i, j));
            ig_cells[k] = i * ny + j;
            image_xy[2 * k] = ip.x;
            image_xy[2 * k + 1] = ip.y;
            if (!host_ghost(G, cell_type, interp, ip, donors + 4 * k, coef + 10 * k, transp + 4 * k, cand))
                short_window = true;
        }
    *n_ig = count;
    if (count > cap || short_window) return CFS_E_ARG;
    return CFS_OK;
}

extern "C" int32_t cfs_circle_setup_device(const cfs_config* cfg, double cx, double cy, double radius,
                                           int32_t n_lagrangian, int32_t n_layers, int32_t interp, int32_t cap,
                                           uint8_t* cell_type, int32_t* n_ig, int32_t* ig_cells, int32_t* donors,
                                           double* image_xy, double* coef, int32_t* transp, int32_t* n_host) {
    if (!setup_args_ok(cfg, n_lagrangian, n_layers, radius, interp, cap, cell_type, n_ig, ig_cells, donors,
                       image_xy, coef, transp))
        return CFS_E_ARG;
    const int nx = cfg->nx, ny = cfg->ny;
    const size_t cells = (size_t)nx * ny;
    const Grid G{nx, ny, cfg->lx / nx, cfg->ly / ny, cfg->x0, cfg->y0};
    const Polygon poly(cx, cy, radius, n_lagrangian);
    const Box B = body_box(cfg, cx, cy, radius, G.dx, G.dy);
    const int box_cells = (B.i1 - B.i0 + 1) * (B.j1 - B.j0 + 1);
    // ghosts lie in the box: box_cells bounds their number, so every buffer
    // is sized before the ghosts are counted and one allocation serves all
    const size_t nk = (size_t)box_cells;
    size_t o = 0;
    const size_t oT = o;      o += up256(cells);
    const size_t oPoly = o;   o += up256(sizeof(P2) * 2 * (size_t)n_lagrangian);
    const size_t oList = o;   o += up256(sizeof(int32_t) * nk);
    const size_t oCount = o;  o += up256(sizeof(int32_t));
    const size_t oXy = o;     o += up256(sizeof(double) * 2 * nk);
    const size_t oCoef = o;   o += up256(sizeof(double) * 10 * nk);
    const size_t oDon = o;    o += up256(sizeof(int32_t) * 4 * nk);
    const size_t oTr = o;     o += up256(sizeof(int32_t) * 4 * nk);
    const size_t oFlag = o;   o += up256(sizeof(int32_t) * nk);
    // the set-up's allocation, null-stream work and free, made one at a time
    // with the library's other set-up calls and stream captures
    cfs::SetupLock lk(cfs::hip_setup_mutex());
    DevArena mem;
    IB_HIP(hipMalloc(&mem.base, o));
    uint8_t* dT = reinterpret_cast<uint8_t*>(mem.base + oT);
    P2* dPoly = reinterpret_cast<P2*>(mem.base + oPoly);
    int32_t* dList = reinterpret_cast<int32_t*>(mem.base + oList);
    int32_t* dCount = reinterpret_cast<int32_t*>(mem.base + oCount);
    double* dXy = reinterpret_cast<double*>(mem.base + oXy);
    double* dCoef = reinterpret_cast<double*>(mem.base + oCoef);
    int32_t* dDon = reinterpret_cast<int32_t*>(mem.base + oDon);
    int32_t* dTr = reinterpret_cast<int32_t*>(mem.base + oTr);
    int32_t* dFlag = reinterpret_cast<int32_t*>(mem.base + oFlag);
    IB_HIP(hipMemsetAsync(dT, CFS_CELL_FLUID, cells, 0));
    IB_HIP(hipMemcpyAsync(dPoly, poly.start.data(), sizeof(P2) * n_lagrangian, hipMemcpyHostToDevice, 0));
    IB_HIP(hipMemcpyAsync(dPoly + n_lagrangian, poly.dir.data(), sizeof(P2) * n_lagrangian, hipMemcpyHostToDevice,
                          0));
    const int blocks = std::min((box_cells + 255) / 256, 1024);
    hipLaunchKernelGGL(ib_classify_kernel, dim3(blocks), dim3(256), 0, 0, dT, G, B, cx, cy, radius);
    IB_HIP(hipGetLastError());
    hipLaunchKernelGGL(ib_mark_kernel, dim3(blocks), dim3(256), 0, 0, dT, G, B, (int)n_layers);
    IB_HIP(hipGetLastError());
    hipLaunchKernelGGL(ib_compact_kernel, dim3(1), dim3(1024), 0, 0, dT, G, B, dList, box_cells, dCount);
    IB_HIP(hipGetLastError());
    // every ghost's tables (box_cells threads; those past the count return at once)
    hipLaunchKernelGGL(ib_ghost_kernel_counted, dim3((box_cells + 127) / 128), dim3(128), 0, 0, dT, G, dPoly,
                       dPoly + n_lagrangian, (int)n_lagrangian, dList, dCount, (int)interp, dXy, dDon, dCoef, dTr,
                       dFlag);
    IB_HIP(hipGetLastError());
    int32_t count = 0;
    IB_HIP(hipMemcpy(&count, dCount, sizeof(int32_t), hipMemcpyDeviceToHost));
    const int k_n = std::min(count, cap);
    IB_HIP(hipMemcpy(cell_type, dT, cells, hipMemcpyDeviceToHost));
    *n_ig = count;
    if (n_host) *n_host = 0;
    if (count > cap) return CFS_E_ARG;
    if (k_n == 0) return CFS_OK;
    std::vector<int32_t> flag((size_t)k_n);
    IB_HIP(hipMemcpy(ig_cells, dList, sizeof(int32_t) * k_n, hipMemcpyDeviceToHost));
    IB_HIP(hipMemcpy(image_xy, dXy, sizeof(double) * 2 * k_n, hipMemcpyDeviceToHost));
    IB_HIP(hipMemcpy(coef, dCoef, sizeof(double) * 10 * k_n, hipMemcpyDeviceToHost));
    IB_HIP(hipMemcpy(donors, dDon, sizeof(int32_t) * 4 * k_n, hipMemcpyDeviceToHost));
    IB_HIP(hipMemcpy(transp, dTr, sizeof(int32_t) * 4 * k_n, hipMemcpyDeviceToHost));
    IB_HIP(hipMemcpy(flag.data(), dFlag, sizeof(int32_t) * k_n, hipMemcpyDeviceToHost));
    // ghosts with tied candidates: the host's candidate sort decides
    bool short_window = false;
    int finished_here = 0;
    std::vector<std::pair<double, int>> cand;
    for (int k = 0; k < k_n; ++k) {
        if (flag[k] == 0) continue;
        if (flag[k] == 2) {
            short_window = true;
            continue;
        }
        const P2 ip{image_xy[2 * k], image_xy[2 * k + 1]};
        if (!host_ghost(G, cell_type, interp, ip, donors + 4 * k, coef + 10 * k, transp + 4 * k, cand))
            short_window = true;
        ++finished_here;
    }
    if (n_host) *n_host = finished_here;
    return short_window ? CFS_E_ARG : CFS_OK;
}
