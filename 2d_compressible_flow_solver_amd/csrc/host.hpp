// host.hpp -- host-side pieces shared by the single-solve (cfs_api.hip) and
// the batched (batch.hip) entry points: configuration checks, the stage plan
// of the two TVD-RK schemes, the per-stage kernel arguments of one problem's
// device buffers, and the immersed-boundary index tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/cfs.h"
#include "kernels.hpp"
#include "setup_lock.h"

namespace cfs {

// Sub-stage of the global stage counter i (TVD2rdRK.cpp:11, TVD3rdRK.cpp:15).
inline int rk_substage(int rk, int counter) {
    if (rk == 2) return (counter % 2 == 0) ? 1 : 0;
    return (counter % 3 == 0) ? 2 : (counter % 3 == 1) ? 0 : 1;
}

inline int check_config(const cfs_config* c) {
    if (!c) return CFS_E_ARG;
    if (c->nx < 3 || c->ny < 3) return CFS_E_ARG;
    if ((long long)c->nx * c->ny > (1ll << 31) - 1) return CFS_E_ARG;
    if (!(c->lx > 0) || !(c->ly > 0) || !(c->gamma > 1)) return CFS_E_ARG;
    for (int k = 0; k < 4; ++k)
        if (c->bc[k] < CFS_BC_NEUMANN || c->bc[k] > CFS_BC_PERIODIC) return CFS_E_ARG;
    if ((c->bc[0] == CFS_BC_PERIODIC) != (c->bc[1] == CFS_BC_PERIODIC)) return CFS_E_ARG;
    if ((c->bc[2] == CFS_BC_PERIODIC) != (c->bc[3] == CFS_BC_PERIODIC)) return CFS_E_ARG;
    if (c->rk != 2 && c->rk != 3) return CFS_E_ARG;
    if (c->precision != CFS_PREC_FP64 && c->precision != CFS_PREC_FP32_BULK) return CFS_E_ARG;
    if (c->positivity < 0 || c->positivity > 2) return CFS_E_ARG;
    if (c->hll_bounds != 0 && c->hll_bounds != 1) return CFS_E_ARG;
    if (c->viscous != 0 && c->viscous != 1) return CFS_E_ARG;
    if (c->viscous) {
        if (!(c->mu >= 0) || !(c->mu < 1e300) || !(c->prandtl > 0) || !(c->ma_ref > 0)) return CFS_E_ARG;
        if (!(c->prandtl < 1e300) || !(c->ma_ref < 1e150)) return CFS_E_ARG;
        if (c->positivity == 2) return CFS_E_UNSUPPORTED;   // the limiter covers the inviscid flux only
    }
    return CFS_OK;
}

// Device scalar block of one problem (doubles): [0] dt, [1..2] time by step
// parity, [3..4] CFL bound bits by step parity, [5] CFL bound for a one-off
// dt query, [6] ints: halt, status, [7] ints: steps, unused.
constexpr int SCAL_DOUBLES = 8;

struct ScalPtrs {
    double* p;
    double* dt() const { return p; }
    double* t(int parity) const { return p + 1 + parity; }
    unsigned long long* acc(int parity) const { return reinterpret_cast<unsigned long long*>(p + 3 + parity); }
    unsigned long long* acc_tmp() const { return reinterpret_cast<unsigned long long*>(p + 5); }
    int* halt() const { return reinterpret_cast<int*>(p + 6); }
    int* status() const { return reinterpret_cast<int*>(p + 6) + 1; }
    int* steps() const { return reinterpret_cast<int*>(p + 7); }
};

// The device buffers of one problem, wherever they were allocated.
struct ProblemDev {
    const cfs_config* cfg = nullptr;
    long long plane = 0;          // nx*ny
    double* planes = nullptr;     // 3 plane-sets of 4*plane doubles
    int tiles_i = 0, tiles_j = 0;
    uint8_t* type = nullptr;      // nullptr: every cell is fluid
    int n_ig = 0;
    int ib_interp = CFS_IB_REFERENCE;
    int* ig_cells = nullptr;
    int* ig_donors = nullptr;
    double* ig_xy = nullptr;
    double* ig_ldl = nullptr;
    int* ig_transp = nullptr;
    int* ig_index = nullptr;
    uint8_t* tile_ib = nullptr;
    double* scal = nullptr;       // SCAL_DOUBLES
    int* exact = nullptr;         // 2 counters by stage parity

    double* set(int k) const { return planes + (size_t)k * 4 * plane; }
    ScalPtrs s() const { return ScalPtrs{scal}; }
};

// Last-level cache of the device (MI355X Infinity Cache, 256 MiB): state
// plane-sets beyond it are streamed with non-temporal accesses in the RK
// update (StageArgs::stream_nt).
constexpr long long LLC_BYTES = 256ll << 20;

// How a stage obtains dt: fixed (dt scalar set by the host), or as part of a
// CFL step (first stage derives dt from the bound of the entering state, the
// last stage accumulates the bound of the leaving state).
struct StageTime {
    int dt_mode = 0;
    bool fuse_cfl = false;
    double cfl = 0.0, t_end = 0.0;
};

// Kernel arguments of sub-stage `sub`.  Plane-sets: 0 holds U^n at a step
// boundary; RK3: 0->1, 1(+0)->2, 2(+0,1)->0; RK2: 0->1, 1(+0)->0.  Returns the
// plane-set the stage writes.
inline int fill_stage_args(StageArgs& a, const ProblemDev& p, int sub, int step_parity, int stage_parity,
                           const StageTime& tm) {
    const cfs_config& c = *p.cfg;
    a = StageArgs{};
    double* A = p.set(0);
    double* B = p.set(1);
    double* Cp = p.set(2);
    a.type = p.type;
    a.plane = p.plane;
    a.nx = c.nx;
    a.ny = c.ny;
    a.tiles_j = p.tiles_j;
    a.ntiles = p.tiles_i * p.tiles_j;
    a.dx = c.lx / c.nx;
    a.dy = c.ly / c.ny;
    a.inv_dx = 1.0 / a.dx;
    a.inv_dy = 1.0 / a.dy;
    int e;
    a.stream_nt = 3ll * 32 * p.plane > LLC_BYTES ? 1 : 0;   // three plane-sets of 32 B per cell
    a.dx_pow2 = std::frexp(a.dx, &e) == 0.5;   // then 1/dx is exact and x/dx == x*(1/dx)
    a.dy_pow2 = std::frexp(a.dy, &e) == 0.5;
    a.gamma = c.gamma;
    a.positivity = c.positivity;
    a.hll_bounds = c.hll_bounds;
    if (c.viscous) {
        a.mu = c.mu;
        a.ma2 = c.ma_ref * c.ma_ref;
        a.kappa = c.mu / ((c.gamma - 1) * (c.ma_ref * c.ma_ref) * c.prandtl);
    }
    a.exact_count = p.exact + stage_parity;
    a.exact_other = p.exact + (1 - stage_parity);
    for (int k = 0; k < 4; ++k) {
        a.bc[k] = c.bc[k];
        for (int q = 0; q < 4; ++q) a.bcs[k][q] = c.bc_state[k][q];
    }
    if (p.n_ig > 0) {
        a.tile_ib = p.tile_ib;
        a.ig_index = p.ig_index;
        a.ig_donors = p.ig_donors;
        a.ig_coef = p.ig_ldl;
        a.ig_transp = p.ig_transp;
        a.ig_xy = p.ig_xy;
        a.ib_interp = p.ib_interp;
    }
    const ScalPtrs s = p.s();
    a.dt_mode = tm.dt_mode;
    a.dt = s.dt();
    a.acc_in = s.acc(step_parity);
    a.acc_out = s.acc(1 - step_parity);
    a.fuse_cfl = tm.fuse_cfl ? 1 : 0;
    a.t_in = s.t(step_parity);
    a.t_out = s.t(1 - step_parity);
    a.halt = s.halt();
    a.status = s.status();
    a.steps = s.steps();
    a.cfl = tm.cfl;
    a.t_end = tm.t_end;
    if (c.rk == 3) {
        if (sub == 0) { a.in = A; a.out = B; a.form = 0; a.c0 = 0.0; a.c1 = 1.0; a.c2 = 1.0; return 1; }
        if (sub == 1) { a.in = B; a.u0 = A; a.out = Cp; a.form = 1; a.c0 = 0.75; a.c1 = 0.25; a.c2 = 0.25; return 2; }
        a.in = Cp; a.u0 = A; a.u1 = B; a.out = A; a.form = 2; a.c0 = 1.0 / 3.0; a.c1 = 2.0 / 3.0; a.c2 = 2.0 / 3.0;
        return 0;
    }
    if (sub == 0) { a.in = A; a.out = B; a.form = 0; a.c0 = 0.0; a.c1 = 1.0; a.c2 = 1.0; return 1; }
    a.in = B; a.u0 = A; a.out = A; a.form = 1; a.c0 = 0.5; a.c1 = 0.5; a.c2 = 0.5;
    return 0;
}

// The CFL-step time mode of sub-stage k of an rk-stage step.
inline StageTime cfl_stage_time(int rk, int k, double cfl, double t_end) {
    StageTime tm;
    tm.dt_mode = (k == 0) ? 1 : 2;
    tm.fuse_cfl = (k == rk - 1);
    tm.cfl = cfl;
    tm.t_end = t_end;
    return tm;
}

// Tile shapes of the stage kernels (kernels.hpp), by index.
struct ShapeDims {
    int ti, tj, nt;
};
inline ShapeDims shape_dims(int k) {
    if (k == 1) return ShapeDims{Shape1::TI, Shape1::TJ, Shape1::NT};
    if (k == 2) return ShapeDims{Shape2::TI, Shape2::TJ, Shape2::NT};
    if (k == 3) return ShapeDims{Shape3::TI, Shape3::TJ, Shape3::NT};
    return ShapeDims{Shape0::TI, Shape0::TJ, Shape0::NT};
}
inline long long tiles_for(const cfs_config& c, int k) {
    const ShapeDims d = shape_dims(k);
    return (long long)((c.nx + d.ti - 1) / d.ti) * ((c.ny + d.tj - 1) / d.tj);
}

// Per-tile flag of tile shape k: "the staged region (tile + 3-cell halo,
// periodic wrap) holds an immersed ghost".
inline void ib_tile_flags(const cfs_config& c, int k, int n_ig, const int32_t* ig_cells,
                          std::vector<uint8_t>& tiles) {
    const ShapeDims d = shape_dims(k);
    const int nx = c.nx, ny = c.ny;
    const int tiles_j = (ny + d.tj - 1) / d.tj;
    tiles.assign((size_t)tiles_for(c, k), 0);
    for (int n = 0; n < n_ig; ++n) {
        const int gi = ig_cells[n] / ny, gj = ig_cells[n] % ny;
        for (int di = -HALO; di <= HALO; ++di)
            for (int dj = -HALO; dj <= HALO; ++dj) {
                int a = gi + di, b = gj + dj;
                if (c.bc[0] == CFS_BC_PERIODIC) a = (a % nx + nx) % nx;
                if (c.bc[2] == CFS_BC_PERIODIC) b = (b % ny + ny) % ny;
                if (a < 0 || a >= nx || b < 0 || b >= ny) continue;
                tiles[(size_t)(a / d.ti) * tiles_j + b / d.tj] = 1;
            }
    }
}

// Checks immersed-boundary tables against the cell types and builds the
// per-cell row index (-1 off the ghosts) and the per-tile flags of tile shape
// `shape` (ib_tile_flags).
inline int build_ib_index(const cfs_config& c, int shape, int interp, const uint8_t* cell_type,
                          int n_ig, const int32_t* ig_cells, const int32_t* donors, const int32_t* transp,
                          std::vector<int>& index, std::vector<uint8_t>& tiles) {
    if (!cell_type || n_ig < 0) return CFS_E_ARG;
    if (interp != CFS_IB_REFERENCE && interp != CFS_IB_BILINEAR_POSITIVE) return CFS_E_ARG;
    if (n_ig > 0 && (!ig_cells || !donors || !transp)) return CFS_E_ARG;
    const int nx = c.nx, ny = c.ny;
    const long long plane = (long long)nx * ny;
    index.assign((size_t)plane, -1);
    for (long long k = 0; k < plane; ++k)
        if (cell_type[k] > CFS_CELL_IMMERSED_GHOST) return CFS_E_ARG;
    for (int k = 0; k < n_ig; ++k) {
        if (ig_cells[k] < 0 || ig_cells[k] >= plane) return CFS_E_ARG;
        if (cell_type[ig_cells[k]] != CFS_CELL_IMMERSED_GHOST || index[ig_cells[k]] >= 0) return CFS_E_ARG;
        index[ig_cells[k]] = k;
        for (int d = 0; d < 4; ++d) {
            const int cc = donors[4 * k + d];
            if (cc < 0 || cc >= plane || cell_type[cc] != CFS_CELL_FLUID) return CFS_E_ARG;
            const int t = transp[4 * k + d];
            if (interp == CFS_IB_REFERENCE && (t < d || t > 3)) return CFS_E_ARG;
        }
    }
    for (long long k = 0; k < plane; ++k)
        if (cell_type[k] == CFS_CELL_IMMERSED_GHOST && index[k] < 0) return CFS_E_ARG;   // ghost without a row
    ib_tile_flags(c, shape, n_ig, ig_cells, tiles);
    return CFS_OK;
}

// Launch helpers on one problem's buffers.
inline int launch_ib_fill(const ProblemDev& p, double* U, hipStream_t stream) {
    if (p.n_ig == 0) return CFS_OK;
    IBArgs a{};
    a.U = U;
    a.plane = p.plane;
    a.n = p.n_ig;
    a.interp = p.ib_interp;
    a.cells = p.ig_cells;
    a.donors = p.ig_donors;
    a.ipxy = p.ig_xy;
    a.ldl = p.ig_ldl;
    a.transp = p.ig_transp;
    hipLaunchKernelGGL(ib_fill_kernel, dim3((p.n_ig + 255) / 256), dim3(256), 0, stream, a);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

// The CFL bound of plane-set `k` into acc (cleared first, by a kernel).
inline int launch_cfl(const ProblemDev& p, int k, unsigned long long* acc, hipStream_t stream) {
    const cfs_config& c = *p.cfg;
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, stream, acc, 0ull);
    hipLaunchKernelGGL(cfl_kernel, dim3(2048), dim3(256), 0, stream, p.set(k), p.plane, p.type, c.gamma,
                       c.lx / c.nx, c.ly / c.ny, c.hll_bounds, c.precision == CFS_PREC_FP32_BULK ? 1 : 0, acc);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

}  // namespace cfs
