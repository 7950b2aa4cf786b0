// cfs_api.hip -- the C-ABI of include/cfs.h on top of the gfx950 kernels.
//
// A cfs_solver owns: three FP64 state plane-sets (RK ping-pong), an optional
// cell-type plane and immersed-boundary tables (plus a per-cell ghost index
// and a per-tile "near the body" flag for the fused fill), a block of device
// scalars (dt, time, CFL bounds, halt/status/step counters), scratch for the
// checksum tree, one HIP stream and a cached HIP graph of two RK steps.
// Nothing here falls back to a CPU path: every compute entry point launches
// HIP kernels and fails with CFS_E_HIP when HIP does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cfs.h"
#include "kernels.hpp"

using namespace cfs;

namespace {

// device scalar block (doubles): [0] dt, [1..2] time by step parity,
// [3..4] CFL bound bits by step parity, [5] CFL bound for cfs_compute_dt,
// [6] ints: halt, status, [7] ints: steps, unused
constexpr int SCAL_DOUBLES = 8;

struct GraphKey {
    bool valid = false;
    bool use_cfl = false;
    double cfl = 0.0, t_end = 0.0;
    int step_parity = 0, stage_parity = 0;
};

}  // namespace

struct cfs_solver {
    cfs_config cfg{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    long long plane = 0;          // nx*ny
    double* planes = nullptr;     // 3 plane-sets of 4*plane doubles
    int cur = 0;                  // plane-set holding the current state
    uint8_t* type = nullptr;      // cell types (nullptr: all fluid)
    // immersed boundary
    int n_ig = 0;
    int ib_interp = CFS_IB_REFERENCE;
    int* ig_cells = nullptr;
    int* ig_donors = nullptr;
    double* ig_xy = nullptr;
    double* ig_ldl = nullptr;
    int* ig_transp = nullptr;
    int* ig_index = nullptr;      // per cell: row of the IB tables, -1 elsewhere
    uint8_t* tile_ib = nullptr;   // per tile: load region holds immersed ghosts
    double* scal = nullptr;       // device scalar block
    double* tree = nullptr;       // checksum scratch
    long long tree_len = 0;
    int* exact = nullptr;         // per-stage-parity counters of tiles outside the x/3, x/6 exact range
    int parity = 0;               // stage parity (exact counters)
    int step_parity = 0;          // step parity (time and CFL-bound slots)
    int counter = 1;
    int tiles_i = 0, tiles_j = 0;
    bool kernel_timing = false;   // bracket stage kernels with HIP events in cfs_advance
    double last_kernel_ms = 0.0;
    std::vector<hipEvent_t> ev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    GraphKey graph_key;
};

namespace {

#define HIP_OK(x)                                  \
    do {                                           \
        if ((x) != hipSuccess) return CFS_E_HIP;   \
    } while (0)

int rk_substage(int rk, int counter) {
    if (rk == 2) return (counter % 2 == 0) ? 1 : 0;
    return (counter % 3 == 0) ? 2 : (counter % 3 == 1) ? 0 : 1;
}

double* plane_set(cfs_solver* s, int k) { return s->planes + (size_t)k * 4 * s->plane; }
double* dt_ptr(cfs_solver* s) { return s->scal; }
double* t_ptr(cfs_solver* s, int p) { return s->scal + 1 + p; }
unsigned long long* acc_ptr(cfs_solver* s, int p) { return reinterpret_cast<unsigned long long*>(s->scal + 3 + p); }
unsigned long long* acc_tmp_ptr(cfs_solver* s) { return reinterpret_cast<unsigned long long*>(s->scal + 5); }
int* halt_ptr(cfs_solver* s) { return reinterpret_cast<int*>(s->scal + 6); }
int* status_ptr(cfs_solver* s) { return reinterpret_cast<int*>(s->scal + 6) + 1; }
int* steps_ptr(cfs_solver* s) { return reinterpret_cast<int*>(s->scal + 7); }

int check_config(const cfs_config* c) {
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
    if (c->positivity != 0 && c->positivity != 1) return CFS_E_ARG;
    if (c->hll_bounds != 0 && c->hll_bounds != 1) return CFS_E_ARG;
    if (c->viscous != 0 && c->viscous != 1) return CFS_E_ARG;
    if (c->viscous) {
        if (!(c->mu >= 0) || !(c->mu < 1e300) || !(c->prandtl > 0) || !(c->ma_ref > 0)) return CFS_E_ARG;
        if (!(c->prandtl < 1e300) || !(c->ma_ref < 1e150)) return CFS_E_ARG;
    }
    return CFS_OK;
}

void drop_graph(cfs_solver* s) {
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    s->graph_key.valid = false;
}

int launch_ib_fill(cfs_solver* s, double* U) {
    if (s->n_ig == 0) return CFS_OK;
    IBArgs a{};
    a.U = U;
    a.plane = s->plane;
    a.n = s->n_ig;
    a.interp = s->ib_interp;
    a.cells = s->ig_cells;
    a.donors = s->ig_donors;
    a.ipxy = s->ig_xy;
    a.ldl = s->ig_ldl;
    a.transp = s->ig_transp;
    hipLaunchKernelGGL(ib_fill_kernel, dim3((s->n_ig + 255) / 256), dim3(256), 0, s->stream, a);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

// How a stage obtains dt: fixed (dt scalar set by the host), or as part of a
// CFL step (first stage derives dt from the bound of the entering state, the
// last stage accumulates the bound of the leaving state).
struct StageTime {
    int dt_mode = 0;
    bool fuse_cfl = false;
    double cfl = 0.0, t_end = 0.0;
};

// One RK stage.  Plane-sets: A = 0 holds U^n at a step boundary;
// RK3: A->B, B(+A)->C, C(+A,B)->A; RK2: A->B, B(+A)->A.
int launch_stage(cfs_solver* s, int counter, bool timed, const StageTime& tm) {
    const cfs_config& c = s->cfg;
    const int sub = rk_substage(c.rk, counter);
    StageArgs a{};
    double* A = plane_set(s, 0);
    double* B = plane_set(s, 1);
    double* Cp = plane_set(s, 2);
    a.type = s->type;
    a.plane = s->plane;
    a.nx = c.nx;
    a.ny = c.ny;
    a.tiles_j = s->tiles_j;
    a.dx = c.lx / c.nx;
    a.dy = c.ly / c.ny;
    a.inv_dx = 1.0 / a.dx;
    a.inv_dy = 1.0 / a.dy;
    int e;
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
    a.exact_count = s->exact + s->parity;
    a.exact_other = s->exact + (1 - s->parity);
    for (int k = 0; k < 4; ++k) {
        a.bc[k] = c.bc[k];
        for (int q = 0; q < 4; ++q) a.bcs[k][q] = c.bc_state[k][q];
    }
    if (s->n_ig > 0) {
        a.tile_ib = s->tile_ib;
        a.ig_index = s->ig_index;
        a.ig_donors = s->ig_donors;
        a.ig_coef = s->ig_ldl;
        a.ig_transp = s->ig_transp;
        a.ig_xy = s->ig_xy;
        a.ib_interp = s->ib_interp;
    }
    a.dt_mode = tm.dt_mode;
    a.dt = dt_ptr(s);
    a.acc_in = acc_ptr(s, s->step_parity);
    a.acc_out = acc_ptr(s, 1 - s->step_parity);
    a.fuse_cfl = tm.fuse_cfl ? 1 : 0;
    a.t_in = t_ptr(s, s->step_parity);
    a.t_out = t_ptr(s, 1 - s->step_parity);
    a.halt = halt_ptr(s);
    a.status = status_ptr(s);
    a.steps = steps_ptr(s);
    a.cfl = tm.cfl;
    a.t_end = tm.t_end;
    int next;
    if (c.rk == 3) {
        if (sub == 0) { a.in = A; a.out = B; a.form = 0; a.c0 = 0.0; a.c1 = 1.0; a.c2 = 1.0; next = 1; }
        else if (sub == 1) { a.in = B; a.u0 = A; a.out = Cp; a.form = 1; a.c0 = 0.75; a.c1 = 0.25; a.c2 = 0.25; next = 2; }
        else { a.in = Cp; a.u0 = A; a.u1 = B; a.out = A; a.form = 2; a.c0 = 1.0 / 3.0; a.c1 = 2.0 / 3.0; a.c2 = 2.0 / 3.0; next = 0; }
    } else {
        if (sub == 0) { a.in = A; a.out = B; a.form = 0; a.c0 = 0.0; a.c1 = 1.0; a.c2 = 1.0; next = 1; }
        else { a.in = B; a.u0 = A; a.out = A; a.form = 1; a.c0 = 0.5; a.c1 = 0.5; a.c2 = 0.5; next = 0; }
    }
    const dim3 grid((unsigned)(s->tiles_i * s->tiles_j));
    if (timed) {
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        s->ev.push_back(e0);
        s->ev.push_back(e1);
        HIP_OK(hipEventRecord(e0, s->stream));
    }
    if (c.precision == CFS_PREC_FP64) {
        if (c.viscous) hipLaunchKernelGGL((stage_kernel<double, true>), grid, dim3(BLOCK), 0, s->stream, a);
        else hipLaunchKernelGGL((stage_kernel<double, false>), grid, dim3(BLOCK), 0, s->stream, a);
    } else {
        if (c.viscous) hipLaunchKernelGGL((stage_kernel<float, true>), grid, dim3(BLOCK), 0, s->stream, a);
        else hipLaunchKernelGGL((stage_kernel<float, false>), grid, dim3(BLOCK), 0, s->stream, a);
    }
    HIP_OK(hipGetLastError());
    if (timed) HIP_OK(hipEventRecord(s->ev.back(), s->stream));
    s->parity ^= 1;
    s->cur = next;
    return CFS_OK;
}

// One RK step (all sub-stages) in the given time mode.
int launch_step(cfs_solver* s, bool timed, bool use_cfl, double cfl, double t_end) {
    for (int k = 0; k < s->cfg.rk; ++k) {
        StageTime tm;
        if (use_cfl) {
            tm.dt_mode = (k == 0) ? 1 : 2;
            tm.fuse_cfl = (k == s->cfg.rk - 1);
            tm.cfl = cfl;
            tm.t_end = t_end;
        }
        const int rc = launch_stage(s, s->counter, timed, tm);
        if (rc) return rc;
        s->counter++;
    }
    if (use_cfl) s->step_parity ^= 1;
    return CFS_OK;
}

int set_dt_device(cfs_solver* s, double dt) {
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, dt_ptr(s), dt);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

int free_ib(cfs_solver* s) {
    (void)hipFree(s->ig_cells);
    (void)hipFree(s->ig_donors);
    (void)hipFree(s->ig_xy);
    (void)hipFree(s->ig_ldl);
    (void)hipFree(s->ig_transp);
    (void)hipFree(s->ig_index);
    (void)hipFree(s->tile_ib);
    s->ig_cells = s->ig_donors = s->ig_transp = s->ig_index = nullptr;
    s->ig_xy = s->ig_ldl = nullptr;
    s->tile_ib = nullptr;
    s->n_ig = 0;
    return CFS_OK;
}

// The CFL bound of the current state into acc (cleared first, by a kernel).
int launch_cfl(cfs_solver* s, unsigned long long* acc) {
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream, acc, 0ull);
    hipLaunchKernelGGL(cfl_kernel, dim3(2048), dim3(256), 0, s->stream, plane_set(s, s->cur), s->plane, s->type,
                       s->cfg.gamma, s->cfg.lx / s->cfg.nx, s->cfg.ly / s->cfg.ny, s->cfg.hll_bounds, acc);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

}  // namespace

extern "C" {

int32_t cfs_abi_version(void) { return CFS_ABI_VERSION; }

int32_t cfs_config_size(void) { return (int32_t)sizeof(cfs_config); }

const char* cfs_status_string(int32_t st) {
    switch (st) {
        case CFS_OK: return "CFS_OK";
        case CFS_E_ARG: return "CFS_E_ARG: invalid argument";
        case CFS_E_HIP: return "CFS_E_HIP: HIP runtime error";
        case CFS_E_OOM: return "CFS_E_OOM: device allocation failed";
        case CFS_E_STATE: return "CFS_E_STATE: call out of sequence";
        case CFS_E_NONFINITE: return "CFS_E_NONFINITE: non-finite reduction";
        case CFS_E_UNSUPPORTED: return "CFS_E_UNSUPPORTED: configuration not supported";
        default: return "unknown cfs status";
    }
}

int32_t cfs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int32_t cfs_create(const cfs_config* cfg, void* hip_stream, cfs_solver** out) {
    if (!out) return CFS_E_ARG;
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    cfs_solver* s = new (std::nothrow) cfs_solver();
    if (!s) return CFS_E_OOM;
    s->cfg = *cfg;
    s->plane = (long long)cfg->nx * cfg->ny;
    s->tiles_i = (cfg->nx + TI - 1) / TI;
    s->tiles_j = (cfg->ny + TJ - 1) / TJ;
    if (hip_stream) {
        s->stream = (hipStream_t)hip_stream;
    } else {
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) { delete s; return CFS_E_HIP; }
        s->own_stream = true;
    }
    if (hipMalloc(&s->planes, sizeof(double) * 12 * (size_t)s->plane) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    if (hipMalloc(&s->scal, sizeof(double) * SCAL_DOUBLES) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    s->tree_len = 4 * ((s->plane + 255) / 256 + 1);
    if (hipMalloc(&s->tree, sizeof(double) * 2 * s->tree_len) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    if (hipMalloc(&s->exact, sizeof(int) * 2) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    for (int k = 0; k < SCAL_DOUBLES; ++k)
        hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                           reinterpret_cast<unsigned long long*>(s->scal + k), 0ull);
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream, reinterpret_cast<unsigned long long*>(s->exact),
                       0ull);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) { cfs_destroy(s); return CFS_E_HIP; }
    *out = s;
    return CFS_OK;
}

void cfs_destroy(cfs_solver* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graph(s);
    for (auto e : s->ev) (void)hipEventDestroy(e);
    free_ib(s);
    (void)hipFree(s->type);
    (void)hipFree(s->planes);
    (void)hipFree(s->scal);
    (void)hipFree(s->tree);
    (void)hipFree(s->exact);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

static int32_t after_set_state(cfs_solver* s) {
    s->cur = 0;
    // prologue of the reference's solver: immersed-ghost fill (Run.cpp:106-112);
    // the physical boundary is applied inside every stage's tile load
    int rc = launch_ib_fill(s, plane_set(s, 0));
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_set_state(cfs_solver* s, const double* U) {
    if (!s || !U) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;
    HIP_OK(hipMemcpyAsync(plane_set(s, 0), U, sizeof(double) * 4 * s->plane, hipMemcpyHostToDevice, s->stream));
    return after_set_state(s);
}

int32_t cfs_set_state_device(cfs_solver* s, const double* dU) {
    if (!s || !dU) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;
    HIP_OK(hipMemcpyAsync(plane_set(s, 0), dU, sizeof(double) * 4 * s->plane, hipMemcpyDeviceToDevice, s->stream));
    return after_set_state(s);
}

int32_t cfs_get_state(cfs_solver* s, double* U) {
    if (!s || !U) return CFS_E_ARG;
    HIP_OK(hipMemcpyAsync(U, plane_set(s, s->cur), sizeof(double) * 4 * s->plane, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_get_state_device(cfs_solver* s, double* dU) {
    if (!s || !dU) return CFS_E_ARG;
    HIP_OK(hipMemcpyAsync(dU, plane_set(s, s->cur), sizeof(double) * 4 * s->plane, hipMemcpyDeviceToDevice, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_set_immersed(cfs_solver* s, int32_t interp, const uint8_t* cell_type, int32_t n_ig,
                         const int32_t* ig_cells, const int32_t* donors, const double* image_xy,
                         const double* ldlt, const int32_t* transp) {
    if (!s || !cell_type || n_ig < 0) return CFS_E_ARG;
    if (interp != CFS_IB_REFERENCE && interp != CFS_IB_BILINEAR_POSITIVE) return CFS_E_ARG;
    if (n_ig > 0 && (!ig_cells || !donors || !image_xy || !ldlt || !transp)) return CFS_E_ARG;
    const int nx = s->cfg.nx, ny = s->cfg.ny;
    std::vector<int> index((size_t)s->plane, -1);
    std::vector<uint8_t> tiles((size_t)s->tiles_i * s->tiles_j, 0);
    for (long long c = 0; c < s->plane; ++c)
        if (cell_type[c] > CFS_CELL_IMMERSED_GHOST) return CFS_E_ARG;
    for (int k = 0; k < n_ig; ++k) {
        if (ig_cells[k] < 0 || ig_cells[k] >= s->plane) return CFS_E_ARG;
        if (cell_type[ig_cells[k]] != CFS_CELL_IMMERSED_GHOST || index[ig_cells[k]] >= 0) return CFS_E_ARG;
        index[ig_cells[k]] = k;
        for (int d = 0; d < 4; ++d) {
            const int c = donors[4 * k + d];
            if (c < 0 || c >= s->plane || cell_type[c] != CFS_CELL_FLUID) return CFS_E_ARG;
            const int t = transp[4 * k + d];
            if (interp == CFS_IB_REFERENCE && (t < d || t > 3)) return CFS_E_ARG;
        }
        // every tile whose staged region (tile + 3-cell halo, periodic wrap)
        // holds this ghost fills it while loading
        const int gi = ig_cells[k] / ny, gj = ig_cells[k] % ny;
        for (int di = -HALO; di <= HALO; ++di)
            for (int dj = -HALO; dj <= HALO; ++dj) {
                int a = gi + di, b = gj + dj;
                if (s->cfg.bc[0] == CFS_BC_PERIODIC) a = (a % nx + nx) % nx;
                if (s->cfg.bc[2] == CFS_BC_PERIODIC) b = (b % ny + ny) % ny;
                if (a < 0 || a >= nx || b < 0 || b >= ny) continue;
                tiles[(size_t)(a / TI) * s->tiles_j + b / TJ] = 1;
            }
    }
    for (long long c = 0; c < s->plane; ++c)
        if (cell_type[c] == CFS_CELL_IMMERSED_GHOST && index[c] < 0) return CFS_E_ARG;   // ghost without a row
    HIP_OK(hipStreamSynchronize(s->stream));
    drop_graph(s);
    free_ib(s);
    if (!s->type && hipMalloc(&s->type, (size_t)s->plane) != hipSuccess) return CFS_E_OOM;
    HIP_OK(hipMemcpy(s->type, cell_type, (size_t)s->plane, hipMemcpyHostToDevice));
    if (n_ig > 0) {
        if (hipMalloc(&s->ig_cells, sizeof(int) * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_donors, sizeof(int) * 4 * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_xy, sizeof(double) * 2 * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_ldl, sizeof(double) * 10 * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_transp, sizeof(int) * 4 * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_index, sizeof(int) * (size_t)s->plane) != hipSuccess ||
            hipMalloc(&s->tile_ib, tiles.size()) != hipSuccess) {
            free_ib(s);
            return CFS_E_OOM;
        }
        HIP_OK(hipMemcpy(s->ig_cells, ig_cells, sizeof(int) * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_donors, donors, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_xy, image_xy, sizeof(double) * 2 * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_ldl, ldlt, sizeof(double) * 10 * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_transp, transp, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_index, index.data(), sizeof(int) * (size_t)s->plane, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->tile_ib, tiles.data(), tiles.size(), hipMemcpyHostToDevice));
    }
    s->n_ig = n_ig;
    s->ib_interp = interp;
    return CFS_OK;
}

int32_t cfs_stage(cfs_solver* s, int32_t stage_counter, double dt) {
    if (!s) return CFS_E_ARG;
    if (stage_counter != s->counter) return CFS_E_STATE;
    if (!(dt >= 0.0)) return CFS_E_ARG;
    int rc = set_dt_device(s, dt);
    if (rc) return rc;
    rc = launch_stage(s, stage_counter, false, StageTime{});
    if (rc) return rc;
    s->counter++;
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_step(cfs_solver* s, double dt) {
    if (!s) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;
    if (!(dt >= 0.0)) return CFS_E_ARG;
    int rc = set_dt_device(s, dt);
    if (rc) return rc;
    rc = launch_step(s, false, false, 0.0, 0.0);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_advance(cfs_solver* s, int32_t nsteps, double dt, double cfl, double t_end, double* t_io,
                    int32_t* steps_done) {
    if (!s || nsteps < 0) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;
    const bool use_cfl = !(dt > 0.0);
    if (use_cfl && !(cfl > 0.0)) return CFS_E_ARG;
    const double t0 = t_io ? *t_io : 0.0;
    for (auto e : s->ev) (void)hipEventDestroy(e);
    s->ev.clear();
    const bool timed = s->kernel_timing;
    int rc;
    int done = 0;
    double t = t0;

    if (!use_cfl) {
        // fixed dt; an end time clips the last step on the host
        rc = set_dt_device(s, dt);
        if (rc) return rc;
        for (int n = 0; n < nsteps; ++n) {
            if (t_end > 0.0) {
                double dtx = dt;
                if (!(t + dtx < t_end)) dtx = t_end - t;
                if (!(dtx > 0.0)) break;
                if (dtx != dt && (rc = set_dt_device(s, dtx))) return rc;
                t += dtx;
            } else {
                t += dt;
            }
            rc = launch_step(s, timed, false, 0.0, 0.0);
            if (rc) return rc;
            ++done;
        }
        HIP_OK(hipStreamSynchronize(s->stream));
    } else {
        // CFL steps, all on the device: the bound of the entering state is
        // reduced once here, then every step's first stage derives dt from it and
        // its last stage accumulates the bound of the new state
        const int p0 = s->step_parity;
        hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, t_ptr(s, p0), t0);
        hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                           reinterpret_cast<unsigned long long*>(halt_ptr(s)), 0ull);
        hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                           reinterpret_cast<unsigned long long*>(steps_ptr(s)), 0ull);
        HIP_OK(hipGetLastError());
        rc = launch_cfl(s, acc_ptr(s, p0));
        if (rc) return rc;
        int n = 0;
        if (!timed && nsteps >= 4) {
            // two steps per graph (time and bound slots alternate by step parity)
            GraphKey key;
            key.valid = true;
            key.use_cfl = true;
            key.cfl = cfl;
            key.t_end = t_end;
            key.step_parity = s->step_parity;
            key.stage_parity = s->parity;
            const bool same = s->graph_key.valid && s->graph_key.use_cfl == key.use_cfl &&
                              s->graph_key.cfl == key.cfl && s->graph_key.t_end == key.t_end &&
                              s->graph_key.step_parity == key.step_parity &&
                              s->graph_key.stage_parity == key.stage_parity;
            const int counter0 = s->counter;
            if (!same) {
                drop_graph(s);
                HIP_OK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                int crc = launch_step(s, false, true, cfl, t_end);
                if (!crc) crc = launch_step(s, false, true, cfl, t_end);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(s->stream, &g);
                if (crc) {
                    if (g) (void)hipGraphDestroy(g);
                    return crc;
                }
                HIP_OK(ec);
                s->graph = g;
                HIP_OK(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
                s->graph_key = key;
                s->counter = counter0;   // capture launched nothing
            }
            // the captured launches left step/stage parities and planes where two
            // real steps leave them
            for (; n + 2 <= nsteps; n += 2) {
                HIP_OK(hipGraphLaunch(s->graph_exec, s->stream));
                s->counter += 2 * s->cfg.rk;
            }
            s->cur = 0;
        }
        for (; n < nsteps; ++n) {
            rc = launch_step(s, timed, true, cfl, t_end);
            if (rc) return rc;
        }
        HIP_OK(hipStreamSynchronize(s->stream));
        int ints[2] = {0, 0};
        HIP_OK(hipMemcpy(ints, halt_ptr(s), sizeof(ints), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(&done, steps_ptr(s), sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(&t, t_ptr(s, s->step_parity), sizeof(double), hipMemcpyDeviceToHost));
        if (ints[1]) {
            if (t_io) *t_io = t;
            if (steps_done) *steps_done = done;
            return CFS_E_NONFINITE;
        }
    }
    double tot = 0.0;
    for (size_t k = 0; k + 1 < s->ev.size(); k += 2) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]));
        tot += ms;
    }
    if (!s->ev.empty()) s->last_kernel_ms = tot / (double)(s->ev.size() / 2);
    if (t_io) *t_io = t;
    if (steps_done) *steps_done = done;
    return CFS_OK;
}

int32_t cfs_compute_dt(cfs_solver* s, double cfl, double* dt_out) {
    if (!s || !dt_out || !(cfl > 0.0)) return CFS_E_ARG;
    int rc = launch_cfl(s, acc_tmp_ptr(s));
    if (rc) return rc;
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, acc_tmp_ptr(s), sizeof(bits), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    double smax;
    std::memcpy(&smax, &bits, sizeof(smax));
    if (!(smax > 0.0) || !(smax < 1.0e308)) return CFS_E_NONFINITE;
    *dt_out = cfl / smax;
    return CFS_OK;
}

int32_t cfs_checksums(cfs_solver* s, double out[4]) {
    if (!s || !out) return CFS_E_ARG;
    const double* src = plane_set(s, s->cur);
    long long n = s->plane, stride = s->plane;
    const uint8_t* type = s->type;
    double* bufs[2] = {s->tree, s->tree + s->tree_len};
    int which = 0;
    while (true) {
        const long long nb = (n + 255) / 256;
        double* dst = bufs[which];
        hipLaunchKernelGGL(tree_sum_kernel, dim3((unsigned)nb, 4), dim3(256), 0, s->stream, src, n, stride, type,
                           dst, nb);
        HIP_OK(hipGetLastError());
        src = dst;
        stride = nb;
        n = nb;
        type = nullptr;
        which ^= 1;
        if (nb == 1) break;
    }
    double tmp[4];
    HIP_OK(hipMemcpyAsync(tmp, src, sizeof(double) * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    for (int q = 0; q < 4; ++q) out[q] = tmp[q];
    return CFS_OK;
}

int32_t cfs_calibrate_copy(cfs_solver* s, int32_t repeats) {
    if (!s || repeats < 0) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;   // spare sets hold RK stages mid-step
    const double2* src = reinterpret_cast<const double2*>(plane_set(s, s->cur));
    double2* dst = reinterpret_cast<double2*>(plane_set(s, (s->cur + 1) % 3));
    const long long n = 2 * s->plane;   // 4*plane doubles = 2*plane double2
    for (int k = 0; k < repeats; ++k) {
        hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s->stream, src, dst, n);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_set_kernel_timing(cfs_solver* s, int32_t on) {
    if (!s) return CFS_E_ARG;
    s->kernel_timing = on != 0;
    return CFS_OK;
}

int32_t cfs_last_redo_tiles(cfs_solver* s, int32_t* n) {
    if (!s || !n) return CFS_E_ARG;
    int v = 0;   // the counter of the last stage: its parity was toggled after launch
    HIP_OK(hipMemcpyAsync(&v, s->exact + (1 - s->parity), sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    *n = v;
    return CFS_OK;
}

int32_t cfs_stage_counter(const cfs_solver* s) { return s ? s->counter : -1; }

int32_t cfs_last_kernel_ms(const cfs_solver* s, double* ms) {
    if (!s || !ms) return CFS_E_ARG;
    *ms = s->last_kernel_ms;
    return CFS_OK;
}

}  // extern "C"
