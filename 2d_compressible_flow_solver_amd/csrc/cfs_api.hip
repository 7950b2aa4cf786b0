// cfs_api.hip -- the C-ABI of include/cfs.h on top of the gfx950 kernels.
//
// A cfs_solver owns: three FP64 state plane-sets (RK ping-pong), an optional
// cell-type plane and immersed-boundary tables, device scalars for dt, time,
// the CFL accumulator and a status word, scratch for the checksum tree, and
// one HIP stream.  Nothing here falls back to a CPU path: every compute entry
// point launches HIP kernels and fails with CFS_E_HIP when HIP does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cfs.h"
#include "kernels.hpp"

using namespace cfs;

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
    // device scalars: [0] dt, [1] t, [2] acc (u64), [3] status (int)
    double* scal = nullptr;
    double* tree = nullptr;       // checksum scratch
    long long tree_len = 0;
    int* redo = nullptr;          // per-stage-parity counters of tiles that took IEEE x/3, x/6
    int parity = 0;
    bool kernel_timing = false;   // bracket stage kernels with HIP events in cfs_advance
    int counter = 1;
    int tiles_i = 0, tiles_j = 0;
    double last_kernel_ms = 0.0;
    // instrumentation: events around stage kernels
    std::vector<hipEvent_t> ev;
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
double* t_ptr(cfs_solver* s) { return s->scal + 1; }
unsigned long long* acc_ptr(cfs_solver* s) { return reinterpret_cast<unsigned long long*>(s->scal + 2); }
int* status_ptr(cfs_solver* s) { return reinterpret_cast<int*>(s->scal + 3); }

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
    if (c->viscous != 0) return CFS_E_UNSUPPORTED;
    return CFS_OK;
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

// One RK stage with dt read from the device scalar.  Plane-sets: A = 0 holds
// U^n at a step boundary; RK3: A->B, B(+A)->C, C(+A,B)->A; RK2: A->B, B(+A)->A.
int launch_stage(cfs_solver* s, int counter, bool timed) {
    const cfs_config& c = s->cfg;
    const int sub = rk_substage(c.rk, counter);
    StageArgs a{};
    double* A = plane_set(s, 0);
    double* B = plane_set(s, 1);
    double* Cp = plane_set(s, 2);
    a.type = s->type;
    a.dt = dt_ptr(s);
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
    a.exact_count = s->redo + s->parity;
    a.exact_other = s->redo + (1 - s->parity);
    for (int k = 0; k < 4; ++k) {
        a.bc[k] = c.bc[k];
        for (int q = 0; q < 4; ++q) a.bcs[k][q] = c.bc_state[k][q];
    }
    int next;
    if (c.rk == 3) {
        if (sub == 0) { a.in = A; a.out = B; a.form = 0; a.c0 = 0.0; a.c1 = 1.0; a.c2 = 1.0; next = 1; }
        else if (sub == 1) { a.in = B; a.u0 = A; a.out = Cp; a.form = 1; a.c0 = 0.75; a.c1 = 0.25; a.c2 = 0.25; next = 2; }
        else { a.in = Cp; a.u0 = A; a.u1 = B; a.out = A; a.form = 2; a.c0 = 1.0 / 3.0; a.c1 = 2.0 / 3.0; a.c2 = 2.0 / 3.0; next = 0; }
    } else {
        if (sub == 0) { a.in = A; a.out = B; a.form = 0; a.c0 = 0.0; a.c1 = 1.0; a.c2 = 1.0; next = 1; }
        else { a.in = B; a.u0 = A; a.out = A; a.form = 1; a.c0 = 0.5; a.c1 = 0.5; a.c2 = 0.5; next = 0; }
    }
    int rc = launch_ib_fill(s, const_cast<double*>(a.in));
    if (rc) return rc;
    const dim3 grid((unsigned)(s->tiles_i * s->tiles_j));
    if (timed) {
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        s->ev.push_back(e0);
        s->ev.push_back(e1);
        HIP_OK(hipEventRecord(e0, s->stream));
    }
    if (c.precision == CFS_PREC_FP64)
        hipLaunchKernelGGL(stage_kernel<double>, grid, dim3(BLOCK), 0, s->stream, a);
    else
        hipLaunchKernelGGL(stage_kernel<float>, grid, dim3(BLOCK), 0, s->stream, a);
    HIP_OK(hipGetLastError());
    if (timed) HIP_OK(hipEventRecord(s->ev.back(), s->stream));
    s->parity ^= 1;
    s->cur = next;
    return CFS_OK;
}

int set_dt_device(cfs_solver* s, double dt) {
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, dt_ptr(s), dt);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

int launch_cfl(cfs_solver* s, double cfl, double t_end) {
    const double* U = plane_set(s, s->cur);
    const int blocks = 2048;
    hipLaunchKernelGGL(cfl_kernel, dim3(blocks), dim3(256), 0, s->stream, U, s->plane, s->type,
                       s->cfg.gamma, s->cfg.lx / s->cfg.nx, s->cfg.ly / s->cfg.ny, s->cfg.hll_bounds, acc_ptr(s));
    HIP_OK(hipGetLastError());
    DtArgs d{};
    d.acc = acc_ptr(s);
    d.dt = dt_ptr(s);
    d.t = t_ptr(s);
    d.cfl = cfl;
    d.t_end = t_end;
    d.status = status_ptr(s);
    hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(1), 0, s->stream, d);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

int free_ib(cfs_solver* s) {
    hipFree(s->ig_cells); hipFree(s->ig_donors); hipFree(s->ig_xy); hipFree(s->ig_ldl); hipFree(s->ig_transp);
    s->ig_cells = s->ig_donors = s->ig_transp = nullptr;
    s->ig_xy = s->ig_ldl = nullptr;
    s->n_ig = 0;
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
    if (hipMalloc(&s->scal, sizeof(double) * 8) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    s->tree_len = 4 * ((s->plane + 255) / 256 + 1);
    if (hipMalloc(&s->tree, sizeof(double) * 2 * s->tree_len) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    if (hipMalloc(&s->redo, sizeof(int) * 2) != hipSuccess) {
        cfs_destroy(s);
        return CFS_E_OOM;
    }
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream, reinterpret_cast<unsigned long long*>(s->redo),
                       0ull);
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, dt_ptr(s), 0.0);
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, t_ptr(s), 0.0);
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream, acc_ptr(s), 0ull);
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                       reinterpret_cast<unsigned long long*>(status_ptr(s)), 0ull);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) { cfs_destroy(s); return CFS_E_HIP; }
    *out = s;
    return CFS_OK;
}

void cfs_destroy(cfs_solver* s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
    for (auto e : s->ev) hipEventDestroy(e);
    free_ib(s);
    hipFree(s->type);
    hipFree(s->planes);
    hipFree(s->scal);
    hipFree(s->tree);
    hipFree(s->redo);
    if (s->own_stream && s->stream) hipStreamDestroy(s->stream);
    delete s;
}

static int32_t after_set_state(cfs_solver* s) {
    s->cur = 0;
    // prologue of the reference's solver: immersed-ghost fill (Run.cpp:106-112);
    // the physical boundary is applied inside every stage's tile load.
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
    for (int k = 0; k < n_ig; ++k) {
        if (ig_cells[k] < 0 || ig_cells[k] >= s->plane) return CFS_E_ARG;
        if (cell_type[ig_cells[k]] != CFS_CELL_IMMERSED_GHOST) return CFS_E_ARG;
        for (int d = 0; d < 4; ++d) {
            const int c = donors[4 * k + d];
            if (c < 0 || c >= s->plane || cell_type[c] != CFS_CELL_FLUID) return CFS_E_ARG;
            const int t = transp[4 * k + d];
            if (interp == CFS_IB_REFERENCE && (t < d || t > 3)) return CFS_E_ARG;
        }
    }
    HIP_OK(hipStreamSynchronize(s->stream));
    free_ib(s);
    if (!s->type && hipMalloc(&s->type, (size_t)s->plane) != hipSuccess) return CFS_E_OOM;
    HIP_OK(hipMemcpy(s->type, cell_type, (size_t)s->plane, hipMemcpyHostToDevice));
    if (n_ig > 0) {
        if (hipMalloc(&s->ig_cells, sizeof(int) * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_donors, sizeof(int) * 4 * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_xy, sizeof(double) * 2 * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_ldl, sizeof(double) * 10 * n_ig) != hipSuccess ||
            hipMalloc(&s->ig_transp, sizeof(int) * 4 * n_ig) != hipSuccess) {
            free_ib(s);
            return CFS_E_OOM;
        }
        HIP_OK(hipMemcpy(s->ig_cells, ig_cells, sizeof(int) * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_donors, donors, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_xy, image_xy, sizeof(double) * 2 * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_ldl, ldlt, sizeof(double) * 10 * n_ig, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->ig_transp, transp, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice));
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
    rc = launch_stage(s, stage_counter, false);
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
    for (int k = 0; k < s->cfg.rk; ++k) {
        rc = launch_stage(s, s->counter, false);
        if (rc) return rc;
        s->counter++;
    }
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
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, t_ptr(s), t0);
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                       reinterpret_cast<unsigned long long*>(status_ptr(s)), 0ull);
    HIP_OK(hipGetLastError());
    for (auto e : s->ev) hipEventDestroy(e);
    s->ev.clear();
    const bool timed = s->kernel_timing;
    int rc;
    if (!use_cfl) {
        double dtx = dt;
        if (t_end > 0.0) {
            // fixed dt with end-time clipping is resolved on the host
            rc = set_dt_device(s, dtx);
            if (rc) return rc;
        } else {
            rc = set_dt_device(s, dtx);
            if (rc) return rc;
        }
    }
    int done = 0;
    double t = t0;
    for (int n = 0; n < nsteps; ++n) {
        if (use_cfl) {
            rc = launch_cfl(s, cfl, t_end);
            if (rc) return rc;
        } else if (t_end > 0.0) {
            double dtx = dt;
            if (!(t + dtx < t_end)) dtx = t_end - t;
            if (!(dtx > 0.0)) break;
            rc = set_dt_device(s, dtx);
            if (rc) return rc;
            t += dtx;
        } else {
            t += dt;
        }
        for (int k = 0; k < s->cfg.rk; ++k) {
            rc = launch_stage(s, s->counter, timed);
            if (rc) return rc;
            s->counter++;
        }
        ++done;
        if (use_cfl && t_end > 0.0) {
            // stop once the device time reached t_end (one host sync per step)
            double tt = 0.0;
            HIP_OK(hipMemcpyAsync(&tt, t_ptr(s), sizeof(double), hipMemcpyDeviceToHost, s->stream));
            HIP_OK(hipStreamSynchronize(s->stream));
            t = tt;
            if (!(t < t_end)) break;
        }
    }
    HIP_OK(hipStreamSynchronize(s->stream));
    int status = 0;
    HIP_OK(hipMemcpy(&status, status_ptr(s), sizeof(int), hipMemcpyDeviceToHost));
    if (use_cfl) HIP_OK(hipMemcpy(&t, t_ptr(s), sizeof(double), hipMemcpyDeviceToHost));
    double tot = 0.0;
    for (size_t k = 0; k + 1 < s->ev.size(); k += 2) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]));
        tot += ms;
    }
    s->last_kernel_ms = s->ev.empty() ? 0.0 : tot / (double)(s->ev.size() / 2);
    if (t_io) *t_io = t;
    if (steps_done) *steps_done = done;
    if (status) return CFS_E_NONFINITE;
    return CFS_OK;
}

int32_t cfs_compute_dt(cfs_solver* s, double cfl, double* dt_out) {
    if (!s || !dt_out || !(cfl > 0.0)) return CFS_E_ARG;
    const double* U = plane_set(s, s->cur);
    hipLaunchKernelGGL(cfl_kernel, dim3(2048), dim3(256), 0, s->stream, U, s->plane, s->type, s->cfg.gamma,
                       s->cfg.lx / s->cfg.nx, s->cfg.ly / s->cfg.ny, s->cfg.hll_bounds, acc_ptr(s));
    HIP_OK(hipGetLastError());
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, acc_ptr(s), sizeof(bits), hipMemcpyDeviceToHost, s->stream));
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream, acc_ptr(s), 0ull);
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
    HIP_OK(hipMemcpyAsync(&v, s->redo + (1 - s->parity), sizeof(int), hipMemcpyDeviceToHost, s->stream));
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
