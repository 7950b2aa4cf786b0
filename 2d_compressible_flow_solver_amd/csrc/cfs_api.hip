// cfs_api.hip -- the C-ABI of include/cfs.h on top of the gfx950 kernels.
//
// A cfs_solver owns: three FP64 state plane-sets (RK ping-pong), an optional
// cell-type plane and immersed-boundary tables (plus a per-cell ghost index
// and a per-tile "near the body" flag for the fused fill), a block of device
// scalars (dt, time, CFL bounds, halt/status/step counters), scratch for the
// checksum tree, one HIP stream and a cached HIP graph of RK steps.
// Nothing here falls back to a CPU path: every compute entry point launches
// HIP kernels and fails with CFS_E_HIP when HIP does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/cfs.h"
#include "host.hpp"
#include "kernels.hpp"

using namespace cfs;

// Device scratch of cfs_checksums_device_noalloc: the partial sums of its
// fixed-order tree.  Level one holds 4 x ceil(cells/256) <= 4 x 2^18 sums,
// level two 4 x ceil(cells/256^2) <= 4 x 2^10; the later levels (at most 4 x 4
// sums, then the 4 results, written to the caller's output) reuse level one's
// region.  Every entry a level reads was written by the level before it in the
// same call, so nothing needs clearing between calls.  Host code reaches the
// array only through the per-device address from hipGetSymbolAddress.
namespace {
constexpr long long NOALLOC_L1 = 4 * (CFS_NOALLOC_MAX_CELLS / 256);
constexpr long long NOALLOC_L2 = 4 * (CFS_NOALLOC_MAX_CELLS / (256 * 256));
}  // namespace
static __device__ __align__(256) double scratch[NOALLOC_L1 + NOALLOC_L2];

namespace {

struct GraphKey {
    bool valid = false;
    bool use_cfl = false;
    double cfl = 0.0, t_end = 0.0;
    int step_parity = 0, stage_parity = 0;
    int steps = 0;   // RK steps captured
};

}  // namespace

struct cfs_solver {
    cfs_config cfg{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    ProblemDev dev;               // device buffers (planes, types, IB tables, scalars)
    int cur = 0;                  // plane-set holding the current state
    double* tree = nullptr;       // checksum scratch
    double* rhs = nullptr;        // cfs_last_rhs output (allocated on first use) + 2 ints of counters
    long long tree_len = 0;
    int parity = 0;               // stage parity (exact-range counters)
    int step_parity = 0;          // step parity (time and CFL-bound slots)
    int counter = 1;
    bool kernel_timing = false;   // bracket stage kernels with HIP events in cfs_advance
    double last_kernel_ms = 0.0;
    std::vector<hipEvent_t> ev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    GraphKey graph_key;
    int shape = 0;                              // tile shape of the stage kernels
    int cus = 0;                                // compute units of the handle's device (0: unknown)
    uint8_t* tile_ib_shape[NUM_SHAPES] = {};    // per-tile IB flags of every shape
};

namespace {

ScalPtrs scal(cfs_solver* s) { return s->dev.s(); }

void drop_graph(cfs_solver* s) {
    SetupLock lk(hip_setup_mutex());
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    s->graph_key = GraphKey{};
}

template <typename R, bool VISC, bool LIMIT, typename Sh>
void launch_stage_kernel(dim3 grid, hipStream_t stream, const StageArgs& a) {
    hipLaunchKernelGGL((stage_kernel<R, VISC, LIMIT, Sh>), grid, dim3(Sh::NT), 0, stream, a);
}

template <typename R, bool VISC, bool LIMIT>
void launch_shape(int shape, dim3 grid, hipStream_t stream, const StageArgs& a) {
    if (VISC || shape == 0) launch_stage_kernel<R, VISC, LIMIT, Shape0>(grid, stream, a);
    else if (shape == 1) launch_stage_kernel<R, false, LIMIT, Shape1>(grid, stream, a);
    else if (shape == 2) launch_stage_kernel<R, false, LIMIT, Shape2>(grid, stream, a);
    else launch_stage_kernel<R, false, LIMIT, Shape3>(grid, stream, a);
}

// the kernel variant of a configuration: precision, viscous terms, limiter,
// tile shape (the viscous variant has the 15 x 16 shape only)
void dispatch_stage(const cfs_config& c, int shape, dim3 grid, hipStream_t stream, const StageArgs& a) {
    const bool limit = c.positivity == 2;
    if (c.precision == CFS_PREC_FP64) {
        if (c.viscous) launch_shape<double, true, false>(0, grid, stream, a);
        else if (limit) launch_shape<double, false, true>(shape, grid, stream, a);
        else launch_shape<double, false, false>(shape, grid, stream, a);
    } else {
        if (c.viscous) launch_shape<float, true, false>(0, grid, stream, a);
        else if (limit) launch_shape<float, false, true>(shape, grid, stream, a);
        else launch_shape<float, false, false>(shape, grid, stream, a);
    }
}

dim3 stage_grid(const cfs_solver* s) { return dim3((unsigned)(8 * ((s->dev.tiles_i * s->dev.tiles_j + 7) / 8))); }

// Grids of at most 16 tiles per CU read each tile's first staging pass
// before the time step is known: there every workgroup starts in the first
// round, so a stage costs one tile's latency chain and the load overlaps the
// time-step scalars' (Lax-Liu 200^2 in CFL steps: 13.6 -> 12.9 us per stage,
// 400^2: 22.5 -> 22.3 us).  On larger grids the first round's latency is
// hidden by the later rounds, while a launch replayed past the end time
// would read every tile (a 4096^2 grid: 179 instead of 49 us per such launch).
bool stage_early(const cfs_solver* s) {
    return s->cus > 0 && (long long)s->dev.tiles_i * s->dev.tiles_j <= 16ll * s->cus;
}

// The tile shape a new handle starts with: the dual-face shape 3 for
// inviscid grids of at most 5000 tiles of 15 x 16 (up to about 1100^2
// cells) without the positivity limiter, where its two chains per thread
// beat shape 0 (Lax-Liu 400^2: 7%); shape 0 above (the 2048 x 1024 cylinder:
// 1% the other way) and with the limiter (served Riemann problems 512^2 to
// 1024^2: shape 3 is 0-1% slower in FP64, up to 7% in FP32-bulk).
// cfs_autotune measures instead.
int default_shape(const cfs_config& c) {
    return (!c.viscous && c.positivity != 2 && tiles_for(c, 0) <= 5000) ? 3 : 0;
}

// Makes tile shape k current: tile counts and the IB tile flags of that shape.
void apply_shape(cfs_solver* s, int k) {
    const ShapeDims d = shape_dims(k);
    s->shape = k;
    s->dev.tiles_i = (s->cfg.nx + d.ti - 1) / d.ti;
    s->dev.tiles_j = (s->cfg.ny + d.tj - 1) / d.tj;
    s->dev.tile_ib = s->tile_ib_shape[k];
}

// One RK stage of the global stage counter.
int launch_stage(cfs_solver* s, int counter, bool timed, const StageTime& tm) {
    const cfs_config& c = s->cfg;
    StageArgs a;
    const int next = fill_stage_args(a, s->dev, rk_substage(c.rk, counter), s->step_parity, s->parity, tm);
    a.stage_early = stage_early(s) ? 1 : 0;
    const dim3 grid = stage_grid(s);
    if (timed) {
        SetupLock lk(hip_setup_mutex());
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        s->ev.push_back(e0);
        s->ev.push_back(e1);
        HIP_OK(hipEventRecord(e0, s->stream));
    }
    dispatch_stage(c, s->shape, grid, s->stream, a);
    HIP_OK(hipGetLastError());
    if (timed) HIP_OK(hipEventRecord(s->ev.back(), s->stream));
    s->parity ^= 1;
    s->cur = next;
    return CFS_OK;
}

// One RK step (all sub-stages) in the given time mode.
int launch_step(cfs_solver* s, bool timed, bool use_cfl, double cfl, double t_end) {
    for (int k = 0; k < s->cfg.rk; ++k) {
        const StageTime tm = use_cfl ? cfl_stage_time(s->cfg.rk, k, cfl, t_end) : StageTime{};
        const int rc = launch_stage(s, s->counter, timed, tm);
        if (rc) return rc;
        s->counter++;
    }
    if (use_cfl) s->step_parity ^= 1;
    return CFS_OK;
}

// steps between two reads of the device halt flag in a CFL advance
constexpr int HALT_POLL = 64;
// RK steps per graph replay of a long CFL advance (even: the time and bound
// slots alternate by step parity; a divisor of HALT_POLL).  Consecutive
// replays leave a few microseconds between them, about 2% of a 400^2 step
// with two steps per graph.
constexpr int GRAPH_STEPS = 8;
static_assert(GRAPH_STEPS % 2 == 0 && HALT_POLL % GRAPH_STEPS == 0, "graph steps");

int read_halt(cfs_solver* s, bool* halted) {
    int h = 0;
    HIP_OK(hipMemcpyAsync(&h, scal(s).halt(), sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    *halted = h != 0;
    return CFS_OK;
}

int set_dt_device(cfs_solver* s, double dt) {
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, scal(s).dt(), dt);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

void free_ib(cfs_solver* s) {
    SetupLock lk(hip_setup_mutex());
    ProblemDev& d = s->dev;
    (void)hipFree(d.ig_cells);
    (void)hipFree(d.ig_donors);
    (void)hipFree(d.ig_xy);
    (void)hipFree(d.ig_ldl);
    (void)hipFree(d.ig_transp);
    (void)hipFree(d.ig_index);
    for (int k = 0; k < NUM_SHAPES; ++k) {
        (void)hipFree(s->tile_ib_shape[k]);
        s->tile_ib_shape[k] = nullptr;
    }
    d.ig_cells = d.ig_donors = d.ig_transp = d.ig_index = nullptr;
    d.ig_xy = d.ig_ldl = nullptr;
    d.tile_ib = nullptr;
    d.n_ig = 0;
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
        case CFS_E_FULL: return "CFS_E_FULL: no free batch slot or arena space";
        case CFS_E_LIMIT: return "CFS_E_LIMIT: input larger than the entry point's fixed device scratch";
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
    SetupLock lk(hip_setup_mutex());
    cfs_solver* s = new (std::nothrow) cfs_solver();
    if (!s) return CFS_E_OOM;
    s->cfg = *cfg;
    ProblemDev& d = s->dev;
    d.cfg = &s->cfg;
    d.plane = (long long)cfg->nx * cfg->ny;
    apply_shape(s, default_shape(*cfg));
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            s->cus = cus;
    }
    if (hip_stream) {
        s->stream = (hipStream_t)hip_stream;
    } else {
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) { delete s; return CFS_E_HIP; }
        s->own_stream = true;
    }
    if (hipMalloc(&d.planes, sizeof(double) * 12 * (size_t)d.plane) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    if (hipMalloc(&d.scal, sizeof(double) * SCAL_DOUBLES) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    s->tree_len = 4 * ((d.plane + 255) / 256 + 1);
    if (hipMalloc(&s->tree, sizeof(double) * 2 * s->tree_len) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    if (hipMalloc(&d.exact, sizeof(int) * 2) != hipSuccess) { cfs_destroy(s); return CFS_E_OOM; }
    for (int k = 0; k < SCAL_DOUBLES; ++k)
        hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                           reinterpret_cast<unsigned long long*>(d.scal + k), 0ull);
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream, reinterpret_cast<unsigned long long*>(d.exact),
                       0ull);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) { cfs_destroy(s); return CFS_E_HIP; }
    *out = s;
    return CFS_OK;
}

void cfs_destroy(cfs_solver* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    SetupLock lk(hip_setup_mutex());
    drop_graph(s);
    for (auto e : s->ev) (void)hipEventDestroy(e);
    free_ib(s);
    (void)hipFree(s->dev.type);
    (void)hipFree(s->dev.planes);
    (void)hipFree(s->dev.scal);
    (void)hipFree(s->tree);
    (void)hipFree(s->rhs);
    (void)hipFree(s->dev.exact);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

static int32_t after_set_state(cfs_solver* s) {
    s->cur = 0;
    // prologue of the reference's solver: immersed-ghost fill (Run.cpp:106-112);
    // the physical boundary is applied inside every stage's tile load
    int rc = launch_ib_fill(s->dev, s->dev.set(0), s->stream);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_set_state(cfs_solver* s, const double* U) {
    if (!s || !U) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;
    HIP_OK(hipMemcpyAsync(s->dev.set(0), U, sizeof(double) * 4 * s->dev.plane, hipMemcpyHostToDevice, s->stream));
    return after_set_state(s);
}

int32_t cfs_set_state_device(cfs_solver* s, const double* dU) {
    if (!s || !dU) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;
    HIP_OK(hipMemcpyAsync(s->dev.set(0), dU, sizeof(double) * 4 * s->dev.plane, hipMemcpyDeviceToDevice, s->stream));
    return after_set_state(s);
}

int32_t cfs_get_state(cfs_solver* s, double* U) {
    if (!s || !U) return CFS_E_ARG;
    HIP_OK(hipMemcpyAsync(U, s->dev.set(s->cur), sizeof(double) * 4 * s->dev.plane, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_get_state_device(cfs_solver* s, double* dU) {
    if (!s || !dU) return CFS_E_ARG;
    HIP_OK(hipMemcpyAsync(dU, s->dev.set(s->cur), sizeof(double) * 4 * s->dev.plane, hipMemcpyDeviceToDevice, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_set_immersed(cfs_solver* s, int32_t interp, const uint8_t* cell_type, int32_t n_ig,
                         const int32_t* ig_cells, const int32_t* donors, const double* image_xy,
                         const double* ldlt, const int32_t* transp) {
    if (!s) return CFS_E_ARG;
    if (n_ig > 0 && (!image_xy || !ldlt)) return CFS_E_ARG;
    ProblemDev& d = s->dev;
    std::vector<int> index;
    std::vector<uint8_t> tiles[NUM_SHAPES];
    int rc = build_ib_index(s->cfg, 0, interp, cell_type, n_ig, ig_cells, donors, transp, index, tiles[0]);
    if (rc) return rc;
    for (int k = 1; k < NUM_SHAPES; ++k) ib_tile_flags(s->cfg, k, n_ig, ig_cells, tiles[k]);
    HIP_OK(hipStreamSynchronize(s->stream));
    SetupLock lk(hip_setup_mutex());
    drop_graph(s);
    free_ib(s);
    if (!d.type && hipMalloc(&d.type, (size_t)d.plane) != hipSuccess) return CFS_E_OOM;
    // copies on the handle's stream (the legacy default stream would meet
    // other threads' streams), waited for before the host tables go away
    HIP_OK(hipMemcpyAsync(d.type, cell_type, (size_t)d.plane, hipMemcpyHostToDevice, s->stream));
    if (n_ig > 0) {
        if (hipMalloc(&d.ig_cells, sizeof(int) * n_ig) != hipSuccess ||
            hipMalloc(&d.ig_donors, sizeof(int) * 4 * n_ig) != hipSuccess ||
            hipMalloc(&d.ig_xy, sizeof(double) * 2 * n_ig) != hipSuccess ||
            hipMalloc(&d.ig_ldl, sizeof(double) * 10 * n_ig) != hipSuccess ||
            hipMalloc(&d.ig_transp, sizeof(int) * 4 * n_ig) != hipSuccess ||
            hipMalloc(&d.ig_index, sizeof(int) * (size_t)d.plane) != hipSuccess) {
            free_ib(s);
            return CFS_E_OOM;
        }
        for (int k = 0; k < NUM_SHAPES; ++k)
            if (hipMalloc(&s->tile_ib_shape[k], tiles[k].size()) != hipSuccess) {
                free_ib(s);
                return CFS_E_OOM;
            }
        HIP_OK(hipMemcpyAsync(d.ig_cells, ig_cells, sizeof(int) * n_ig, hipMemcpyHostToDevice, s->stream));
        HIP_OK(hipMemcpyAsync(d.ig_donors, donors, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice, s->stream));
        HIP_OK(hipMemcpyAsync(d.ig_xy, image_xy, sizeof(double) * 2 * n_ig, hipMemcpyHostToDevice, s->stream));
        HIP_OK(hipMemcpyAsync(d.ig_ldl, ldlt, sizeof(double) * 10 * n_ig, hipMemcpyHostToDevice, s->stream));
        HIP_OK(hipMemcpyAsync(d.ig_transp, transp, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice, s->stream));
        HIP_OK(hipMemcpyAsync(d.ig_index, index.data(), sizeof(int) * (size_t)d.plane, hipMemcpyHostToDevice,
                              s->stream));
        for (int k = 0; k < NUM_SHAPES; ++k)
            HIP_OK(hipMemcpyAsync(s->tile_ib_shape[k], tiles[k].data(), tiles[k].size(), hipMemcpyHostToDevice,
                                  s->stream));
    }
    HIP_OK(hipStreamSynchronize(s->stream));
    d.n_ig = n_ig;
    d.ib_interp = interp;
    apply_shape(s, s->shape);
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
    {
        SetupLock lk(hip_setup_mutex());
        for (auto e : s->ev) (void)hipEventDestroy(e);
        s->ev.clear();
    }
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
        hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s->stream, scal(s).t(p0), t0);
        hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                           reinterpret_cast<unsigned long long*>(scal(s).halt()), 0ull);
        hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream,
                           reinterpret_cast<unsigned long long*>(scal(s).steps()), 0ull);
        HIP_OK(hipGetLastError());
        rc = launch_cfl(s->dev, s->cur, scal(s).acc(p0), s->stream);
        if (rc) return rc;
        int n = 0;
        if (!timed && nsteps >= 4) {
            // GRAPH_STEPS steps per graph (two for shorter advances)
            const int gsteps = nsteps >= 4 * GRAPH_STEPS ? GRAPH_STEPS : 2;
            GraphKey key;
            key.valid = true;
            key.use_cfl = true;
            key.cfl = cfl;
            key.t_end = t_end;
            key.step_parity = s->step_parity;
            key.stage_parity = s->parity;
            key.steps = gsteps;
            const bool same = s->graph_key.valid && s->graph_key.use_cfl == key.use_cfl &&
                              s->graph_key.cfl == key.cfl && s->graph_key.t_end == key.t_end &&
                              s->graph_key.step_parity == key.step_parity &&
                              s->graph_key.stage_parity == key.stage_parity && s->graph_key.steps == key.steps;
            const int counter0 = s->counter;
            if (!same) {
                SetupLock lk(hip_setup_mutex());
                drop_graph(s);
                HIP_OK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                int crc = 0;
                for (int k = 0; k < gsteps && !crc; ++k) crc = launch_step(s, false, true, cfl, t_end);
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
            // the captured launches left step/stage parities and planes where
            // gsteps real steps leave them.  Every HALT_POLL steps the halt flag
            // is read back: once the end time is reached (or the bound is not
            // finite) the remaining steps would do nothing, so they are not launched.
            for (; n + gsteps <= nsteps; n += gsteps) {
                if (n > 0 && n % HALT_POLL == 0) {
                    bool halted = false;
                    if ((rc = read_halt(s, &halted))) return rc;
                    if (halted) break;
                }
                HIP_OK(hipGraphLaunch(s->graph_exec, s->stream));
                s->counter += gsteps * s->cfg.rk;
            }
            s->cur = 0;
            if (n + gsteps <= nsteps) n = nsteps;   // stopped at a halt
        }
        for (; n < nsteps; ++n) {
            if (n > 0 && n % HALT_POLL == 0) {
                bool halted = false;
                if ((rc = read_halt(s, &halted))) return rc;
                if (halted) break;
            }
            rc = launch_step(s, timed, true, cfl, t_end);
            if (rc) return rc;
        }
        int ints[2] = {0, 0};
        HIP_OK(hipMemcpyAsync(ints, scal(s).halt(), sizeof(ints), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipMemcpyAsync(&done, scal(s).steps(), sizeof(int), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipMemcpyAsync(&t, scal(s).t(s->step_parity), sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
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
    int rc = launch_cfl(s->dev, s->cur, scal(s).acc_tmp(), s->stream);
    if (rc) return rc;
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, scal(s).acc_tmp(), sizeof(bits), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    double smax;
    std::memcpy(&smax, &bits, sizeof(smax));
    if (!(smax > 0.0) || !(smax < 1.0e308)) return CFS_E_NONFINITE;
    *dt_out = cfl / smax;
    return CFS_OK;
}

int32_t cfs_checksums(cfs_solver* s, double out[4]) {
    if (!s || !out) return CFS_E_ARG;
    const double* src = s->dev.set(s->cur);
    long long n = s->dev.plane, stride = s->dev.plane;
    const uint8_t* type = s->dev.type;
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
    int rc = CFS_OK;
    for (int q = 0; q < 4; ++q) {
        out[q] = tmp[q];
        if (!std::isfinite(tmp[q])) rc = CFS_E_NONFINITE;   // a NaN or Inf somewhere in the state
    }
    return rc;
}

}  // extern "C"

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mutex;
double* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up once per device.
int noalloc_scratch(double** out) {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return CFS_E_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(noalloc_mutex);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return CFS_E_HIP;
        noalloc_ptr[dev] = static_cast<double*>(p);
    }
    *out = noalloc_ptr[dev];
    return CFS_OK;
}

}  // namespace

extern "C" {

int32_t cfs_checksums_device_noalloc(const double* dU, int64_t cells, const uint8_t* d_type, double* d_out,
                                     void* hip_stream) {
    if (!dU || !d_out || cells < 1) return CFS_E_ARG;
    if (cells > CFS_NOALLOC_MAX_CELLS) return CFS_E_LIMIT;
    double* base = nullptr;
    const int rc = noalloc_scratch(&base);
    if (rc) return rc;
    const hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    // the tree of cfs_checksums, level by level; the last level (one sum per
    // variable) goes to d_out
    double* bufs[2] = {base, base + NOALLOC_L1};
    const double* src = dU;
    long long n = cells, stride = cells;
    const uint8_t* type = d_type;
    int which = 0;
    while (true) {
        const long long nb = (n + 255) / 256;
        double* dst = nb == 1 ? d_out : bufs[which];
        hipLaunchKernelGGL(tree_sum_kernel, dim3((unsigned)nb, 4), dim3(256), 0, stream, src, n, stride, type, dst,
                           nb);
        HIP_OK(hipGetLastError());
        if (nb == 1) break;
        src = dst;
        stride = nb;
        n = nb;
        type = nullptr;
        which ^= 1;
    }
    return CFS_OK;
}

int32_t cfs_calibrate_copy(cfs_solver* s, int32_t repeats) {
    if (!s || repeats < 0) return CFS_E_ARG;
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;   // spare sets hold RK stages mid-step
    const double2* src = reinterpret_cast<const double2*>(s->dev.set(s->cur));
    double2* dst = reinterpret_cast<double2*>(s->dev.set((s->cur + 1) % 3));
    const long long n = 2 * s->dev.plane;   // 4*plane doubles = 2*plane double2
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
    HIP_OK(hipMemcpyAsync(&v, s->dev.exact + (1 - s->parity), sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    *n = v;
    return CFS_OK;
}

int32_t cfs_stage_counter(const cfs_solver* s) { return s ? s->counter : -1; }

int32_t cfs_last_rhs(cfs_solver* s, double* R) {
    if (!s || !R) return CFS_E_ARG;
    if (s->counter <= 1) return CFS_E_STATE;
    const cfs_config& c = s->cfg;
    const size_t bytes = sizeof(double) * 4 * (size_t)s->dev.plane;
    {
        SetupLock lk(hip_setup_mutex());
        if (!s->rhs && hipMalloc(&s->rhs, bytes + 2 * sizeof(int)) != hipSuccess) {
            s->rhs = nullptr;
            return CFS_E_OOM;
        }
    }
    // the last stage's input plane-set is intact after it (every stage writes
    // another set): evaluate its RHS again, with its immersed-ghost fill
    StageArgs a;
    fill_stage_args(a, s->dev, rk_substage(c.rk, s->counter - 1), s->step_parity, 0, StageTime{});
    a.form = 3;
    a.out = s->rhs;
    a.u0 = a.u1 = nullptr;
    int* cnt = reinterpret_cast<int*>(s->rhs + 4 * (size_t)s->dev.plane);
    a.exact_count = cnt;
    a.exact_other = cnt + 1;
    dispatch_stage(c, s->shape, stage_grid(s), s->stream, a);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(R, s->rhs, bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return CFS_OK;
}

int32_t cfs_tile_shape_info(int32_t shape, int32_t* ti, int32_t* tj, int32_t* threads) {
    if (shape < 0 || shape >= NUM_SHAPES) return CFS_E_ARG;
    const ShapeDims d = shape_dims(shape);
    if (ti) *ti = d.ti;
    if (tj) *tj = d.tj;
    if (threads) *threads = d.nt;
    return CFS_OK;
}

int32_t cfs_set_tile_shape(cfs_solver* s, int32_t shape) {
    if (!s || shape < 0 || shape >= NUM_SHAPES) return CFS_E_ARG;
    if (s->cfg.viscous && shape != 0) return CFS_E_UNSUPPORTED;
    HIP_OK(hipStreamSynchronize(s->stream));
    if (shape != s->shape) drop_graph(s);
    apply_shape(s, shape);
    return CFS_OK;
}

int32_t cfs_get_tile_shape(const cfs_solver* s, int32_t* shape) {
    if (!s || !shape) return CFS_E_ARG;
    *shape = s->shape;
    return CFS_OK;
}

int32_t cfs_autotune(cfs_solver* s, int32_t reps, int32_t* best, double* ms) {
    if (!s || reps < 1 || reps > 10000) return CFS_E_ARG;
    // the timed stage writes the spare plane-set 1, which holds no stage data
    // only at a step boundary
    if (rk_substage(s->cfg.rk, s->counter) != 0) return CFS_E_STATE;
    const int nshape = s->cfg.viscous ? 1 : NUM_SHAPES;
    const int shape0 = s->shape;
    int* cnt = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CFS_OK;
    double t[NUM_SHAPES];
    int k_best = 0;
    {
        SetupLock lk(hip_setup_mutex());
        if (hipMalloc(&cnt, 2 * sizeof(int)) != hipSuccess) return CFS_E_OOM;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = CFS_E_HIP;
    }
    if (!rc) rc = set_dt_device(s, 0.0);
    // sub-stage 0 of the current state, fixed dt = 0, into plane-set 1: the
    // stage kernel's whole work (tile load with the boundary and IB fills,
    // faces, update, stores) without changing the solve.  Two passes over the
    // shapes, each shape's faster one kept: the first shape timed on an idle
    // GPU reads slow while the clocks ramp up (5-7% on a 0.25 ms stage).
    for (int pass = 0; pass < 2 && !rc; ++pass) {
        for (int k = 0; k < nshape && !rc; ++k) {
            apply_shape(s, k);
            StageArgs a;
            fill_stage_args(a, s->dev, 0, s->step_parity, 0, StageTime{});
            a.stage_early = stage_early(s) ? 1 : 0;
            a.exact_count = cnt;
            a.exact_other = cnt + 1;
            const dim3 grid = stage_grid(s);
            for (int w = 0; w < 2; ++w) dispatch_stage(s->cfg, k, grid, s->stream, a);
            if (hipEventRecord(e0, s->stream) != hipSuccess) { rc = CFS_E_HIP; break; }
            for (int n = 0; n < reps; ++n) dispatch_stage(s->cfg, k, grid, s->stream, a);
            float msk = 0.f;
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s->stream) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&msk, e0, e1) != hipSuccess) {
                rc = CFS_E_HIP;
                break;
            }
            const double tk = msk / reps;
            if (pass == 0 || tk < t[k]) t[k] = tk;
        }
    }
    for (int k = 0; k < nshape && !rc; ++k) {
        if (ms) ms[k] = t[k];
        if (k == 0 || t[k] < t[k_best]) k_best = k;
    }
    (void)hipStreamSynchronize(s->stream);
    {
        SetupLock lk(hip_setup_mutex());
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(cnt);
    }
    if (rc) {
        apply_shape(s, shape0);
        return rc;
    }
    for (int k = nshape; k < NUM_SHAPES && ms; ++k) ms[k] = -1.0;
    drop_graph(s);
    apply_shape(s, k_best);
    if (best) *best = k_best;
    return CFS_OK;
}

int32_t cfs_last_kernel_ms(const cfs_solver* s, double* ms) {
    if (!s || !ms) return CFS_E_ARG;
    *ms = s->last_kernel_ms;
    return CFS_OK;
}

}  // extern "C"
