// batch.hip -- batched solves (cfs_batch_* of include/cfs.h).
//
// A cfs_batch owns one device arena (problem planes, cell types, IB tables),
// a slot table of per-problem device scalars, the device array of per-problem
// stage arguments (2 step parities x 3 sub-stages per slot), the device list
// of (slot, tile) pairs the persistent stage kernel walks, one HIP stream and
// HIP graphs of two and of eight RK steps (consecutive replays leave a few
// microseconds between them).  The graphs' kernel arguments are only device
// pointers and the launch shape, so admitting and retiring problems (which
// rewrite the device arrays between replays) never invalidates them.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/cfs.h"
#include "host.hpp"
#include "arena.h"
#include "serve_lane.h"
#include "kernels.hpp"

using namespace cfs;

namespace {

constexpr size_t ALIGN = 256;
inline size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// Arena layout of one problem: planes | type | ig_index | tile_ib | IB tables.
struct Layout {
    size_t planes = 0, type = 0, ig_index = 0, tile_ib = 0;
    size_t ig_cells = 0, donors = 0, xy = 0, ldl = 0, transp = 0, total = 0;
};

Layout layout_of(const cfs_config& c, int n_ig, bool has_ib) {
    const size_t plane = (size_t)c.nx * c.ny;
    const size_t tiles = (size_t)tiles_for(c, 0);   // the smallest tile shape: the most tiles
    Layout L;
    size_t o = 0;
    L.planes = o;
    o += align_up(sizeof(double) * 12 * plane);
    if (has_ib) {
        L.type = o;
        o += align_up(plane);
        L.ig_index = o;
        o += align_up(sizeof(int) * plane);
        L.tile_ib = o;
        o += align_up(tiles);
        L.ig_cells = o;
        o += align_up(sizeof(int) * (size_t)n_ig);
        L.donors = o;
        o += align_up(sizeof(int) * 4 * (size_t)n_ig);
        L.xy = o;
        o += align_up(sizeof(double) * 2 * (size_t)n_ig);
        L.ldl = o;
        o += align_up(sizeof(double) * 10 * (size_t)n_ig);
        L.transp = o;
        o += align_up(sizeof(int) * 4 * (size_t)n_ig);
    }
    L.total = o;
    return L;
}

struct Slot {
    bool used = false;
    cfs_config cfg{};
    ProblemDev dev;
    size_t off = 0, bytes = 0;
    int ntiles = 0;
};

constexpr int DESCS_PER_SLOT = 6;   // 2 step parities x 3 sub-stages

}  // namespace

struct cfs_batch {
    cfs_batch_config bc{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    char* arena = nullptr;
    Arena alloc;
    std::vector<Slot> slots;
    double* scal = nullptr;           // [max_problems][SCAL_DOUBLES]
    int* exact = nullptr;             // [max_problems][2]
    StageArgs* descs = nullptr;       // [max_problems][2][3]
    StageArgs* descs_host = nullptr;  // pinned staging for one slot
    int2* tiles = nullptr;            // (slot, tile) list
    int* sched = nullptr;             // tile counters of the batched kernel (BATCH_SCHED_INTS)
    int* ntiles = nullptr;
    long long tiles_cap = 0;
    int2* tiles_host = nullptr;       // pinned staging of the list
    double* scal_host = nullptr;      // pinned copy for polling
    int grid = 0;
    int resident = 0;                 // workgroups of the stage kernel the CUs hold at once
    int device = 0;                   // the device current at creation
    int parity = 0;                   // step parity of the next step
    bool dirty = false;               // tile list to rebuild before the next step
    hipGraph_t graph = nullptr;            // two RK steps
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph8 = nullptr;           // eight RK steps
    hipGraphExec_t graph8_exec = nullptr;
};

namespace {

template <typename R, bool LIMIT, typename Sh>
int launch_batch_shape(cfs_batch* b, int parity, int sub) {
    hipLaunchKernelGGL((batch_stage_kernel<R, LIMIT, Sh>), dim3((unsigned)b->grid), dim3(Sh::NT), 0, b->stream,
                       b->descs, b->tiles, b->ntiles, parity, sub, b->sched);
    HIP_OK(hipGetLastError());
    return CFS_OK;
}

template <typename R, bool LIMIT>
int launch_batch_stage(cfs_batch* b, int parity, int sub) {
    const int k = b->bc.tile_shape;
    return k == 1 ? launch_batch_shape<R, LIMIT, Shape1>(b, parity, sub)
                  : k == 2 ? launch_batch_shape<R, LIMIT, Shape2>(b, parity, sub)
                  : k == 3 ? launch_batch_shape<R, LIMIT, Shape3>(b, parity, sub)
                           : launch_batch_shape<R, LIMIT, Shape0>(b, parity, sub);
}

template <typename R, bool LIMIT, typename Sh>
hipError_t blocks_per_cu(int* per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, batch_stage_kernel<R, LIMIT, Sh>, Sh::NT, 0);
}

template <typename R, bool LIMIT>
hipError_t blocks_per_cu_shape(int k, int* per_cu) {
    return k == 1 ? blocks_per_cu<R, LIMIT, Shape1>(per_cu)
                  : k == 2 ? blocks_per_cu<R, LIMIT, Shape2>(per_cu)
                  : k == 3 ? blocks_per_cu<R, LIMIT, Shape3>(per_cu) : blocks_per_cu<R, LIMIT, Shape0>(per_cu);
}

int launch_batch_step(cfs_batch* b, int parity) {
    const bool fp64 = b->bc.precision == CFS_PREC_FP64, lim = b->bc.limiter != 0;
    for (int k = 0; k < b->bc.rk; ++k) {
        const int rc = fp64 ? (lim ? launch_batch_stage<double, true>(b, parity, k)
                                   : launch_batch_stage<double, false>(b, parity, k))
                            : (lim ? launch_batch_stage<float, true>(b, parity, k)
                                   : launch_batch_stage<float, false>(b, parity, k));
        if (rc) return rc;
    }
    return CFS_OK;
}

// (Re)writes the device tile list: every used slot's tiles, slot by slot, in
// tile order (neighbouring tiles stay adjacent in the list).
int rebuild_tiles(cfs_batch* b) {
    long long n = 0;
    for (const Slot& s : b->slots)
        if (s.used) n += s.ntiles;
    if (n > b->tiles_cap) return CFS_E_FULL;   // checked before the staging buffer is written
    n = 0;
    for (int p = 0; p < (int)b->slots.size(); ++p) {
        const Slot& s = b->slots[p];
        if (!s.used) continue;
        for (int t = 0; t < s.ntiles; ++t) b->tiles_host[n++] = make_int2(p, t);
    }
    if (n) HIP_OK(hipMemcpyAsync(b->tiles, b->tiles_host, sizeof(int2) * n, hipMemcpyHostToDevice, b->stream));
    const int nn = (int)n;
    HIP_OK(hipMemcpyAsync(b->ntiles, &nn, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));   // staging buffers are reused
    b->dirty = false;
    return CFS_OK;
}

void drop_graph(cfs_batch* b) {
    SetupLock lk(hip_setup_mutex());
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->graph8_exec) (void)hipGraphExecDestroy(b->graph8_exec);
    if (b->graph8) (void)hipGraphDestroy(b->graph8);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph8_exec = nullptr;
    b->graph8 = nullptr;
}

long long tiles_of(const cfs_batch* b, const cfs_config& c) { return tiles_for(c, b->bc.tile_shape); }

}  // namespace

extern "C" {

int64_t cfs_batch_footprint(const cfs_config* cfg, int32_t n_ig) {
    if (check_config(cfg) || n_ig < 0) return -1;
    return (int64_t)layout_of(*cfg, n_ig, n_ig > 0).total;
}

int32_t cfs_batch_create(const cfs_batch_config* bc, void* hip_stream, cfs_batch** out) {
    if (!out) return CFS_E_ARG;
    *out = nullptr;
    if (!bc || (bc->rk != 2 && bc->rk != 3)) return CFS_E_ARG;
    if (bc->precision != CFS_PREC_FP64 && bc->precision != CFS_PREC_FP32_BULK) return CFS_E_ARG;
    if (bc->max_problems < 1 || bc->max_problems > (1 << 20) || bc->arena_bytes < 1 || bc->workgroups < 0)
        return CFS_E_ARG;
    if (bc->limiter != 0 && bc->limiter != 1) return CFS_E_ARG;
    if (bc->tile_shape < 0 || bc->tile_shape >= NUM_SHAPES) return CFS_E_ARG;
    SetupLock lk(hip_setup_mutex());
    cfs_batch* b = new (std::nothrow) cfs_batch();
    if (!b) return CFS_E_OOM;
    b->bc = *bc;
    b->slots.resize((size_t)bc->max_problems);
    if (hip_stream) {
        b->stream = (hipStream_t)hip_stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return CFS_E_HIP; }
        b->own_stream = true;
    }
    // persistent grid, sized from the stage kernel's occupancy
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) {
        const bool fp64 = bc->precision == CFS_PREC_FP64, lim = bc->limiter != 0;
        const int k = bc->tile_shape;
        e = fp64 ? (lim ? blocks_per_cu_shape<double, true>(k, &per_cu) : blocks_per_cu_shape<double, false>(k, &per_cu))
                 : (lim ? blocks_per_cu_shape<float, true>(k, &per_cu) : blocks_per_cu_shape<float, false>(k, &per_cu));
    }
    if (e != hipSuccess) { cfs_batch_destroy(b); return CFS_E_HIP; }
    // the workgroups the CUs hold at once: they take the tiles in list order
    // from a counter per XCD range, so tiles of uneven cost (body, limiter,
    // halted problems) even out without more workgroups than fit
    b->device = dev;
    b->resident = cus * (per_cu > 0 ? per_cu : 1);
    int g = bc->workgroups > 0 ? bc->workgroups : b->resident;
    b->grid = ((g + 7) / 8) * 8;
    // tile list capacity: twice the full tiles an arena of problem planes
    // holds (96 B per cell) plus 16 ragged tiles per slot; admission past it
    // is CFS_E_FULL
    const ShapeDims sd = shape_dims(bc->tile_shape);
    b->tiles_cap = bc->arena_bytes / (96ll * sd.ti * sd.tj / 2) + 16ll * bc->max_problems;
    const size_t np = (size_t)bc->max_problems;
    if (hipMalloc(&b->arena, (size_t)bc->arena_bytes) != hipSuccess ||
        hipMalloc(&b->scal, sizeof(double) * SCAL_DOUBLES * np) != hipSuccess ||
        hipMalloc(&b->exact, sizeof(int) * 2 * np) != hipSuccess ||
        hipMalloc(&b->descs, sizeof(StageArgs) * DESCS_PER_SLOT * np) != hipSuccess ||
        hipMalloc(&b->tiles, sizeof(int2) * (size_t)b->tiles_cap) != hipSuccess ||
        hipMalloc(&b->ntiles, sizeof(int)) != hipSuccess ||
        hipMalloc(&b->sched, sizeof(int) * BATCH_SCHED_INTS) != hipSuccess ||
        hipHostMalloc(&b->descs_host, sizeof(StageArgs) * DESCS_PER_SLOT, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&b->tiles_host, sizeof(int2) * (size_t)b->tiles_cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&b->scal_host, sizeof(double) * SCAL_DOUBLES * np, hipHostMallocDefault) != hipSuccess) {
        cfs_batch_destroy(b);
        return CFS_E_OOM;
    }
    b->alloc.reset((size_t)bc->arena_bytes);
    if (hipMemsetAsync(b->scal, 0, sizeof(double) * SCAL_DOUBLES * np, b->stream) != hipSuccess ||
        hipMemsetAsync(b->exact, 0, sizeof(int) * 2 * np, b->stream) != hipSuccess ||
        hipMemsetAsync(b->ntiles, 0, sizeof(int), b->stream) != hipSuccess ||
        hipMemsetAsync(b->sched, 0, sizeof(int) * BATCH_SCHED_INTS, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        cfs_batch_destroy(b);
        return CFS_E_HIP;
    }
    *out = b;
    return CFS_OK;
}

void cfs_batch_destroy(cfs_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    SetupLock lk(hip_setup_mutex());
    drop_graph(b);
    (void)hipFree(b->arena);
    (void)hipFree(b->scal);
    (void)hipFree(b->exact);
    (void)hipFree(b->descs);
    (void)hipFree(b->tiles);
    (void)hipFree(b->ntiles);
    (void)hipFree(b->sched);
    (void)hipHostFree(b->descs_host);
    (void)hipHostFree(b->tiles_host);
    (void)hipHostFree(b->scal_host);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int32_t batch_check_problem(const cfs_batch* b, const cfs_config* cfg, const cfs_immersed* ib, double cfl,
                            double t_end) {
    if (!b) return CFS_E_ARG;
    int rc = check_config(cfg);
    if (rc) return rc;
    if (cfg->rk != b->bc.rk || cfg->precision != b->bc.precision) return CFS_E_ARG;
    if (cfg->viscous) return CFS_E_UNSUPPORTED;
    if (cfg->positivity == 2 && !b->bc.limiter) return CFS_E_ARG;   // the batch runs the kernel without the limiter
    if (!(cfl > 0.0) || !(cfl < 1e300) || t_end != t_end) return CFS_E_ARG;
    if (ib != nullptr && (ib->n_ig <= 0 || !ib->image_xy || !ib->coef)) return CFS_E_ARG;
    return CFS_OK;
}

int64_t batch_problem_tiles(const cfs_batch* b, const cfs_config* cfg) { return tiles_of(b, *cfg); }

int64_t batch_resident_workgroups(const cfs_batch* b) { return b->resident; }

int32_t batch_device_free_bytes(const cfs_batch* b, int64_t* bytes) {
    if (hipSetDevice(b->device) != hipSuccess) return CFS_E_HIP;
    size_t free_b = 0, total = 0;
    if (hipMemGetInfo(&free_b, &total) != hipSuccess) return CFS_E_HIP;
    *bytes = (int64_t)free_b;
    return CFS_OK;
}

int32_t batch_make_device_current(const cfs_batch* b) {
    return hipSetDevice(b->device) == hipSuccess ? CFS_OK : CFS_E_HIP;
}

int32_t cfs_batch_admit(cfs_batch* b, const cfs_config* cfg, const double* U, const cfs_immersed* ib, double cfl,
                        double t_end, int32_t* slot) {
    if (!b || !U || !slot) return CFS_E_ARG;
    int rc = batch_check_problem(b, cfg, ib, cfl, t_end);
    if (rc) return rc;
    const bool has_ib = ib != nullptr;
    const int n_ig = has_ib ? ib->n_ig : 0;
    const ShapeDims sd = shape_dims(b->bc.tile_shape);
    const int tiles_i = (cfg->nx + sd.ti - 1) / sd.ti, tiles_j = (cfg->ny + sd.tj - 1) / sd.tj;
    std::vector<int> index;
    std::vector<uint8_t> tflags;
    if (has_ib) {
        rc = build_ib_index(*cfg, b->bc.tile_shape, ib->interp, ib->cell_type, n_ig, ib->ig_cells, ib->donors,
                            ib->transp, index, tflags);
        if (rc) return rc;
    }
    int p = -1;
    for (int k = 0; k < (int)b->slots.size(); ++k)
        if (!b->slots[k].used) { p = k; break; }
    if (p < 0) return CFS_E_FULL;
    long long used_tiles = 0;
    for (const Slot& s : b->slots)
        if (s.used) used_tiles += s.ntiles;
    if (used_tiles + tiles_of(b, *cfg) > b->tiles_cap) return CFS_E_FULL;
    const Layout L = layout_of(*cfg, n_ig, has_ib);
    size_t off = 0;
    if (!b->alloc.alloc(L.total, &off)) return CFS_E_FULL;

    Slot& s = b->slots[p];
    s = Slot{};
    s.cfg = *cfg;
    s.off = off;
    s.bytes = L.total;
    s.ntiles = (int)tiles_of(b, *cfg);
    ProblemDev& d = s.dev;
    char* base = b->arena + off;
    d.cfg = &s.cfg;
    d.plane = (long long)cfg->nx * cfg->ny;
    d.tiles_i = tiles_i;
    d.tiles_j = tiles_j;
    d.planes = reinterpret_cast<double*>(base + L.planes);
    d.scal = b->scal + (size_t)SCAL_DOUBLES * p;
    d.exact = b->exact + 2 * p;
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(b->stream);
        b->alloc.release(s.off, s.bytes);
        s = Slot{};
        return code;
    };
    // wait for queued steps: they read the arrays rewritten below
    if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(CFS_E_HIP);
    if (has_ib) {
        d.type = reinterpret_cast<uint8_t*>(base + L.type);
        if (hipMemcpyAsync(d.type, ib->cell_type, (size_t)d.plane, hipMemcpyHostToDevice, b->stream) != hipSuccess)
            return fail(CFS_E_HIP);
        if (n_ig > 0) {
            d.n_ig = n_ig;
            d.ib_interp = ib->interp;
            d.ig_index = reinterpret_cast<int*>(base + L.ig_index);
            d.tile_ib = reinterpret_cast<uint8_t*>(base + L.tile_ib);
            d.ig_cells = reinterpret_cast<int*>(base + L.ig_cells);
            d.ig_donors = reinterpret_cast<int*>(base + L.donors);
            d.ig_xy = reinterpret_cast<double*>(base + L.xy);
            d.ig_ldl = reinterpret_cast<double*>(base + L.ldl);
            d.ig_transp = reinterpret_cast<int*>(base + L.transp);
            if (hipMemcpyAsync(d.ig_index, index.data(), sizeof(int) * (size_t)d.plane, hipMemcpyHostToDevice,
                               b->stream) != hipSuccess ||
                hipMemcpyAsync(d.tile_ib, tflags.data(), tflags.size(), hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(d.ig_cells, ib->ig_cells, sizeof(int) * n_ig, hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(d.ig_donors, ib->donors, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(d.ig_xy, ib->image_xy, sizeof(double) * 2 * n_ig, hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(d.ig_ldl, ib->coef, sizeof(double) * 10 * n_ig, hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(d.ig_transp, ib->transp, sizeof(int) * 4 * n_ig, hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess)
                return fail(CFS_E_HIP);
        }
    }
    // state, immersed-ghost prologue fill, scalars, bound of the entering state
    if (hipMemcpyAsync(d.set(0), U, sizeof(double) * 4 * (size_t)d.plane, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemsetAsync(d.scal, 0, sizeof(double) * SCAL_DOUBLES, b->stream) != hipSuccess ||
        hipMemsetAsync(d.exact, 0, sizeof(int) * 2, b->stream) != hipSuccess)
        return fail(CFS_E_HIP);
    if (launch_ib_fill(d, d.set(0), b->stream) || launch_cfl(d, 0, d.s().acc(b->parity), b->stream))
        return fail(CFS_E_HIP);
    // stage arguments for both step parities
    for (int q = 0; q < 2; ++q)
        for (int k = 0; k < 3; ++k) {
            StageArgs& a = b->descs_host[q * 3 + k];
            if (k < cfg->rk) fill_stage_args(a, d, k, q, k & 1, cfl_stage_time(cfg->rk, k, cfl, t_end));
            else a = StageArgs{};
            // the batch's resident state as a whole, not one problem's, decides
            // whether the update streams past the last-level cache
            a.stream_nt = b->bc.arena_bytes > LLC_BYTES ? 1 : 0;
        }
    if (hipMemcpyAsync(b->descs + (size_t)DESCS_PER_SLOT * p, b->descs_host, sizeof(StageArgs) * DESCS_PER_SLOT,
                       hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return fail(CFS_E_HIP);
    s.used = true;
    b->dirty = true;
    *slot = p;
    return CFS_OK;
}

// A graph of `steps` RK steps (even: the step parities alternate), starting at
// parity 0.
static int capture_steps(cfs_batch* b, int steps, hipGraph_t* graph, hipGraphExec_t* exec) {
    SetupLock lk(hip_setup_mutex());
    HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int crc = 0;
    for (int k = 0; k < steps && !crc; ++k) crc = launch_batch_step(b, k & 1);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (crc) {
        if (g) (void)hipGraphDestroy(g);
        return crc;
    }
    HIP_OK(ec);
    *graph = g;
    HIP_OK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return CFS_OK;
}

int32_t cfs_batch_step(cfs_batch* b, int32_t nsteps) {
    if (!b || nsteps < 0) return CFS_E_ARG;
    int rc;
    if (b->dirty && (rc = rebuild_tiles(b))) return rc;
    int n = nsteps;
    if (b->parity == 1 && n > 0) {
        if ((rc = launch_batch_step(b, 1))) return rc;
        b->parity = 0;
        --n;
    }
    if (n >= 8) {
        if (!b->graph8_exec && (rc = capture_steps(b, 8, &b->graph8, &b->graph8_exec))) return rc;
        for (; n >= 8; n -= 8) HIP_OK(hipGraphLaunch(b->graph8_exec, b->stream));
    }
    if (n >= 2) {
        if (!b->graph_exec && (rc = capture_steps(b, 2, &b->graph, &b->graph_exec))) return rc;
        for (; n >= 2; n -= 2) HIP_OK(hipGraphLaunch(b->graph_exec, b->stream));
    }
    if (n == 1) {
        if ((rc = launch_batch_step(b, 0))) return rc;
        b->parity = 1;
    }
    return CFS_OK;
}

int32_t cfs_batch_poll(cfs_batch* b, cfs_batch_info* info, int32_t cap, int32_t* n) {
    if (!b || !n || cap < 0 || (cap > 0 && !info)) return CFS_E_ARG;
    const size_t np = b->slots.size();
    HIP_OK(hipMemcpyAsync(b->scal_host, b->scal, sizeof(double) * SCAL_DOUBLES * np, hipMemcpyDeviceToHost,
                          b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    int m = 0;
    for (size_t p = 0; p < np; ++p) {
        const Slot& s = b->slots[p];
        if (!s.used) continue;
        if (m < cap) {
            const ScalPtrs sp{b->scal_host + SCAL_DOUBLES * p};
            cfs_batch_info& r = info[m];
            r.slot = (int32_t)p;
            r.state = *sp.status() ? CFS_SLOT_FAILED : (*sp.halt() ? CFS_SLOT_DONE : CFS_SLOT_RUNNING);
            r.steps = *sp.steps();
            r.nx = s.cfg.nx;
            r.ny = s.cfg.ny;
            r.t = *sp.t(b->parity);
        }
        ++m;
    }
    *n = m;
    return CFS_OK;
}

int32_t cfs_batch_retire(cfs_batch* b, int32_t slot, double* U) {
    if (!b || slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot].used) return CFS_E_ARG;
    Slot& s = b->slots[slot];
    if (U)
        HIP_OK(hipMemcpyAsync(U, s.dev.set(0), sizeof(double) * 4 * (size_t)s.dev.plane, hipMemcpyDeviceToHost,
                              b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    b->alloc.release(s.off, s.bytes);
    s = Slot{};
    b->dirty = true;
    return CFS_OK;
}

int32_t cfs_batch_capacity(const cfs_batch* b, int64_t* arena_bytes, int32_t* max_problems) {
    if (!b || !arena_bytes || !max_problems) return CFS_E_ARG;
    *arena_bytes = b->bc.arena_bytes;
    *max_problems = b->bc.max_problems;
    return CFS_OK;
}

int32_t cfs_batch_autotune(const cfs_batch_config* bc, const cfs_request* sample, int32_t n, int32_t steps,
                           int32_t* best, double* ms_per_step) {
    if (!bc || !sample || n < 1 || steps < 1 || steps > 100000) return CFS_E_ARG;
    int k_best = 0;
    double t_best = 0.0;
    for (int k = 0; k < NUM_SHAPES; ++k) {
        cfs_batch_config bck = *bc;
        bck.tile_shape = k;
        cfs_batch* b = nullptr;
        int rc = cfs_batch_create(&bck, nullptr, &b);
        if (rc) return rc;
        for (int r = 0; r < n && !rc; ++r) {
            int32_t slot = 0;
            rc = cfs_batch_admit(b, sample[r].cfg, sample[r].U, sample[r].ib, sample[r].cfl, 0.0, &slot);
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        float ms = 0.f;
        {
            SetupLock lk(hip_setup_mutex());
            if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = CFS_E_HIP;
        }
        if (!rc) rc = cfs_batch_step(b, 2);
        if (!rc && hipEventRecord(e0, b->stream) != hipSuccess) rc = CFS_E_HIP;
        if (!rc) rc = cfs_batch_step(b, steps);
        if (!rc && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = CFS_E_HIP;
        {
            SetupLock lk(hip_setup_mutex());
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
        cfs_batch_destroy(b);
        if (rc) return rc;
        const double t = (double)ms / steps;
        if (ms_per_step) ms_per_step[k] = t;
        if (k == 0 || t < t_best) {
            k_best = k;
            t_best = t;
        }
    }
    if (best) *best = k_best;
    return CFS_OK;
}

int32_t cfs_batch_free_bytes(const cfs_batch* b, int64_t* bytes) {
    if (!b || !bytes) return CFS_E_ARG;
    *bytes = (int64_t)b->alloc.free_bytes();
    return CFS_OK;
}

}  // extern "C"
