/*
 * cfs.h -- C-ABI of the MI355X (gfx950) per-stage hot path of the 2D
 * compressible-flow solver: WENO5 reconstruction, HLL flux, TVD-RK stage
 * update, physical-boundary fill, ghost-cell immersed-boundary fill and the
 * CFL time-step reduction.
 *
 * The reference binds this path through C++ virtual "plugins" held as global
 * singletons and called once per cell / per face from its stage loop
 * (Run.cpp:153-204).  That granularity cannot cross a device boundary, so this
 * ABI sits one level up, at stage / step granularity.  Each entry point below
 * names the reference interface whose work it takes over.
 *
 * Conventions
 *   - plain C, no exceptions cross the boundary; every call returns a
 *     cfs_status (0 = success, negative = error), cfs_status_string() names it;
 *   - the handle owns every device buffer and one HIP stream; host arrays are
 *     caller-owned and copied;
 *   - a handle is not re-entrant; distinct handles may be used concurrently;
 *   - state arrays are FP64, structure-of-arrays [4][nx][ny], x-major
 *     (index i*ny + j), variables (rho, rho*u, rho*v, E) -- the layout the
 *     reference writes its snapshots in (source/io/Output.cpp:60-74).
 */
#ifndef CFS_H
#define CFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFS_ABI_VERSION 3

typedef enum cfs_status {
    CFS_OK = 0,
    CFS_E_ARG = -1,         /* invalid argument (null pointer, bad size, bad enum) */
    CFS_E_HIP = -2,         /* a HIP runtime call failed */
    CFS_E_OOM = -3,         /* device allocation failed */
    CFS_E_STATE = -4,       /* call out of sequence (e.g. stage counter mismatch) */
    CFS_E_NONFINITE = -5,   /* a reduction met NaN/Inf (e.g. negative pressure) */
    CFS_E_UNSUPPORTED = -6, /* configuration not supported by this build */
    CFS_E_FULL = -7,        /* batch: no free slot or arena space for the problem now */
    CFS_E_LIMIT = -8        /* input larger than the fixed device scratch of a no-allocation entry point */
} cfs_status;

/* Physical boundary kinds.  NEUMANN and DIRICHLET follow
 * source/boundary/NeummanBoundary.cpp:31-45 and DirichletBoundary.cpp:32-46
 * (one ghost cell per boundary face); PERIODIC has no reference counterpart
 * (the reference's Cartesian topology is non-periodic,
 * source/parallel/CartesianParallel.cpp:7) and must be set on both sides of
 * an axis. */
enum { CFS_BC_NEUMANN = 0, CFS_BC_DIRICHLET = 1, CFS_BC_PERIODIC = 2 };
/* side order of the reference's boundary vector (Reimman2D.cpp:41-51) */
enum { CFS_SIDE_LEFT = 0, CFS_SIDE_RIGHT = 1, CFS_SIDE_BOTTOM = 2, CFS_SIDE_TOP = 3 };
/* arithmetic of the reconstruct/flux/RHS part of a stage; the conserved state,
 * the RK combination, the CFL reduction, checksums and the immersed-boundary
 * fill are FP64 in both. */
enum { CFS_PREC_FP64 = 0, CFS_PREC_FP32_BULK = 1 };
/* Immersed-ghost interpolation from the 4 donor cells.
 * REFERENCE: as ImagePoint::InterpolatePhase (ImagePoint.cpp:22-69) computes
 *   it: Eigen's LDLT of the donor matrix M (rows [x*y, x, y, 1]) reads only the
 *   lower triangle, so the system solved is sym(tril(M)) w = f, and the value
 *   is w0*x*y + w1*x + w2*y + w3 at the image point.  This is not a bilinear
 *   fit and can give negative densities: the reference's own cylinder case
 *   turns non-finite within a few stages at most resolutions.
 * BILINEAR_POSITIVE: the bilinear fit the reference intends, solved exactly
 *   in cell-local coordinates, used when the donor system is non-degenerate and
 *   all its weights are >= 0; otherwise inverse-squared-distance weights.  Both
 *   branches are convex combinations (admissible states stay admissible). */
enum { CFS_IB_REFERENCE = 0, CFS_IB_BILINEAR_POSITIVE = 1 };
/* cell classification of source/meshes/Cell.h:14 (CellType) */
enum { CFS_CELL_FLUID = 0, CFS_CELL_SOLID = 1, CFS_CELL_IMMERSED_GHOST = 2 };

typedef struct cfs_config {
    int32_t nx, ny;          /* cells along x and y (CartesianMesh::setnumberCells*); at least 3 each,
                                nx*ny < 2^31 (a periodic wrap of the 3-cell halo stays in one period) */
    double lx, ly;           /* domain lengths; dx = lx/nx (CartesianMesh.cpp:150-153) */
    double x0, y0;           /* domain origin; centre of (i,j) = ((i+0.5)*dx + x0, (j+0.5)*dy + y0) */
    double gamma;            /* ratio of specific heats (param.h:5) */
    int32_t bc[4];           /* CFS_BC_* per side, order L,R,B,T */
    double bc_state[4][4];   /* Dirichlet ghost state per side, conserved (rho, rho*u, rho*v, E) */
    int32_t rk;              /* 2 = TVD2rdRK (Run.cpp:47), 3 = TVD3rdRK as written */
    int32_t precision;       /* CFS_PREC_* */
    int32_t viscous;         /* 0 = Euler; 1 = Navier-Stokes terms (4th/6th-order central) */
    int32_t positivity;      /* 0 = reference semantics.  1 = a face state whose WENO
                                reconstruction is not admissible (rho > 0 and p > 0 fail)
                                is replaced by the adjacent cell's state (first order);
                                the reference turns such a face into NaN through
                                sqrt(gamma*p/rho) whenever p/rho < 0.  2 = as 1, plus
                                the positivity-preserving flux limiter (each face flux
                                blended with the Lax-Friedrichs flux so that the
                                forward-Euler update keeps rho, p > 0; Hu, Adams & Shu
                                2013); inviscid configurations only. */
    int32_t hll_bounds;      /* 0 = reference HLL outer wave-speed bounds
                                SL = min(mL - aL, u* - a*L), SR = max(mL + aL, u* + a*R)
                                with m the normal MOMENTUM and the left state in SR
                                (RiemmanSolverHLL.cpp:41-42); 1 = velocity bounds
                                SL = min(uL - aL, u* - a*L), SR = max(uR + aR, u* + a*R). */
    double mu, prandtl, ma_ref; /* constant viscosity, Prandtl number, reference Mach number */
} cfs_config;

typedef struct cfs_solver cfs_solver;

/* ---- library ----------------------------------------------------------- */
int32_t     cfs_abi_version(void);
/* sizeof(cfs_config) as compiled into the library (binding layout check) */
int32_t     cfs_config_size(void);
const char* cfs_status_string(int32_t status);
/* number of visible HIP devices (0 without a GPU; never an error) */
int32_t     cfs_device_count(void);

/* ---- solver lifetime ---------------------------------------------------- */
/* Allocates the device state planes for one grid.  Replaces the reference's
 * mesh/plugin construction for the hot path (Run::initialize, Run.cpp:10-90:
 * CartesianMesh::allocateCells/allocateCellInterfaces/BindCellInterfacesWithCells,
 * the EulerMod/Weno5th/RiemmanSolverHLL/TVD*RK singletons, the boundary
 * objects of Initialcondition::setBounbdaryCondition).
 * hip_stream may be NULL: the handle then creates its own stream. */
int32_t cfs_create(const cfs_config* cfg, void* hip_stream, cfs_solver** out);
void    cfs_destroy(cfs_solver* s);

/* Host <-> device state copies ([4][nx][ny] FP64).  set_state is allowed only
 * at a step boundary (the reference fills the initial condition before its
 * stage loop, Run.cpp:99-105).  get_state returns the current stage's state. */
int32_t cfs_set_state(cfs_solver* s, const double* U);
int32_t cfs_get_state(cfs_solver* s, double* U);
/* Same with device pointers (no host staging); the pointer must hold 4*nx*ny doubles. */
int32_t cfs_set_state_device(cfs_solver* s, const double* dU);
int32_t cfs_get_state_device(cfs_solver* s, double* dU);

/* ---- immersed boundary --------------------------------------------------- */
/* Host-side set-up of a circle immersed boundary, the work of
 * CartesianMesh::setImmersedBoundary (CartesianMesh.cpp:357-374):
 * Circle::Initialize (Circle.cpp:12-41), DistinguishWallandFluid (:43-51),
 * SetGhostcell (:53-103), SetImagepoint (:105-138) and
 * FindNearest4FluidCellsIndexed (:160-215), plus the factorisation that
 * ImagePoint::InterpolatePhase (ImagePoint.cpp:20-51) repeats every stage
 * (done once here: the geometry is fixed).
 * interp is CFS_IB_* and selects what coef holds.
 * Outputs (caller-allocated):
 *   cell_type [nx*ny]         CFS_CELL_* per cell
 *   ig_cells  [cap]           linear index i*ny+j of each immersed-ghost cell, i-major order
 *   donors    [cap*4]         linear indices of the 4 donor fluid cells (distance order)
 *   image_xy  [cap*2]         image-point coordinates
 *   coef      [cap*10]        REFERENCE: pivoted LDL^T of the donor matrix, 6
 *                             strict-lower L entries (10,20,21,30,31,32) then D[0..3];
 *                             BILINEAR_POSITIVE: the 4 donor weights, then zeros
 *   transp    [cap*4]         REFERENCE: pivot transpositions; otherwise 0,1,2,3
 * *n_ig receives the number of immersed-ghost cells; CFS_E_ARG if it exceeds
 * cap (then only *n_ig is meaningful) or if an image point has fewer than 4
 * fluid donors in its window. Needs no GPU. */
int32_t cfs_circle_setup(const cfs_config* cfg, double cx, double cy, double radius,
                         int32_t n_lagrangian, int32_t n_layers, int32_t interp, int32_t cap,
                         uint8_t* cell_type, int32_t* n_ig, int32_t* ig_cells,
                         int32_t* donors, double* image_xy, double* coef,
                         int32_t* transp);

/* cfs_circle_setup on the GPU (the device current on the calling thread):
 * the same outputs, bit for bit.  Kernels classify the cells of the body's
 * bounding box, mark the ghost layers, list the ghosts in i-major order and
 * compute each ghost's image point, donors and coefficients; a ghost whose
 * five nearest donor candidates include two at equal distance (where the
 * reference's std::sort decides their order) is finished on the host with that
 * sort, and *n_host (optional) receives how many were.  Same errors as
 * cfs_circle_setup, and CFS_E_HIP.  Reference: Circle.cpp:12-215 (the host
 * set-up it replaces at admission time). */
int32_t cfs_circle_setup_device(const cfs_config* cfg, double cx, double cy, double radius,
                                int32_t n_lagrangian, int32_t n_layers, int32_t interp, int32_t cap,
                                uint8_t* cell_type, int32_t* n_ig, int32_t* ig_cells,
                                int32_t* donors, double* image_xy, double* coef,
                                int32_t* transp, int32_t* n_host);

/* Uploads the immersed-boundary tables produced by cfs_circle_setup.  From
 * then on every stage first fills the immersed-ghost cells from their image
 * points (Run.cpp:156-162 -> Cell::comImmersedGhostPhase, Cell.cpp:172-181)
 * and updates fluid cells only (Run.cpp:193-200). */
int32_t cfs_set_immersed(cfs_solver* s, int32_t interp, const uint8_t* cell_type, int32_t n_ig,
                         const int32_t* ig_cells, const int32_t* donors,
                         const double* image_xy, const double* coef,
                         const int32_t* transp);

/* ---- the per-stage hot path ---------------------------------------------- */
/* One RK stage; stage_counter is the reference's global loop counter i
 * (Run.cpp:153, starting at 1) that TVD2rdRK/TVD3rdRK::advance key their
 * sub-stage on (TVD2rdRK.cpp:11, TVD3rdRK.cpp:15).  It must equal the
 * handle's next counter (CFS_E_STATE otherwise).  Replaces one iteration of
 * Run.cpp:153-204: IB fill, (halo), WENO over all faces, HLL over all faces,
 * RK update of fluid cells, physical-boundary fill. */
int32_t cfs_stage(cfs_solver* s, int32_t stage_counter, double dt);
/* All sub-stages of one RK step with time step dt (2 or 3 stages). */
int32_t cfs_step(cfs_solver* s, double dt);
/* nsteps RK steps.  dt > 0: fixed step.  dt <= 0: each step uses
 * dt = cfl / s_max, s_max = max over fluid cells of
 *   (a_x + c)/dx + (a_y + c)/dy,  c = sqrt(gamma*p/rho),
 * a_x = max(|u|,|rho*u|) with the reference HLL bounds (hll_bounds = 0: they
 * use the momentum where the velocity belongs, RiemmanSolverHLL.cpp:41-42),
 * a_x = |u| with velocity bounds; likewise a_y.  Reduced on the device (no
 * reference counterpart: the reference uses a fixed dt, param.h:20).  FP64
 * evaluates it with IEEE divisions; FP32-bulk, still in FP64, with one refined
 * reciprocal of rho and multiplications by 1/dx, 1/dy (within a few ulp).
 * t_end > 0 clips the last step so that the time does not pass t_end and stops there.
 * On return *t_io is advanced by the simulated time (may be NULL). */
int32_t cfs_advance(cfs_solver* s, int32_t nsteps, double dt, double cfl,
                    double t_end, double* t_io, int32_t* steps_done);
/* The CFL time step of the current state: dt = cfl / s_max (s_max as for
 * cfs_advance).  FP64 max reduction; CFS_E_NONFINITE if s_max is not finite
 * or not positive. */
int32_t cfs_compute_dt(cfs_solver* s, double cfl, double* dt_out);
/* Fixed-order FP64 tree sums of (rho, rho*u, rho*v, E) over fluid cells.
 * A sum that is not finite (a NaN or Inf in the state) is written to out and
 * reported as CFS_E_NONFINITE. */
int32_t cfs_checksums(cfs_solver* s, double out[4]);
/* cfs_checksums of a state that the caller holds in device memory, with no
 * handle and no device allocation: the tree's partial sums live in a fixed
 * file-scope device array of the library (about 8.3 MB per GPU, held until
 * the process ends).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops
 * that check conservation totals of resident states).
 *   dU     [4][cells] FP64 state on the current device (x-major planes)
 *   d_type [cells] cell types (CFS_CELL_*; only fluid cells are summed), or NULL
 *   d_out  4 doubles of device memory that receive the sums of
 *          (rho, rho*u, rho*v, E), the same bits as cfs_checksums (the call
 *          does not wait for them, so it cannot report non-finite sums)
 *   hip_stream  stream the kernels are queued on (NULL: the null stream);
 *          the call returns without waiting.
 * cells must lie in [1, CFS_NOALLOC_MAX_CELLS]: larger inputs are CFS_E_LIMIT
 * before anything is queued.  Every call on one GPU shares that GPU's scratch
 * array: calls whose work overlaps on one GPU (other streams, other threads)
 * race and must be ordered by the caller. */
#define CFS_NOALLOC_MAX_CELLS (1ll << 26)
int32_t cfs_checksums_device_noalloc(const double* dU, int64_t cells, const uint8_t* d_type, double* d_out,
                                     void* hip_stream);
/* Copies the current state plane-set into a spare one `repeats` times with a
 * plain streaming copy kernel (32*nx*ny bytes read and written per launch).
 * No reference counterpart: it calibrates hardware byte counters against a
 * kernel of known traffic. */
int32_t cfs_calibrate_copy(cfs_solver* s, int32_t repeats);
/* Number of 15x16-cell tiles of the last FP64 stage that computed their
 * faces with plain IEEE division in WENO5 because they hold a value outside
 * the range (0, or |v| in [2^-900, 2^200]) of the kernel's faster division
 * forms (x/3 and x/6 by two fmas, the weights through one refined
 * reciprocal), which are bit-identical to IEEE division inside it.  Either
 * way the stage is bit-identical to IEEE evaluation; 0 in normal flows. */
int32_t cfs_last_redo_tiles(cfs_solver* s, int32_t* n);
/* The right-hand side L(U) = -(F_{i+1/2} - F_{i-1/2})/dx - (G_{j+1/2} -
 * G_{j-1/2})/dy that the last stage applied, [4][nx][ny], on fluid cells (0
 * elsewhere): what the reference keeps per cell from Cell::comRHS
 * (Cell.cpp:110-125) and writes as FluxRho/FluxMu/FluxMv/FluxEnergy
 * (Output.cpp:66-69).  Recomputed from the last stage's input, which every
 * stage leaves intact.  CFS_E_STATE before the first stage. */
int32_t cfs_last_rhs(cfs_solver* s, double* R);
/* next stage counter (1 + stages done) */
int32_t cfs_stage_counter(const cfs_solver* s);
/* Kernel timing switch (default off): when on, cfs_advance brackets every
 * stage kernel with HIP events on the handle's stream (a few microseconds of
 * queue latency per stage) and cfs_last_kernel_ms reports their mean. */
int32_t cfs_set_kernel_timing(cfs_solver* s, int32_t on);
/* Mean duration (ms) of the stage kernel over the last cfs_advance call made
 * with kernel timing on (0 if none). */
int32_t cfs_last_kernel_ms(const cfs_solver* s, double* ms_per_stage);

/* ---- tile shapes and the autotuner ---------------------------------------- */
/* The stage kernels come in CFS_NUM_TILE_SHAPES tile shapes: a workgroup of
 * `threads` threads updates ti x tj cells (ti along x, tj along y).  Results do
 * not depend on the shape (every face and cell is computed the same way);
 * speed does, with the grid, the precision and the kernel variant.  Shape 0
 * (15 x 16, 256 threads) is the only one of the viscous variant; shape 3 has
 * shape 0's tile with each thread computing its x-face and its y-face in one
 * pass (two independent chains per thread).  A new handle starts with shape
 * 3 for inviscid grids of at most 5000 shape-0 tiles without the positivity
 * limiter and with shape 0 otherwise.  No reference counterpart: the
 * reference has no tiling. */
#define CFS_NUM_TILE_SHAPES 4
int32_t cfs_tile_shape_info(int32_t shape, int32_t* ti, int32_t* tj, int32_t* threads);
/* Selects the tile shape of later stages (CFS_E_UNSUPPORTED for a viscous
 * configuration and shape != 0).  Waits for queued work. */
int32_t cfs_set_tile_shape(cfs_solver* s, int32_t shape);
int32_t cfs_get_tile_shape(const cfs_solver* s, int32_t* shape);
/* Autotuner: times `reps` launches of the stage kernel of every tile shape on
 * the handle's current state and configuration (sub-stage 0 with dt = 0
 * written to a spare plane-set: the state is not changed), in two passes
 * over the shapes, selects the fastest and reports it in *best (may be NULL)
 * and the mean milliseconds per launch of each shape's faster pass in
 * ms[CFS_NUM_TILE_SHAPES] (may be NULL; -1 for shapes the configuration has
 * not).  At a step boundary only (CFS_E_STATE). */
int32_t cfs_autotune(cfs_solver* s, int32_t reps, int32_t* best, double* ms);

/* ---- batched solves ------------------------------------------------------- */
/* Many independent problems of uneven sizes advanced together: each call of
 * cfs_batch_step runs one persistent stage kernel per RK sub-stage over the
 * tiles of every resident problem (two steps per captured HIP graph).  Each
 * problem keeps its own CFL time step, end time, boundary conditions and
 * immersed boundary, and its result is bit-identical to solving it alone
 * with cfs_advance(nsteps, 0, cfl, t_end).  Problems live in one device arena
 * sized at creation: admission is by device-memory footprint
 * (cfs_batch_footprint), and a full arena or slot table is CFS_E_FULL, not an
 * error of the batch.  No reference counterpart: the reference solves one
 * problem per run (main.cpp); this is the data-parallel form of its stage
 * loop (Run.cpp:153-204) for many small solves. */
typedef struct cfs_batch cfs_batch;

typedef struct cfs_batch_config {
    int32_t rk;              /* 2 or 3, shared by every problem of the batch */
    int32_t precision;       /* CFS_PREC_FP64 | CFS_PREC_FP32_BULK, shared */
    int32_t max_problems;    /* slot table size */
    int32_t workgroups;      /* persistent grid; 0: CUs x resident workgroups per CU */
    int64_t arena_bytes;     /* device memory for problem state and tables */
    int32_t limiter;         /* 1: kernel with the positivity limiter (problems may use positivity = 2) */
    int32_t tile_shape;      /* stage-kernel tile shape, 0 .. CFS_NUM_TILE_SHAPES-1
                                (cfs_tile_shape_info; cfs_batch_autotune picks one) */
} cfs_batch_config;

/* Immersed-boundary tables of one problem, as cfs_set_immersed takes them. */
typedef struct cfs_immersed {
    int32_t interp;
    int32_t n_ig;
    const uint8_t* cell_type;
    const int32_t* ig_cells;
    const int32_t* donors;
    const double* image_xy;
    const double* coef;
    const int32_t* transp;
} cfs_immersed;

enum { CFS_SLOT_FREE = 0, CFS_SLOT_RUNNING = 1, CFS_SLOT_DONE = 2, CFS_SLOT_FAILED = 3 };

typedef struct cfs_batch_info {
    int32_t slot;
    int32_t state;           /* CFS_SLOT_* */
    int32_t steps;           /* RK steps taken */
    int32_t nx, ny;
    double t;                /* simulated time reached */
} cfs_batch_info;

int32_t cfs_batch_create(const cfs_batch_config* bc, void* hip_stream, cfs_batch** out);
void cfs_batch_destroy(cfs_batch* b);
/* Arena bytes a problem occupies: n_ig = 0 without an immersed boundary,
 * else the number of immersed ghosts of its tables. */
int64_t cfs_batch_footprint(const cfs_config* cfg, int32_t n_ig);
/* Admits a problem with initial state U ([4][nx][ny], copied), optional
 * immersed-boundary tables (NULL: none), CFL number cfl > 0 and end time
 * t_end (<= 0: runs until retired), starting at t = 0.  The problem's rk and
 * precision must be the batch's; viscous problems are CFS_E_UNSUPPORTED;
 * immersed-boundary tables need n_ig > 0.
 * On success *slot names it.  Waits for the batch's queued work. */
int32_t cfs_batch_admit(cfs_batch* b, const cfs_config* cfg, const double* U, const cfs_immersed* ib,
                        double cfl, double t_end, int32_t* slot);
/* Queues nsteps RK steps of every resident problem and returns without
 * waiting.  Problems past their end time (or with a non-finite bound) stay
 * halted until retired. */
int32_t cfs_batch_step(cfs_batch* b, int32_t nsteps);
/* Waits for the queued work and reports every occupied slot (at most cap
 * entries into info; *n = number of occupied slots). */
int32_t cfs_batch_poll(cfs_batch* b, cfs_batch_info* info, int32_t cap, int32_t* n);
/* Copies a problem's state out (U may be NULL) and frees its slot and arena
 * space.  Waits for the queued work. */
int32_t cfs_batch_retire(cfs_batch* b, int32_t slot, double* U);
/* Free arena bytes (the largest admissible footprint is at most this). */
int32_t cfs_batch_free_bytes(const cfs_batch* b, int64_t* bytes);
/* Arena size and slot count the batch was created with. */
int32_t cfs_batch_capacity(const cfs_batch* b, int64_t* arena_bytes, int32_t* max_problems);

struct cfs_request;
/* Autotuner of the batched kernel: for every tile shape, creates a batch of
 * configuration *bc with that shape (bc->tile_shape is ignored), admits the n
 * sample problems of `sample` (cfg, U, ib and cfl are read; each runs without
 * an end time), runs 2 warm-up steps and times `steps` RK steps of the whole
 * batch with HIP events.  *best receives the fastest shape, ms_per_step
 * [CFS_NUM_TILE_SHAPES] the milliseconds per step of each (may be NULL).
 * CFS_E_FULL if the sample does not fit the arena.  No reference
 * counterpart. */
int32_t cfs_batch_autotune(const cfs_batch_config* bc, const struct cfs_request* sample, int32_t n, int32_t steps,
                           int32_t* best, double* ms_per_step);

/* ---- serving loop --------------------------------------------------------- */
/* A request: one solve to t_end, arriving at arrival_s seconds after the
 * start of cfs_serve.  Outputs are written by cfs_serve. */
typedef struct cfs_request {
    const cfs_config* cfg;
    const double* U;          /* initial state [4][nx][ny] */
    const cfs_immersed* ib;   /* NULL: no immersed boundary */
    double cfl, t_end;        /* t_end > 0 */
    double arrival_s;
    double* U_out;            /* NULL, or [4][nx][ny] for the final state */
    /* outputs */
    double admitted_s;        /* admission time (seconds from the start) */
    double finished_s;        /* time its completion was observed */
    double t;                 /* simulated time reached */
    int32_t steps;
    int32_t state;            /* CFS_SLOT_DONE | CFS_SLOT_FAILED; CFS_SLOT_FREE if never admitted */
} cfs_request;

/* Runs requests through a batch with continuous batching: requests that have
 * arrived are admitted in arrival order while their footprint fits the arena
 * and a slot is free (a request that does not fit waits; later ones that fit
 * may pass it), every resident problem advances steps_per_poll RK steps per
 * round, and finished problems are retired at the end of the round they
 * finish in, freeing their space for waiting requests.  A request that fills
 * the GPU on its own (at least as many tiles as the GPU holds stage-kernel
 * workgroups at once) is solved instead by worker threads through the
 * single-solve entry points (cfs_create, cfs_advance; its own stream and
 * device memory), up to two such requests at a time in arrival order,
 * alongside the rounds; its result is the same as in the batch.  The solo
 * lane admits by footprint as well: its handles may hold at most nine tenths
 * of the device memory free when cfs_serve starts, and a worker keeps its
 * handle for a next request of the same configuration.  Returns when every
 * request has finished; a request the batch would refuse is CFS_E_ARG (or
 * CFS_E_UNSUPPORTED), and one whose footprint exceeds the whole arena without
 * filling the GPU, or the solo lane's budget when it does, is CFS_E_FULL,
 * before anything runs.  No reference counterpart. */
int32_t cfs_serve(cfs_batch* b, cfs_request* reqs, int32_t n, int32_t steps_per_poll);

#ifdef __cplusplus
}
#endif
#endif /* CFS_H */
