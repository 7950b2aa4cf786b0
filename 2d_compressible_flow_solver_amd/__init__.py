"""Python host side of the MI355X per-stage hot path (ctypes over libcfs.so).

The C-ABI is include/cfs.h.  This module mirrors the reference's driver-level
objects for scripting and tests:

  Config / config_*()       the mesh + plugin choices of Run::initialize
                            (Run.cpp:10-90) and param.h, as one cfs_config
  Solver                    a device-resident solve; stage()/step() replace one
                            iteration / one RK step of Run::solver (Run.cpp:153-204)
  circle_setup()            CartesianMesh::setImmersedBoundary (CartesianMesh.cpp:357-374)
  riemann_quadrants(), cylinder_freestream(), kelvin_helmholtz()
                            initial conditions (Initialcondition::fillIn,
                            source/initialcondition/*.cpp), host-side inputs

The package directory name starts with a digit, so load it with
importlib.import_module("2d_compressible_flow_solver_amd").

There is no CPU fallback: every compute call goes to the HIP library and
raises CfsError when it is missing or a HIP call fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfs.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cfs.h")

CFS_OK = 0
CFS_E_ARG = -1
CFS_E_HIP = -2
CFS_E_OOM = -3
CFS_E_STATE = -4
CFS_E_NONFINITE = -5
CFS_E_UNSUPPORTED = -6
CFS_E_FULL = -7
CFS_E_LIMIT = -8
NOALLOC_MAX_CELLS = 1 << 26
SLOT_FREE, SLOT_RUNNING, SLOT_DONE, SLOT_FAILED = 0, 1, 2, 3

BC_NEUMANN, BC_DIRICHLET, BC_PERIODIC = 0, 1, 2
SIDE_LEFT, SIDE_RIGHT, SIDE_BOTTOM, SIDE_TOP = 0, 1, 2, 3
PREC_FP64, PREC_FP32_BULK = 0, 1
CELL_FLUID, CELL_SOLID, CELL_IMMERSED_GHOST = 0, 1, 2
IB_REFERENCE, IB_BILINEAR_POSITIVE = 0, 1


class CfsError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _lib().cfs_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__("%s%s" % (what + ": " if what else "", msg))


class Config(ctypes.Structure):
    """cfs_config of include/cfs.h (same field order and types)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double),
        ("x0", ctypes.c_double), ("y0", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("bc", ctypes.c_int32 * 4),
        ("bc_state", (ctypes.c_double * 4) * 4),
        ("rk", ctypes.c_int32), ("precision", ctypes.c_int32), ("viscous", ctypes.c_int32),
        ("positivity", ctypes.c_int32), ("hll_bounds", ctypes.c_int32),
        ("mu", ctypes.c_double), ("prandtl", ctypes.c_double), ("ma_ref", ctypes.c_double),
    ]

    @property
    def dx(self):
        return self.lx / self.nx

    @property
    def dy(self):
        return self.ly / self.ny

    def centres(self):
        """Cell centres ((i+0.5)*dx + x0, (j+0.5)*dy + y0), arrays [nx][ny]
        (CartesianMesh::SetElement, CartesianMesh.cpp:82-97)."""
        i = np.arange(self.nx, dtype=np.float64)
        j = np.arange(self.ny, dtype=np.float64)
        x = (i + 0.5) * self.dx + self.x0
        y = (j + 0.5) * self.dy + self.y0
        return np.meshgrid(x, y, indexing="ij")


def make_config(nx, ny, lx, ly, bc=(BC_NEUMANN,) * 4, bc_state=None, gamma=1.4, rk=3,
                precision=PREC_FP64, x0=0.0, y0=0.0, viscous=0, mu=0.0, prandtl=0.72, ma_ref=1.0,
                positivity=0, hll_bounds=0):
    c = Config()
    c.nx, c.ny, c.lx, c.ly, c.x0, c.y0, c.gamma = nx, ny, lx, ly, x0, y0, gamma
    for k in range(4):
        c.bc[k] = bc[k]
        st = (bc_state[k] if bc_state is not None and bc_state[k] is not None else (0.0, 0.0, 0.0, 0.0))
        for q in range(4):
            c.bc_state[k][q] = st[q]
    c.rk, c.precision, c.viscous, c.positivity, c.hll_bounds = rk, precision, viscous, positivity, hll_bounds
    c.mu, c.prandtl, c.ma_ref = mu, prandtl, ma_ref
    return c


def conserved(rho, u, v, p, gamma=1.4):
    """(rho, u, v, p) -> (rho, rho*u, rho*v, E), E by EulerPhase::computeEnergyfromOther
    (EulerPhase.cpp:124-128): e = p/(gamma-1); E = e + 0.5*(mu*mu + mv*mv)/rho."""
    rho = np.asarray(rho, dtype=np.float64)
    mu = rho * np.asarray(u, dtype=np.float64)
    mv = rho * np.asarray(v, dtype=np.float64)
    e = np.asarray(p, dtype=np.float64) / (gamma - 1)
    E = e + 0.5 * (mu * mu + mv * mv) / rho
    return rho, mu, mv, E


# ---------------------------------------------------------------------------
# named workloads (the configurations the hot path is measured and tested on)

LAX_LIU_3 = {  # (rho, u, v, p) per quadrant
    "NE": (1.5, 0.0, 0.0, 1.5),
    "NW": (0.5323, 1.206, 0.0, 0.3),
    "SW": (0.138, 1.206, 1.206, 0.029),
    "SE": (0.5323, 0.0, 1.206, 0.3),
}


def riemann_quadrants(cfg, states=LAX_LIU_3, split=(0.5, 0.5)):
    """Four-quadrant Riemann initial state [4][nx][ny] (Reimman2D::fillIn,
    Reimman2D.cpp:5-39, with the split point as a parameter: the reference
    hard-codes 1.6)."""
    x, y = cfg.centres()
    sx, sy = split
    U = np.empty((4, cfg.nx, cfg.ny))
    for key, m in (("SW", (x < sx) & (y < sy)), ("NW", (x < sx) & (y >= sy)),
                   ("SE", (x >= sx) & (y < sy)), ("NE", (x >= sx) & (y >= sy))):
        cons = conserved(*states[key], gamma=cfg.gamma)
        for q in range(4):
            U[q][m] = cons[q]
    return U


def config_riemann(n=400, rk=3, precision=PREC_FP64):
    """Lax-Liu configuration 3 on [0,1]^2, transmissive boundaries."""
    return make_config(n, n, 1.0, 1.0, bc=(BC_NEUMANN,) * 4, rk=rk, precision=precision)


CYLINDER_FREESTREAM = (1.4, 3.0, 0.0, 1.0)   # rho, u, v, p: Mach 3 at c = 1


def config_cylinder(nx=2048, ny=1024, lx=4.0, ly=2.0, freestream=CYLINDER_FREESTREAM, rk=3,
                    precision=PREC_FP64, gamma=1.4, positivity=1, hll_bounds=1):
    """Supersonic inflow on the left (Dirichlet), extrapolation elsewhere
    (Cylinder::setBounbdaryCondition, Cylinder.cpp:19-40)."""
    inflow = tuple(float(a) for a in conserved(*freestream, gamma=gamma))
    return make_config(nx, ny, lx, ly, bc=(BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN, BC_NEUMANN),
                       bc_state=(inflow, None, None, None), rk=rk, precision=precision, gamma=gamma,
                       positivity=positivity, hll_bounds=hll_bounds)


def cylinder_freestream(cfg, cell_type, freestream=CYLINDER_FREESTREAM):
    """Freestream on fluid cells, zero elsewhere (Cylinder::fillIn applied to
    fluid cells only, Run.cpp:99-105)."""
    cons = conserved(*freestream, gamma=cfg.gamma)
    U = np.zeros((4, cfg.nx, cfg.ny))
    fluid = cell_type.reshape(cfg.nx, cfg.ny) == CELL_FLUID
    for q in range(4):
        U[q][fluid] = cons[q]
    return U


def splitmix64(z):
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):   # arithmetic mod 2^64
        z = np.asarray(z, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def cell_noise(seed, nx, ny):
    """Counter-based uniform noise in [-1, 1) per cell: a function of
    (seed, i, j) only, so any producer of the same grid draws the same values.
    u = top 53 bits of splitmix64(splitmix64(seed) ^ (i << 32 | j)) / 2^53."""
    i = np.arange(nx, dtype=np.uint64)[:, None]
    j = np.arange(ny, dtype=np.uint64)[None, :]
    key = splitmix64(np.uint64(seed)) ^ ((i << np.uint64(32)) | j)
    u = (splitmix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return 2.0 * u - 1.0


KH_SEED = 0x5EED


def config_kelvin_helmholtz(n=4096, precision=PREC_FP64, rk=3, mu=1e-5, prandtl=0.72, ma_ref=1.0,
                            gamma=1.4):
    """Viscous shear layer on the periodic unit square: Navier-Stokes terms
    with constant mu = 1/Re (Re = 1e5), Pr = 0.72, T = gamma Ma^2 p/rho."""
    return make_config(n, n, 1.0, 1.0, bc=(BC_PERIODIC,) * 4, gamma=gamma, rk=rk, precision=precision,
                       viscous=1, mu=mu, prandtl=prandtl, ma_ref=ma_ref)


def kelvin_helmholtz(cfg, seed=KH_SEED, amp=0.01, noise=1e-3):
    """rho = 2, u = +0.5 in 0.25 < y < 0.75; rho = 1, u = -0.5 elsewhere;
    p = 2.5; v = amp*sin(4 pi x) + noise * cell_noise(seed).  [4][nx][ny]."""
    x, y = cfg.centres()
    band = (y > 0.25) & (y < 0.75)
    rho = np.where(band, 2.0, 1.0)
    u = np.where(band, 0.5, -0.5)
    v = amp * np.sin(4.0 * np.pi * x) + noise * cell_noise(seed, cfg.nx, cfg.ny)
    return np.stack(conserved(rho, u, v, np.full_like(rho, 2.5), gamma=cfg.gamma))


class BatchConfig(ctypes.Structure):
    """cfs_batch_config of include/cfs.h."""
    _fields_ = [("rk", ctypes.c_int32), ("precision", ctypes.c_int32), ("max_problems", ctypes.c_int32),
                ("workgroups", ctypes.c_int32), ("arena_bytes", ctypes.c_int64), ("limiter", ctypes.c_int32),
                ("tile_shape", ctypes.c_int32)]


class Immersed(ctypes.Structure):
    """cfs_immersed of include/cfs.h."""
    _fields_ = [("interp", ctypes.c_int32), ("n_ig", ctypes.c_int32), ("cell_type", ctypes.c_void_p),
                ("ig_cells", ctypes.c_void_p), ("donors", ctypes.c_void_p), ("image_xy", ctypes.c_void_p),
                ("coef", ctypes.c_void_p), ("transp", ctypes.c_void_p)]


class Request(ctypes.Structure):
    """cfs_request of include/cfs.h."""
    _fields_ = [("cfg", ctypes.c_void_p), ("U", ctypes.c_void_p), ("ib", ctypes.c_void_p),
                ("cfl", ctypes.c_double), ("t_end", ctypes.c_double), ("arrival_s", ctypes.c_double),
                ("U_out", ctypes.c_void_p), ("admitted_s", ctypes.c_double), ("finished_s", ctypes.c_double),
                ("t", ctypes.c_double), ("steps", ctypes.c_int32), ("state", ctypes.c_int32)]


class BatchInfo(ctypes.Structure):
    """cfs_batch_info of include/cfs.h."""
    _fields_ = [("slot", ctypes.c_int32), ("state", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("t", ctypes.c_double)]


# ---------------------------------------------------------------------------
_LIB = None


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcfs.so is not built at %s (run __graft_entry__.build())" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    dbl, i32 = ctypes.c_double, ctypes.c_int32
    vp = ctypes.c_void_p
    lib.cfs_abi_version.restype = i32
    lib.cfs_status_string.restype = ctypes.c_char_p
    lib.cfs_status_string.argtypes = [i32]
    lib.cfs_device_count.restype = i32
    lib.cfs_create.argtypes = [P(Config), vp, P(vp)]
    lib.cfs_destroy.argtypes = [vp]
    lib.cfs_destroy.restype = None
    for name in ("cfs_set_state", "cfs_get_state", "cfs_set_state_device", "cfs_get_state_device"):
        getattr(lib, name).argtypes = [vp, vp]
    lib.cfs_circle_setup.argtypes = [P(Config), dbl, dbl, dbl, i32, i32, i32, i32, vp, P(i32), vp, vp, vp, vp,
                                     vp]
    lib.cfs_circle_setup_device.argtypes = [P(Config), dbl, dbl, dbl, i32, i32, i32, i32, vp, P(i32), vp, vp, vp,
                                            vp, vp, P(i32)]
    lib.cfs_set_immersed.argtypes = [vp, i32, vp, i32, vp, vp, vp, vp, vp]
    lib.cfs_stage.argtypes = [vp, i32, dbl]
    lib.cfs_step.argtypes = [vp, dbl]
    lib.cfs_advance.argtypes = [vp, i32, dbl, dbl, dbl, P(dbl), P(i32)]
    lib.cfs_compute_dt.argtypes = [vp, dbl, P(dbl)]
    lib.cfs_checksums.argtypes = [vp, vp]
    lib.cfs_checksums_device_noalloc.argtypes = [vp, ctypes.c_int64, vp, vp, vp]
    lib.cfs_checksums_device_noalloc.restype = i32
    lib.cfs_stage_counter.argtypes = [vp]
    lib.cfs_stage_counter.restype = i32
    lib.cfs_last_kernel_ms.argtypes = [vp, P(dbl)]
    lib.cfs_last_rhs.argtypes = [vp, vp]
    lib.cfs_last_rhs.restype = i32
    lib.cfs_batch_create.argtypes = [P(BatchConfig), vp, P(vp)]
    lib.cfs_batch_destroy.argtypes = [vp]
    lib.cfs_batch_destroy.restype = None
    lib.cfs_batch_footprint.argtypes = [P(Config), i32]
    lib.cfs_batch_footprint.restype = ctypes.c_int64
    lib.cfs_batch_admit.argtypes = [vp, P(Config), vp, P(Immersed), dbl, dbl, P(i32)]
    lib.cfs_batch_step.argtypes = [vp, i32]
    lib.cfs_batch_poll.argtypes = [vp, P(BatchInfo), i32, P(i32)]
    lib.cfs_batch_retire.argtypes = [vp, i32, vp]
    lib.cfs_batch_free_bytes.argtypes = [vp, P(ctypes.c_int64)]
    lib.cfs_batch_capacity.argtypes = [vp, P(ctypes.c_int64), P(i32)]
    lib.cfs_serve.argtypes = [vp, P(Request), i32, i32]
    lib.cfs_tile_shape_info.argtypes = [i32, P(i32), P(i32), P(i32)]
    lib.cfs_set_tile_shape.argtypes = [vp, i32]
    lib.cfs_get_tile_shape.argtypes = [vp, P(i32)]
    lib.cfs_autotune.argtypes = [vp, i32, P(i32), vp]
    lib.cfs_batch_autotune.argtypes = [P(BatchConfig), P(Request), i32, i32, P(i32), vp]
    for name in ("cfs_tile_shape_info", "cfs_set_tile_shape", "cfs_get_tile_shape", "cfs_autotune",
                 "cfs_batch_autotune"):
        getattr(lib, name).restype = i32
    for name in ("cfs_batch_create", "cfs_batch_admit", "cfs_batch_step", "cfs_batch_poll", "cfs_batch_retire",
                 "cfs_batch_free_bytes", "cfs_batch_capacity", "cfs_serve"):
        getattr(lib, name).restype = i32
    for name in ("cfs_create", "cfs_set_state", "cfs_get_state", "cfs_set_state_device",
                 "cfs_get_state_device", "cfs_circle_setup", "cfs_circle_setup_device", "cfs_set_immersed",
                 "cfs_stage",
                 "cfs_step", "cfs_advance", "cfs_compute_dt", "cfs_checksums", "cfs_last_kernel_ms"):
        getattr(lib, name).restype = i32
    _LIB = lib
    return lib


def library():
    """The loaded ctypes library (raises if libcfs.so is not built)."""
    return _lib()


def _check(rc, what):
    if rc != CFS_OK:
        raise CfsError(rc, what)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_count():
    return _lib().cfs_device_count()


NUM_TILE_SHAPES = 4


def tile_shape_info(shape):
    """(ti, tj, threads) of stage-kernel tile shape `shape`."""
    ti, tj, nt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _check(_lib().cfs_tile_shape_info(shape, ctypes.byref(ti), ctypes.byref(tj), ctypes.byref(nt)),
           "cfs_tile_shape_info")
    return ti.value, tj.value, nt.value


def circle_setup(cfg, cx, cy, radius, n_lagrangian=400, n_layers=3, interp=IB_BILINEAR_POSITIVE, device=False):
    """Immersed-boundary tables of a circle.  device=False: host code of
    libcfs.so (no GPU needed); device=True: cfs_circle_setup_device (the same
    tables, computed on the GPU; out["n_host"] = ghosts finished on the host).
    interp: IB_REFERENCE (the reference's LDLT-on-lower-triangle interpolation,
    as written) or IB_BILINEAR_POSITIVE (exact bilinear fit when non-degenerate
    and non-negative, else inverse-distance weights).
    Returns dict(interp, cell_type, n_ig, ig_cells, donors, image_xy, coef, transp)."""
    lib = _lib()
    n = cfg.nx * cfg.ny
    cell_type = np.zeros(n, dtype=np.uint8)
    n_ig = ctypes.c_int32(0)
    n_host = ctypes.c_int32(0)
    name = "cfs_circle_setup_device" if device else "cfs_circle_setup"
    fn = getattr(lib, name)
    extra = (ctypes.byref(n_host),) if device else ()

    def tables(cap):
        return dict(interp=interp, cell_type=cell_type,
                    ig_cells=np.zeros(cap, dtype=np.int32), donors=np.zeros(4 * cap, dtype=np.int32),
                    image_xy=np.zeros(2 * cap), coef=np.zeros(10 * cap), transp=np.zeros(4 * cap, dtype=np.int32))

    def call(cap, out):
        args = (None,) * 5 if out is None else tuple(_ptr(out[k]) for k in
                                                      ("ig_cells", "donors", "image_xy", "coef", "transp"))
        return fn(ctypes.byref(cfg), cx, cy, radius, n_lagrangian, n_layers, interp, cap, _ptr(cell_type),
                  ctypes.byref(n_ig), *args, *extra)

    if device:
        # one call with room for the ghosts a ring of n_layers + 2 cells
        # around the body can hold (a second only if that was short)
        h = min(cfg.lx / cfg.nx, cfg.ly / cfg.ny)
        cap = int(2.0 * np.pi * radius * (n_layers + 2) / h * 1.5) + 64
        out = tables(cap)
        rc = call(cap, out)
        if rc == CFS_E_ARG and n_ig.value > cap:
            cap = n_ig.value
            out = tables(cap)
            rc = call(cap, out)
    else:
        rc = call(0, None)   # count first
        if rc not in (CFS_OK, CFS_E_ARG):
            _check(rc, name)
        cap = max(n_ig.value, 1)
        out = tables(cap)
        rc = call(cap, out)
    _check(rc, name)
    k = n_ig.value
    out["n_ig"] = k
    if device:
        out["n_host"] = n_host.value
    for key, w in (("ig_cells", 1), ("donors", 4), ("image_xy", 2), ("coef", 10), ("transp", 4)):
        out[key] = out[key][: w * k]
    return out


class Solver:
    """One device-resident solve (a cfs_solver handle)."""

    def __init__(self, cfg, stream=None):
        self.cfg = cfg
        self._lib = _lib()
        h = ctypes.c_void_p()
        _check(self._lib.cfs_create(ctypes.byref(cfg), stream, ctypes.byref(h)), "cfs_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cfs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (4, self.cfg.nx, self.cfg.ny)

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.shape:
            raise ValueError("state must be %s, got %s" % (self.shape, U.shape))
        _check(self._lib.cfs_set_state(self._h, _ptr(U)), "cfs_set_state")

    def get_state(self):
        U = np.empty(self.shape)
        _check(self._lib.cfs_get_state(self._h, _ptr(U)), "cfs_get_state")
        return U

    def set_state_device(self, ptr):
        _check(self._lib.cfs_set_state_device(self._h, ctypes.c_void_p(ptr)), "cfs_set_state_device")

    def get_state_device(self, ptr):
        _check(self._lib.cfs_get_state_device(self._h, ctypes.c_void_p(ptr)), "cfs_get_state_device")

    def set_immersed(self, ib):
        _check(self._lib.cfs_set_immersed(self._h, int(ib["interp"]), _ptr(ib["cell_type"]), int(ib["n_ig"]),
                                          _ptr(ib["ig_cells"]), _ptr(ib["donors"]), _ptr(ib["image_xy"]),
                                          _ptr(ib["coef"]), _ptr(ib["transp"])), "cfs_set_immersed")

    @property
    def counter(self):
        return self._lib.cfs_stage_counter(self._h)

    def stage(self, counter, dt):
        _check(self._lib.cfs_stage(self._h, counter, dt), "cfs_stage")

    def step(self, dt):
        _check(self._lib.cfs_step(self._h, dt), "cfs_step")

    def advance(self, nsteps, dt=0.0, cfl=0.5, t_end=0.0, t=0.0):
        """nsteps RK steps; dt > 0 fixed, else CFL.  Returns (t, steps_done)."""
        tt = ctypes.c_double(t)
        done = ctypes.c_int32(0)
        _check(self._lib.cfs_advance(self._h, nsteps, dt, cfl, t_end, ctypes.byref(tt), ctypes.byref(done)),
               "cfs_advance")
        return tt.value, done.value

    def compute_dt(self, cfl):
        dt = ctypes.c_double(0.0)
        _check(self._lib.cfs_compute_dt(self._h, cfl, ctypes.byref(dt)), "cfs_compute_dt")
        return dt.value

    def checksums(self, check_finite=True):
        """Fixed-order sums of (rho, rho*u, rho*v, E) over fluid cells.  A
        non-finite sum raises CfsError (CFS_E_NONFINITE) unless check_finite
        is False, which returns the sums as they are."""
        out = np.zeros(4)
        rc = self._lib.cfs_checksums(self._h, _ptr(out))
        if rc != CFS_E_NONFINITE or check_finite:
            _check(rc, "cfs_checksums")
        return out

    def set_kernel_timing(self, on=True):
        _check(self._lib.cfs_set_kernel_timing(self._h, 1 if on else 0), "cfs_set_kernel_timing")

    def exact_tiles(self):
        """Tiles of the last FP64 stage that took the IEEE-division form of WENO5
        (a staged value outside the fast form's range); 0 in normal flows."""
        n = ctypes.c_int32(0)
        _check(self._lib.cfs_last_redo_tiles(self._h, ctypes.byref(n)), "cfs_last_redo_tiles")
        return n.value

    def last_rhs(self):
        """L(U) applied by the last stage on fluid cells, [4][nx][ny]."""
        R = np.empty((4, self.cfg.nx, self.cfg.ny))
        _check(self._lib.cfs_last_rhs(self._h, _ptr(R)), "cfs_last_rhs")
        return R

    def last_kernel_ms(self):
        ms = ctypes.c_double(0.0)
        _check(self._lib.cfs_last_kernel_ms(self._h, ctypes.byref(ms)), "cfs_last_kernel_ms")
        return ms.value

    @property
    def tile_shape(self):
        k = ctypes.c_int32(0)
        _check(self._lib.cfs_get_tile_shape(self._h, ctypes.byref(k)), "cfs_get_tile_shape")
        return k.value

    def set_tile_shape(self, shape):
        _check(self._lib.cfs_set_tile_shape(self._h, shape), "cfs_set_tile_shape")

    def autotune(self, reps=20):
        """Times every tile shape on the current state, selects the fastest.
        Returns (best shape, [ms per stage-kernel launch of each shape])."""
        best = ctypes.c_int32(0)
        ms = np.zeros(NUM_TILE_SHAPES)
        _check(self._lib.cfs_autotune(self._h, reps, ctypes.byref(best), _ptr(ms)), "cfs_autotune")
        return best.value, [float(x) for x in ms]


class Batch:
    """Many independent solves advanced together (a cfs_batch handle).

    admit() returns a slot, or None when the arena or slot table is full;
    step() queues RK steps of every resident problem; poll() waits and
    returns {slot: (state, steps, t)}; retire() copies a problem out and
    frees its slot."""

    def __init__(self, rk=3, precision=PREC_FP64, max_problems=256, arena_bytes=1 << 30, workgroups=0,
                 stream=None, limiter=False, tile_shape=0):
        self._lib = _lib()
        self.bc = BatchConfig(rk, precision, max_problems, workgroups, arena_bytes, 1 if limiter else 0,
                              tile_shape)
        h = ctypes.c_void_p()
        _check(self._lib.cfs_batch_create(ctypes.byref(self.bc), stream, ctypes.byref(h)), "cfs_batch_create")
        self._h = h
        self._cfg = {}
        self._info = (BatchInfo * max_problems)()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cfs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def footprint(cfg, n_ig=0):
        return int(_lib().cfs_batch_footprint(ctypes.byref(cfg), n_ig))

    def admit(self, cfg, U, cfl=0.5, t_end=0.0, ib=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (4, cfg.nx, cfg.ny):
            raise ValueError("state shape %s, expected %s" % (U.shape, (4, cfg.nx, cfg.ny)))
        imm = None
        keep = [U]
        if ib is not None:
            arrs = [np.ascontiguousarray(ib["cell_type"], dtype=np.uint8),
                    np.ascontiguousarray(ib["ig_cells"], dtype=np.int32),
                    np.ascontiguousarray(ib["donors"], dtype=np.int32),
                    np.ascontiguousarray(ib["image_xy"], dtype=np.float64),
                    np.ascontiguousarray(ib["coef"], dtype=np.float64),
                    np.ascontiguousarray(ib["transp"], dtype=np.int32)]
            keep += arrs
            imm = Immersed(ib["interp"], int(ib["n_ig"]), *[_ptr(a) for a in arrs])
        slot = ctypes.c_int32(-1)
        rc = self._lib.cfs_batch_admit(self._h, ctypes.byref(cfg), _ptr(U),
                                       ctypes.byref(imm) if imm is not None else None, cfl, t_end,
                                       ctypes.byref(slot))
        if rc == CFS_E_FULL:
            return None
        _check(rc, "cfs_batch_admit")
        self._cfg[slot.value] = cfg
        return slot.value

    def step(self, nsteps=1):
        _check(self._lib.cfs_batch_step(self._h, nsteps), "cfs_batch_step")

    def poll(self):
        n = ctypes.c_int32(0)
        _check(self._lib.cfs_batch_poll(self._h, self._info, len(self._info), ctypes.byref(n)), "cfs_batch_poll")
        return {r.slot: (r.state, r.steps, r.t) for r in self._info[:n.value]}

    def retire(self, slot, copy=True):
        cfg = self._cfg.pop(slot)
        U = np.empty((4, cfg.nx, cfg.ny)) if copy else None
        _check(self._lib.cfs_batch_retire(self._h, slot, _ptr(U) if copy else None), "cfs_batch_retire")
        return U

    def free_bytes(self):
        v = ctypes.c_int64(0)
        _check(self._lib.cfs_batch_free_bytes(self._h, ctypes.byref(v)), "cfs_batch_free_bytes")
        return v.value

    def serve(self, requests, steps_per_poll=5, keep_states=True):
        """Runs cfs_serve over requests = [dict(cfg, U, ib=None, cfl, t_end,
        arrival_s)].  Returns one dict per request with admitted_s,
        finished_s, latency_s, t, steps, state and (keep_states) U."""
        n = len(requests)
        arr, keep, outs = _request_array(requests, keep_states)
        _check(self._lib.cfs_serve(self._h, arr, n, steps_per_poll), "cfs_serve")
        res = []
        for k in range(n):
            r = arr[k]
            res.append(dict(admitted_s=r.admitted_s, finished_s=r.finished_s,
                            latency_s=r.finished_s - requests[k].get("arrival_s", 0.0), t=r.t, steps=r.steps,
                            state=r.state, U=outs[k]))
        del keep
        return res


def _request_array(requests, keep_states=False):
    """cfs_request array of request dicts (cfg, U, ib=None, cfl, t_end,
    arrival_s); returns (array, objects to keep alive, output states)."""
    n = len(requests)
    arr = (Request * n)()
    keep = []
    outs = []
    for k, rq in enumerate(requests):
        cfg = rq["cfg"]
        U = np.ascontiguousarray(rq["U"], dtype=np.float64)
        keep += [cfg, U]
        r = arr[k]
        r.cfg = ctypes.cast(ctypes.pointer(cfg), ctypes.c_void_p)
        r.U = _ptr(U).value
        if rq.get("ib") is not None:
            ib = rq["ib"]
            arrs = [np.ascontiguousarray(ib["cell_type"], dtype=np.uint8),
                    np.ascontiguousarray(ib["ig_cells"], dtype=np.int32),
                    np.ascontiguousarray(ib["donors"], dtype=np.int32),
                    np.ascontiguousarray(ib["image_xy"], dtype=np.float64),
                    np.ascontiguousarray(ib["coef"], dtype=np.float64),
                    np.ascontiguousarray(ib["transp"], dtype=np.int32)]
            imm = Immersed(ib["interp"], int(ib["n_ig"]), *[_ptr(a) for a in arrs])
            keep += arrs + [imm]
            r.ib = ctypes.cast(ctypes.pointer(imm), ctypes.c_void_p)
        r.cfl, r.t_end, r.arrival_s = rq.get("cfl", 0.5), rq["t_end"], rq.get("arrival_s", 0.0)
        out = np.empty((4, cfg.nx, cfg.ny)) if keep_states else None
        outs.append(out)
        r.U_out = _ptr(out).value if keep_states else None
    return arr, keep, outs


def batch_autotune(problems, rk=3, precision=PREC_FP64, limiter=False, steps=10, arena_bytes=1 << 30,
                   max_problems=256, workgroups=0):
    """Tile-shape autotuner of the batched kernel (cfs_batch_autotune): times
    `steps` RK steps of a batch of `problems` (request dicts with cfg, U,
    optional ib, cfl; or (cfg, U) pairs) with every tile shape.  Returns
    (best shape, [ms per step of each shape])."""
    reqs = [p if isinstance(p, dict) else dict(cfg=p[0], U=p[1], ib=None, cfl=0.5, t_end=0.0) for p in problems]
    reqs = [dict(r, t_end=r.get("t_end", 0.0)) for r in reqs]
    arr, keep, _ = _request_array(reqs, keep_states=False)
    bc = BatchConfig(rk, precision, max(max_problems, len(reqs)), workgroups, arena_bytes, 1 if limiter else 0, 0)
    best = ctypes.c_int32(0)
    ms = np.zeros(NUM_TILE_SHAPES)
    _check(_lib().cfs_batch_autotune(ctypes.byref(bc), arr, len(reqs), steps, ctypes.byref(best), _ptr(ms)),
           "cfs_batch_autotune")
    del keep
    return best.value, [float(x) for x in ms]


# ---------------------------------------------------------------------------
# batched workload: many two-dimensional Riemann problems of uneven size

def riemann_batch_problems(seed, count=256, sizes=tuple(range(64, 513, 32)), precision=PREC_FP32_BULK, rk=2,
                           positivity=2, hll_bounds=1):
    """Seeded random four-quadrant Riemann problems on [0,1]^2 with the split at
    (0.5, 0.5) and transmissive boundaries: N ~ U(sizes) square, per quadrant
    rho ~ U[0.1, 2], p ~ logU[0.1, 10], u, v ~ U[-1.5, 1.5].  Returns a list
    of (cfg, U).  The positivity-preserving limiter (positivity = 2) and
    velocity HLL bounds are on by default: with the reference's scheme alone
    about 40% of such random states reach a negative pressure within a few
    steps, with the face-state fallback alone (positivity = 1) about 30% at
    N = 64-512 (four rarefactions meeting at the quadrant corner), with the
    limiter none -- with TVD-RK2 (the reference driver's scheme, Run.cpp:47);
    the reference's RK3 as written (last stage not a convex combination of
    forward-Euler steps) still loses positivity on some of them."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(rng.choice(np.asarray(sizes)))
        states = {}
        for q in ("NE", "NW", "SW", "SE"):
            rho = rng.uniform(0.1, 2.0)
            p = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
            u, v = rng.uniform(-1.5, 1.5, size=2)
            states[q] = (rho, float(u), float(v), p)
        cfg = make_config(n, n, 1.0, 1.0, bc=(BC_NEUMANN,) * 4, rk=rk, precision=precision,
                          positivity=positivity, hll_bounds=hll_bounds)
        out.append((cfg, riemann_quadrants(cfg, states)))
    return out


def serving_requests(seed, count, rate, precision=PREC_FP32_BULK, rk=2, riemann_sizes=tuple(range(128, 1025, 64)),
                     cylinder_ny=tuple(range(128, 513, 64)), mach=(1.5, 4.0), radius=(0.05, 0.2),
                     t_end_riemann=(0.1, 0.5), t_end_cylinder=(0.1, 0.5)):
    """Seeded request stream: Poisson arrivals at `rate` per second; half
    four-quadrant Riemann problems drawn as in riemann_batch_problems (N x N,
    N ~ U(riemann_sizes), t_end ~ U(t_end_riemann)), half cylinders on
    [0,4] x [0,2] (2N x N, N ~ U(cylinder_ny), inflow rho = 1.4, p = 1, Mach ~
    U(mach), radius ~ U(radius) at (1, 1), t_end ~ U(t_end_cylinder)).  The
    positivity-preserving limiter is on (batches serving them need
    limiter=True).  Returns a list of request dicts (cfg, U, ib, cfl, t_end,
    arrival_s)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    reqs = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate))
        if k % 2 == 0:
            cfg, U = riemann_batch_problems(int(rng.integers(1 << 62)), count=1, sizes=riemann_sizes,
                                            precision=precision, rk=rk)[0]
            reqs.append(dict(cfg=cfg, U=U, ib=None, cfl=0.5, t_end=float(rng.uniform(*t_end_riemann)),
                             arrival_s=t))
        else:
            ny = int(rng.choice(np.asarray(cylinder_ny)))
            m = float(rng.uniform(*mach))
            r = float(rng.uniform(*radius))
            fs = (1.4, m, 0.0, 1.0)
            cfg = config_cylinder(2 * ny, ny, 4.0, 2.0, freestream=fs, rk=rk, precision=precision,
                                  positivity=2, hll_bounds=1)
            ib = circle_setup(cfg, 1.0, 1.0, r, 400, 3, interp=IB_BILINEAR_POSITIVE)
            reqs.append(dict(cfg=cfg, U=cylinder_freestream(cfg, ib["cell_type"], freestream=fs), ib=ib, cfl=0.5,
                             t_end=float(rng.uniform(*t_end_cylinder)), arrival_s=t))
    return reqs
