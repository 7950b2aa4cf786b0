"""GPU parity: the HIP library (through the C-ABI) against the CPU oracle.

Tolerance (the contract): per-field L-infinity difference <= 1e-11 relative to
the field's magnitude after 100 steps.  The FP64 kernels reproduce the
reference's operations in source order with IEEE division/sqrt and no FMA
contraction, so the tests also assert the stronger property that the states
are bit-identical.
"""
import ctypes

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu

TOL = 1e-11


def _assert_parity(gpu, ref, what):
    scale = np.maximum(1.0, np.abs(ref).reshape(4, -1).max(axis=1))
    diff = np.abs(gpu - ref).reshape(4, -1).max(axis=1)
    print("%s: max |diff| per field %s" % (what, diff))
    assert np.all(np.isfinite(gpu)), what
    assert np.all(diff <= TOL * scale), "%s: L-inf %s exceeds %g" % (what, diff, TOL)
    assert np.array_equal(gpu.view(np.uint64), ref.view(np.uint64)), "%s: not bit-identical" % what


def _pair(cfg, U, ib=None, circle=None):
    """GPU solver and oracle on the same input.  With circle given, the oracle
    builds its own immersed-boundary set-up (reference interpolation);
    otherwise it takes the library's tables as input data."""
    s = cfs.Solver(cfg)
    o = O.Oracle(cfg, nthreads=4)
    if ib is not None:
        s.set_immersed(ib)
        if circle is not None:
            o.set_circle(*circle)
        else:
            o.set_ib_table(ib)
    s.set_state(U)
    o.set_state(U)
    return s, o


SUBSONIC = (1.0, 1.0, 0.0, 1.0)   # the reference's cylinder inflow (Cylinder.cpp:5-16)


@pytest.mark.parametrize("rk", [3, 2])
def test_riemann_100_steps(rk):
    cfg = cfs.config_riemann(40, rk=rk)
    U = cfs.riemann_quadrants(cfg)
    s, o = _pair(cfg, U)
    _assert_parity(s.get_state(), o.get_state(), "initial")
    dt = 0.5 / o.cfl_smax()
    for n in range(100):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "riemann rk%d 100 steps" % rk)


def test_stage_by_stage_nonsquare():
    # grid not a multiple of the 15x16 tile, unequal dx/dy, RK3 sub-stages one by one
    cfg = cfs.make_config(37, 29, 1.3, 0.7, rk=3)
    U = cfs.riemann_quadrants(cfg, split=(0.6, 0.3))
    s, o = _pair(cfg, U)
    dt = 0.4 / o.cfl_smax()
    for i in range(1, 10):
        s.stage(i, dt)
        o.stage(i, dt)
        _assert_parity(s.get_state(), o.get_state(), "stage %d" % i)


def test_cylinder_ib_positive_dirichlet_100_steps():
    # reference scheme + the positivity-preserving image-point interpolation
    cfg = cfs.config_cylinder(128, 64, 4.0, 2.0, freestream=SUBSONIC, positivity=0, hll_bounds=0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.25, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=SUBSONIC)
    s, o = _pair(cfg, U, ib)
    _assert_parity(s.get_state(), o.get_state(), "cylinder initial (IB filled)")
    dt = 0.5 / o.cfl_smax()
    for n in range(100):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "cylinder 100 steps")


def test_cylinder_ib_reference_interpolation():
    # the reference's own cylinder set-up: [0,2]^2, r = 0.1 at (0.5, 1), inflow
    # (1, 1, 0, 1), TVD2rdRK, fixed dt at the reference's CFL number; the
    # oracle builds its immersed boundary independently and re-factorises the
    # donor matrix at every fill, as the reference does
    cfg = cfs.config_cylinder(500, 500, 2.0, 2.0, freestream=SUBSONIC, positivity=0, hll_bounds=0, rk=2)
    ib = cfs.circle_setup(cfg, 0.5, 1.0, 0.1, interp=cfs.IB_REFERENCE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=SUBSONIC)
    s, o = _pair(cfg, U, ib, (0.5, 1.0, 0.1, 400, 3))
    _assert_parity(s.get_state(), o.get_state(), "reference cylinder initial (IB filled)")
    dt = 5e-5 * 2000 / 500
    for n in range(30):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "reference cylinder 30 RK2 steps")


@pytest.mark.parametrize("hll_bounds", [0, 1])
def test_positivity_and_hll_bound_options(hll_bounds):
    # Mach 3 inflow exercises the inadmissible-face fallback; a few steps only
    cfg = cfs.config_cylinder(128, 64, 4.0, 2.0, positivity=1, hll_bounds=hll_bounds)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.25, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"])
    s, o = _pair(cfg, U, ib)
    assert s.compute_dt(0.5) == 0.5 / o.cfl_smax()
    dt = 0.5 / o.cfl_smax()
    for n in range(4):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "mach 3, positivity, hll_bounds=%d" % hll_bounds)
    cfg = cfs.make_config(48, 48, 1.0, 1.0, positivity=1, hll_bounds=hll_bounds)
    U = cfs.riemann_quadrants(cfg)
    s, o = _pair(cfg, U)
    dt = 0.5 / o.cfl_smax()
    for n in range(100):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "riemann, positivity, hll_bounds=%d" % hll_bounds)


def test_periodic_conservation():
    cfg = cfs.make_config(48, 32, 1.0, 1.0, bc=(cfs.BC_PERIODIC,) * 4, rk=3)
    x, y = cfg.centres()
    rho = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    u = 0.5 * np.cos(2 * np.pi * y)
    v = 0.1 * np.sin(2 * np.pi * x)
    U = np.stack(cfs.conserved(rho, u, v, 2.5 + 0 * x))
    s, o = _pair(cfg, U)
    c0 = s.checksums()
    assert np.array_equal(c0, o.checksums())
    dt = 0.5 / o.cfl_smax()
    for n in range(50):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "periodic 50 steps")
    c1 = s.checksums()
    assert np.array_equal(c1, o.checksums())
    assert np.allclose(c1, c0, rtol=1e-12, atol=1e-12 * np.abs(c0).max())


def test_cfl_dt_and_advance():
    cfg = cfs.config_riemann(40, rk=3)
    U = cfs.riemann_quadrants(cfg)
    s, o = _pair(cfg, U)
    assert s.compute_dt(0.5) == 0.5 / o.cfl_smax()
    t, done = s.advance(20, dt=0.0, cfl=0.5)
    tt = 0.0
    for n in range(20):
        dt = 0.5 / o.cfl_smax()
        o.step(dt)
        tt += dt
    assert done == 20
    assert t == tt
    _assert_parity(s.get_state(), o.get_state(), "cfl advance")
    # end-time clipping stops exactly at t_end; the launches replayed after it
    # (up to the next read of the halt flag) leave the state as it was: it
    # equals the oracle's stepped with the same clipped time steps
    t_end = t + 0.01
    t2, done2 = s.advance(1000, dt=0.0, cfl=0.5, t_end=t_end, t=t)
    assert abs(t2 - t_end) <= 1e-15 and done2 < 1000
    for n in range(done2):
        dt = 0.5 / o.cfl_smax()
        rem = t_end - tt
        if not (dt < rem):
            dt = rem if rem > 0.0 else 0.0
        o.step(dt)
        tt += dt
    assert tt == t2
    _assert_parity(s.get_state(), o.get_state(), "cfl advance to an end time")


def test_cylinder_cfl_advance_fused_ib_and_bound():
    # CFL steps with the immersed-boundary fill done in the tile load and the
    # bound of each new state accumulated in the last stage; graph replay
    cfg = cfs.config_cylinder(128, 64, 4.0, 2.0, freestream=SUBSONIC, positivity=0, hll_bounds=0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.25, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=SUBSONIC)
    s, o = _pair(cfg, U, ib)
    t, done = s.advance(30, dt=0.0, cfl=0.5)
    tt = 0.0
    for n in range(30):
        dt = 0.5 / o.cfl_smax()
        o.step(dt)
        tt += dt
    assert done == 30 and t == tt
    _assert_parity(s.get_state(), o.get_state(), "cylinder CFL advance")


def test_graph_replay_equals_direct_launches():
    cfg = cfs.config_riemann(48, rk=3)
    U = cfs.riemann_quadrants(cfg)
    a = cfs.Solver(cfg)
    b = cfs.Solver(cfg)
    a.set_state(U)
    b.set_state(U)
    b.set_kernel_timing(True)      # event-bracketed direct launches, no graph
    ta, na = a.advance(11, dt=0.0, cfl=0.5)
    tb, nb = b.advance(11, dt=0.0, cfl=0.5)
    assert (ta, na) == (tb, nb)
    assert b.last_kernel_ms() > 0
    assert np.array_equal(a.get_state(), b.get_state())
    # and again from the new state (cached graph, other parity)
    ta, na = a.advance(6, dt=0.0, cfl=0.5, t=ta)
    tb, nb = b.advance(6, dt=0.0, cfl=0.5, t=tb)
    assert (ta, na) == (tb, nb)
    assert np.array_equal(a.get_state(), b.get_state())


def test_state_sequence_errors():
    cfg = cfs.config_riemann(16)
    s = cfs.Solver(cfg)
    s.set_state(cfs.riemann_quadrants(cfg))
    with pytest.raises(cfs.CfsError):
        s.stage(2, 1e-4)          # counter must be 1 first
    s.stage(1, 1e-4)
    with pytest.raises(cfs.CfsError):
        s.step(1e-4)              # not at a step boundary
    with pytest.raises(cfs.CfsError):
        s.set_state(cfs.riemann_quadrants(cfg))


@pytest.mark.parametrize("patch", ["tiny", "huge"])
def test_exact_division_tiles(patch):
    # values outside the range where the kernel's fast WENO5 divisions are
    # bit-identical to IEEE division (momentum of order 1e-300 in patches, or
    # the whole state scaled by 1e65 > 2^200, a flow the Euler equations map
    # onto the unscaled one): those tiles take the IEEE-division path, are
    # counted, and the states stay bit-identical to the oracle
    cfg = cfs.make_config(64, 48, 1.0, 0.75, rk=3, hll_bounds=1 if patch == "huge" else 0)
    U = cfs.riemann_quadrants(cfg)
    if patch == "tiny":
        U[2, 20:30, 10:20] = 1e-300
        U[1, 40:44, 30:34] = -3e-305
    else:
        U *= 1e65
    s, o = _pair(cfg, U)
    dt = 0.4 / o.cfl_smax()
    counted = []
    for i in range(1, 13):
        s.stage(i, dt)
        o.stage(i, dt)
        counted.append(s.exact_tiles())
        _assert_parity(s.get_state(), o.get_state(), "%s patch, stage %d" % (patch, i))
    print("tiles on the IEEE-division path per stage:", counted)
    assert counted[0] > 0
    # a normal-range state: no tile counted
    cfg = cfs.make_config(64, 48, 1.0, 0.75, rk=3)
    U = cfs.riemann_quadrants(cfg)
    s, o = _pair(cfg, U)
    dt = 0.4 / o.cfl_smax()
    for i in range(1, 7):
        s.stage(i, dt)
        o.stage(i, dt)
        assert s.exact_tiles() == 0
    _assert_parity(s.get_state(), o.get_state(), "normal range")


def test_fp32_bulk_l1_density():
    cfg64 = cfs.config_riemann(64, rk=3)
    cfg32 = cfs.config_riemann(64, rk=3, precision=cfs.PREC_FP32_BULK)
    U = cfs.riemann_quadrants(cfg64)
    o = O.Oracle(cfg64, nthreads=4)
    o.set_state(U)
    s = cfs.Solver(cfg32)
    s.set_state(U)
    dt = 0.5 / o.cfl_smax()
    for n in range(100):
        o.step(dt)
        s.step(dt)
    r64 = o.get_state()[0]
    r32 = s.get_state()[0]
    l1 = np.abs(r32 - r64).mean()
    print("FP32-bulk L1(rho) after 100 steps:", l1)
    assert l1 <= 1e-4


def test_fp32_bulk_l1_density_large_values():
    """The Lax-Liu flow with velocities x 100 and pressures x 1e4 (the same
    flow on a time scale 100 times shorter, but for WENO5's unscaled eps):
    energy jumps of ~3e4 make the FP32 WENO5 smoothness terms (eps + beta)^2
    exceed 2^40 at the discontinuities, where the one-reciprocal product form
    of the weights hands over to the reciprocal form, while smooth regions
    keep the product form."""
    a = 1e2
    states = {k: (r, a * u, a * v, a * a * p) for k, (r, u, v, p) in cfs.LAX_LIU_3.items()}
    cfg64 = cfs.config_riemann(64, rk=3)
    cfg32 = cfs.config_riemann(64, rk=3, precision=cfs.PREC_FP32_BULK)
    U = cfs.riemann_quadrants(cfg64, states=states)
    assert np.abs(U[3]).max() > 3e4
    s64, s32 = cfs.Solver(cfg64), cfs.Solver(cfg32)
    s64.set_state(U)
    s32.set_state(U)
    dt = s64.compute_dt(0.5)
    for n in range(100):
        s64.step(dt)
        s32.step(dt)
    r64 = s64.get_state()[0]
    r32 = s32.get_state()[0]
    assert np.all(np.isfinite(r32))
    l1 = np.abs(r32 - r64).mean()
    print("FP32-bulk L1(rho) after 100 steps, velocities x 100, pressures x 1e4:", l1)
    assert l1 <= 1e-4


@pytest.mark.parametrize("rk", [3, 2])
def test_riemann_400_full_size_100_steps(rk):
    # the Lax-Liu configuration at its full 400^2 size: 100 steps at the CFL
    # time step of the initial state, every field bit-identical (the contract
    # is L-inf <= 1e-11 relative after 100 steps)
    cfg = cfs.config_riemann(400, rk=rk)
    U = cfs.riemann_quadrants(cfg)
    s = cfs.Solver(cfg)
    s.set_state(U)
    o = O.Oracle(cfg, nthreads=8)
    o.set_state(U)
    dt = 0.5 / o.cfl_smax()
    s.advance(100, dt=dt)
    for n in range(100):
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "riemann 400^2 rk%d 100 steps" % rk)


def test_full_size_riemann_symmetry():
    # 400^2, the Lax-Liu configuration: symmetric about x = y, bit for bit
    cfg = cfs.config_riemann(400)
    s = cfs.Solver(cfg)
    s.set_state(cfs.riemann_quadrants(cfg))
    dt = s.compute_dt(0.5)
    s.advance(30, dt=dt)
    V = s.get_state()
    assert np.all(np.isfinite(V))
    assert np.array_equal(V[0], V[0].T)
    assert np.array_equal(V[1], V[2].T)
    assert np.array_equal(V[3], V[3].T)


# ---------------------------------------------------------------------------
# Navier-Stokes terms (viscous kernel variant)

def test_viscous_kelvin_helmholtz_periodic():
    # grid not a multiple of the tile; cross terms read the staged corners
    cfg = cfs.config_kelvin_helmholtz(96, mu=2e-3, rk=3)
    cfg.nx, cfg.ny, cfg.ly = 96, 80, 80.0 / 96
    U = cfs.kelvin_helmholtz(cfg)
    s, o = _pair(cfg, U)
    dt = 0.5 / o.cfl_smax()
    for n in range(30):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "viscous KH 30 steps")


@pytest.mark.parametrize("rk", [2, 3])
def test_viscous_mixed_boundaries_and_corners(rk):
    # Dirichlet on two sides, Neumann on the others: corner cells of the
    # cross-derivative stencils take the x-side ghost first
    inflow = tuple(float(a) for a in cfs.conserved(1.0, 0.8, 0.3, 1.2))
    top = tuple(float(a) for a in cfs.conserved(0.6, -0.2, 0.1, 0.7))
    cfg = cfs.make_config(53, 47, 1.0, 0.9, rk=rk, viscous=1, mu=1e-3, prandtl=0.6, ma_ref=1.5,
                          bc=(cfs.BC_DIRICHLET, cfs.BC_NEUMANN, cfs.BC_NEUMANN, cfs.BC_DIRICHLET),
                          bc_state=(inflow, None, None, top))
    U = cfs.riemann_quadrants(cfg, split=(0.45, 0.55))
    s, o = _pair(cfg, U)
    dt = 0.4 / o.cfl_smax()
    for i in range(1, 7):
        s.stage(i, dt)
        o.stage(i, dt)
        _assert_parity(s.get_state(), o.get_state(), "viscous stage %d" % i)
    for n in range(20):
        s.step(dt)
        o.step(dt)
    _assert_parity(s.get_state(), o.get_state(), "viscous mixed boundaries")


def test_viscous_cylinder_ib_cfl_advance():
    # immersed ghosts in the staged corners; CFL steps through the graph
    cfg = cfs.config_cylinder(96, 48, 4.0, 2.0, freestream=SUBSONIC, positivity=0, hll_bounds=0)
    cfg.viscous, cfg.mu, cfg.prandtl, cfg.ma_ref = 1, 1e-3, 0.72, 0.5
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.3, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=SUBSONIC)
    s, o = _pair(cfg, U, ib)
    t, done = s.advance(20, dt=0.0, cfl=0.5)
    tt = 0.0
    for n in range(20):
        dt = 0.5 / o.cfl_smax()
        o.step(dt)
        tt += dt
    assert done == 20 and t == tt
    _assert_parity(s.get_state(), o.get_state(), "viscous cylinder CFL advance")


def test_viscous_fp32_bulk_l1_density():
    cfg64 = cfs.config_kelvin_helmholtz(128, mu=1e-4)
    cfg32 = cfs.config_kelvin_helmholtz(128, mu=1e-4, precision=cfs.PREC_FP32_BULK)
    U = cfs.kelvin_helmholtz(cfg64)
    s64, s32 = cfs.Solver(cfg64), cfs.Solver(cfg32)
    s64.set_state(U)
    s32.set_state(U)
    dt = s64.compute_dt(0.5)
    for n in range(100):
        s64.step(dt)
        s32.step(dt)
    l1 = np.abs(s32.get_state()[0] - s64.get_state()[0]).mean()
    print("viscous FP32-bulk L1(rho) after 100 steps:", l1)
    assert l1 <= 1e-4


# ---------------------------------------------------------------------------
# positivity-preserving flux limiter (positivity = 2)

@pytest.mark.parametrize("rk", [2, 3])
def test_limiter_random_riemann(rk):
    for cfg, U in cfs.riemann_batch_problems(2024, count=3, sizes=(48, 67), precision=cfs.PREC_FP64, rk=rk,
                                             positivity=2):
        s, o = _pair(cfg, U)
        t, done = s.advance(40, dt=0.0, cfl=0.5)
        for n in range(40):
            o.step(0.5 / o.cfl_smax())
        assert done == 40
        _assert_parity(s.get_state(), o.get_state(), "limiter rk%d %dx%d" % (rk, cfg.nx, cfg.ny))


def test_limiter_mach3_cylinder():
    fs = (1.4, 3.0, 0.0, 1.0)
    cfg = cfs.config_cylinder(128, 64, 4.0, 2.0, freestream=fs, positivity=2, hll_bounds=0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.125, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=fs)
    s, o = _pair(cfg, U, ib)
    t, done = s.advance(60, dt=0.0, cfl=0.5)
    for n in range(60):
        o.step(0.5 / o.cfl_smax())
    assert done == 60
    _assert_parity(s.get_state(), o.get_state(), "limiter Mach-3 cylinder")


def test_limiter_fp32_bulk_l1_density():
    cfg64 = cfs.config_riemann(64, rk=3)
    cfg32 = cfs.config_riemann(64, rk=3, precision=cfs.PREC_FP32_BULK)
    cfg64.positivity = cfg32.positivity = 2
    U = cfs.riemann_quadrants(cfg64)
    o = O.Oracle(cfg64, nthreads=4)
    o.set_state(U)
    s = cfs.Solver(cfg32)
    s.set_state(U)
    dt = 0.5 / o.cfl_smax()
    for n in range(100):
        o.step(dt)
        s.step(dt)
    l1 = np.abs(s.get_state()[0] - o.get_state()[0]).mean()
    print("limiter FP32-bulk L1(rho) after 100 steps:", l1)
    assert l1 <= 1e-4


def test_laxliu_anchor_fixture():
    # tests/golden/laxliu3_32x32_rk3_dt1e-3_20steps.json: the committed
    # bit pattern of 20 RK3 steps, reproduced through the C-ABI
    import test_golden as TG
    g, cfg, U0 = TG.laxliu_anchor()
    s = cfs.Solver(cfg)
    s.set_state(U0)
    s.advance(g["steps"], dt=g["dt"])
    U = s.get_state()
    s.close()
    assert TG.digest(U) == g["sha256"]


def test_cylinder_mach3_full_size_cfl_steps():
    # the headline workload at its full 2048 x 1024 size: Mach-3 cylinder with
    # the positivity limiter and the bilinear immersed-boundary tables, 3 CFL
    # RK3 steps on the device against the oracle (its multi-rank form, 16
    # threads, the same CFL time steps), bit for bit
    fs = (1.4, 3.0, 0.0, 1.0)
    cfg = cfs.config_cylinder(2048, 1024, 4.0, 2.0, freestream=fs, positivity=2, hll_bounds=0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.125, 400, 3, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=fs)
    s = cfs.Solver(cfg)
    s.set_immersed(ib)
    s.set_state(U)
    t, done = s.advance(3, dt=0.0, cfl=0.5)
    o = O.Oracle(cfg, nthreads=16)
    o.set_ranks(16)
    o.set_ib_table(ib)
    o.set_state(U)
    tc = 0.0
    for _ in range(3):
        dt = 0.5 / o.cfl_smax()
        o.step(dt)
        tc += dt
    assert done == 3 and t == tc
    _assert_parity(s.get_state(), o.get_state(), "cylinder 2048x1024 3 CFL steps")
    s.close()
    o.close()


def test_shear_layer_full_size_cfl_step():
    # the viscous shear layer at its full 4096^2 size: one CFL RK3 step (three
    # stages of the viscous kernel) against the oracle, bit for bit
    cfg = cfs.config_kelvin_helmholtz(4096)
    U = cfs.kelvin_helmholtz(cfg)
    s = cfs.Solver(cfg)
    s.set_state(U)
    t, done = s.advance(1, dt=0.0, cfl=0.5)
    o = O.Oracle(cfg, nthreads=16)
    o.set_ranks(16)
    o.set_state(U)
    dt = 0.5 / o.cfl_smax()
    o.step(dt)
    assert done == 1 and t == dt
    _assert_parity(s.get_state(), o.get_state(), "shear layer 4096^2 1 CFL step")
    s.close()
    o.close()


@pytest.mark.parametrize("nx,ny", [(3, 3), (3, 17), (17, 3), (4, 5), (15, 16), (16, 15), (31, 33)])
@pytest.mark.parametrize("shape", [0, 1, 2])
def test_small_and_ragged_grids(nx, ny, shape):
    # the smallest grids the library takes (3 cells, narrower than the 6-cell
    # stencil), an exact 15 x 16 tile and grids ragged across tiles, on every
    # tile shape: 6 RK3 steps with transmissive boundaries, 4 viscous steps
    # (tile shape 0), and 6 steps on periodic boundaries, bit for bit
    U = cfs.riemann_quadrants(cfs.make_config(nx, ny, 1.0, 1.0))
    variants = [dict(), dict(viscous=1, mu=1e-3, prandtl=0.7, ma_ref=1.0), dict(bc=(cfs.BC_PERIODIC,) * 4)]
    for kw in variants:
        if kw.get("viscous") and shape != 0:
            continue   # the viscous kernel comes in tile shape 0 only
        cfg = cfs.make_config(nx, ny, 1.0, 1.0, rk=3, **kw)
        s = cfs.Solver(cfg)
        s.set_tile_shape(shape)
        s.set_state(U)
        o = O.Oracle(cfg, nthreads=1)
        o.set_state(U)
        dt = 0.5 / o.cfl_smax()
        n = 4 if kw.get("viscous") else 6
        s.advance(n, dt=dt)
        for _ in range(n):
            o.step(dt)
        G, R = s.get_state(), o.get_state()
        what = "%dx%d shape %d %s" % (nx, ny, shape, sorted(kw))
        if np.all(np.isfinite(R)):
            _assert_parity(G, R, what)
        else:
            # the scheme itself breaks down (4 x 5 periodic): the same entries
            # are non-finite and the others equal bit for bit
            fin = np.isfinite(R)
            assert np.array_equal(np.isfinite(G), fin), what
            assert np.array_equal(G[fin].view(np.uint64), R[fin].view(np.uint64)), what
        s.close()
        o.close()
