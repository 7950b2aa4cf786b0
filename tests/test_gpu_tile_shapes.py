"""Tile shapes and the autotuner on the GPU.  The tile shape decides which
workgroup computes a face or a cell, never how: every shape must give the
oracle's results bit for bit, single solves and batches alike, and the
autotuners must leave the solve untouched."""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu

SHAPES = list(range(cfs.NUM_TILE_SHAPES))


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _cylinder(nx, ny, precision=cfs.PREC_FP64):
    fs = (1.4, 3.0, 0.0, 1.0)
    cfg = cfs.config_cylinder(nx, ny, 4.0, 2.0, freestream=fs, positivity=2, hll_bounds=0, precision=precision)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.25, interp=cfs.IB_BILINEAR_POSITIVE)
    return cfg, ib, cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=fs)


def test_shape_info():
    for k in SHAPES:
        ti, tj, nt = cfs.tile_shape_info(k)
        assert ti * tj <= nt and (ti + 1) * tj + ti * (tj + 1) <= 2 * nt and nt % 64 == 0
    with pytest.raises(cfs.CfsError):
        cfs.tile_shape_info(cfs.NUM_TILE_SHAPES)


@pytest.mark.parametrize("shape", SHAPES)
def test_shape_riemann_matches_oracle(shape):
    # a grid that is a multiple of no tile side, RK3, reference scheme
    cfg = cfs.make_config(70, 53, 1.0, 53 / 70, rk=3)
    U = cfs.riemann_quadrants(cfg)
    s = cfs.Solver(cfg)
    s.set_tile_shape(shape)
    assert s.tile_shape == shape
    s.set_state(U)
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(U)
    dt = 0.4 / o.cfl_smax()
    for _ in range(20):
        s.step(dt)
        o.step(dt)
    assert np.array_equal(_bits(s.get_state()), _bits(o.get_state())), "shape %d" % shape


@pytest.mark.parametrize("shape", SHAPES)
def test_shape_cylinder_limiter_cfl_matches_oracle(shape):
    # immersed boundary (per-shape tile flags), limiter, device CFL steps
    cfg, ib, U = _cylinder(150, 70)
    s = cfs.Solver(cfg)
    s.set_immersed(ib)
    s.set_tile_shape(shape)
    s.set_state(U)
    o = O.Oracle(cfg, nthreads=4)
    o.set_ib_table(ib)
    o.set_state(U)
    t, done = s.advance(30, dt=0.0, cfl=0.5)
    for _ in range(30):
        o.step(0.5 / o.cfl_smax())
    assert done == 30
    assert np.array_equal(_bits(s.get_state()), _bits(o.get_state())), "shape %d" % shape


def test_autotune_leaves_the_solve_unchanged():
    cfg, ib, U = _cylinder(256, 128)
    a = cfs.Solver(cfg)
    a.set_immersed(ib)
    a.set_state(U)
    a.advance(4, dt=0.0, cfl=0.5)
    before = a.get_state()
    best, ms = a.autotune(reps=5)
    print("autotune: best shape", best, "ms per launch", ms)
    assert best in SHAPES and a.tile_shape == best
    assert all(m > 0 for m in ms) and ms[best] == min(ms)
    assert np.array_equal(_bits(a.get_state()), _bits(before))
    ta, na = a.advance(8, dt=0.0, cfl=0.5)
    # the same solve without tuning
    b = cfs.Solver(cfg)
    b.set_immersed(ib)
    b.set_state(U)
    b.advance(4, dt=0.0, cfl=0.5)
    tb, nb = b.advance(8, dt=0.0, cfl=0.5)
    assert (ta, na) == (tb, nb)
    assert np.array_equal(_bits(a.get_state()), _bits(b.get_state()))
    # mid-step the spare plane-sets hold stage data: refused
    a.stage(a.counter, 1e-4)
    with pytest.raises(cfs.CfsError):
        a.autotune(reps=2)


def test_autotune_viscous_has_one_shape():
    cfg = cfs.config_kelvin_helmholtz(64)
    s = cfs.Solver(cfg)
    s.set_state(cfs.kelvin_helmholtz(cfg))
    best, ms = s.autotune(reps=2)
    assert best == 0 and ms[0] > 0 and all(m == -1.0 for m in ms[1:])
    with pytest.raises(cfs.CfsError):
        s.set_tile_shape(1)


@pytest.mark.parametrize("shape", SHAPES)
def test_batch_shape_equals_alone(shape):
    probs = []
    for k, (cfg, U) in enumerate(cfs.riemann_batch_problems(5, count=4, sizes=(37, 64, 96),
                                                            precision=cfs.PREC_FP64, rk=3)):
        probs.append((cfg, U, None, 0.5, 0.01 + 0.003 * k))
    cfg, ib, U = _cylinder(96, 48)
    probs.append((cfg, U, ib, 0.5, 0.05))
    got = {}
    with cfs.Batch(rk=3, precision=cfs.PREC_FP64, max_problems=8, arena_bytes=128 << 20, limiter=True,
                   tile_shape=shape) as b:
        slots = {b.admit(c, u, cfl=cfl, t_end=te, ib=i): k for k, (c, u, i, cfl, te) in enumerate(probs)}
        for _ in range(400):
            b.step(4)
            for slot, (state, steps, t) in b.poll().items():
                assert state != cfs.SLOT_FAILED
                if state == cfs.SLOT_DONE:
                    got[slots[slot]] = (b.retire(slot), t, steps)
            if len(got) == len(probs):
                break
    assert len(got) == len(probs)
    for k, (c, u, i, cfl, te) in enumerate(probs):
        s = cfs.Solver(c)
        if i is not None:
            s.set_immersed(i)
        s.set_state(u)
        t, done = s.advance(10000, dt=0.0, cfl=cfl, t_end=te)
        Ub, tb, steps = got[k]
        assert steps == done and tb == t
        assert np.array_equal(_bits(Ub), _bits(s.get_state())), "shape %d problem %d" % (shape, k)


def test_batch_autotune():
    probs = [(c, u) for c, u in cfs.riemann_batch_problems(9, count=12, sizes=(64, 128, 192),
                                                           precision=cfs.PREC_FP32_BULK, rk=2)]
    best, ms = cfs.batch_autotune(probs, rk=2, precision=cfs.PREC_FP32_BULK, limiter=True, steps=6,
                                  arena_bytes=256 << 20)
    print("batch autotune: best shape", best, "ms per step", ms)
    assert best in SHAPES and all(m > 0 for m in ms) and ms[best] == min(ms)
