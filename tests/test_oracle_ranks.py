"""The oracle's multi-rank mode (the reference's MPI CPU path restated:
2D block decomposition, CartesianParallel.cpp:3-47; 3-deep ghost frames
refreshed every stage, :60-427) against the plain oracle.  A decomposition
changes which worker computes a face, never the values it reads, so every
decomposition must give the same state bit for bit: physical and periodic
boundaries, the immersed boundary, the limiter and the viscous terms."""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_dims_follow_mpi_dims_create():
    cfg = cfs.config_riemann(64)
    o = O.Oracle(cfg)
    # MPI_Dims_create(n, 2): the closest factor pair, larger first
    assert [o.set_ranks(n) for n in (1, 2, 4, 6, 8, 12, 16, 7)] == [
        (1, 1), (2, 1), (2, 2), (3, 2), (4, 2), (4, 3), (4, 4), (7, 1)]


def _compare(cfg, U, ranks, nsteps, ib=None, cfl=0.4, fixed_dt=None):
    a = O.Oracle(cfg, nthreads=1)
    b = O.Oracle(cfg, nthreads=1)
    b.set_ranks(ranks)
    for o in (a, b):
        if ib is not None:
            o.set_ib_table(ib)
        o.set_state(U)
    for _ in range(nsteps):
        sa, sb = a.cfl_smax(), b.cfl_smax()
        assert sa == sb
        dt = fixed_dt if fixed_dt is not None else cfl / sa
        a.step(dt)
        b.step(dt)
    Sa, Sb = a.get_state(), b.get_state()
    assert np.all(np.isfinite(Sa))
    assert np.array_equal(_bits(Sa), _bits(Sb)), "ranks=%d" % ranks
    fa, fb = a.face_flux(), b.face_flux()
    assert np.array_equal(_bits(fa[0]), _bits(fb[0])) and np.array_equal(_bits(fa[1]), _bits(fb[1]))


@pytest.mark.parametrize("ranks", [2, 4, 6, 9])
def test_riemann_decomposition_invariant(ranks):
    # 50 x 37: blocks of uneven size, neither side divisible by every decomposition
    cfg = cfs.make_config(50, 37, 1.0, 0.74, rk=3)
    _compare(cfg, cfs.riemann_quadrants(cfg), ranks, 8)


@pytest.mark.parametrize("ranks", [4, 8])
def test_cylinder_ib_limiter_decomposition_invariant(ranks):
    fs = (1.4, 3.0, 0.0, 1.0)
    cfg = cfs.config_cylinder(96, 48, 4.0, 2.0, freestream=fs, positivity=2, hll_bounds=0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.25, interp=cfs.IB_BILINEAR_POSITIVE)
    _compare(cfg, cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=fs), ranks, 10, ib=ib, cfl=0.5)


@pytest.mark.parametrize("ranks", [4, 6])
def test_viscous_periodic_decomposition_invariant(ranks):
    cfg = cfs.config_kelvin_helmholtz(48)
    _compare(cfg, cfs.kelvin_helmholtz(cfg), ranks, 6)


def test_reference_interpolation_rk2_decomposition_invariant():
    # the reference's own interpolation (it turns the flow non-finite within a
    # few steps, so one step) and its own RK2 driver scheme
    cfg = cfs.config_cylinder(100, 50, 4.0, 2.0, freestream=(1.0, 1.0, 0.0, 1.0), rk=2, positivity=0,
                              hll_bounds=0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.2, interp=cfs.IB_REFERENCE)
    _compare(cfg, cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=(1.0, 1.0, 0.0, 1.0)), 4, 1, ib=ib,
             fixed_dt=5e-4)


def test_rank_count_larger_than_the_grid_is_refused():
    cfg = cfs.make_config(4, 4, 1.0, 1.0)
    o = O.Oracle(cfg)
    with pytest.raises(RuntimeError):
        o.set_ranks(7)   # 7 x 1 blocks on 4 cells
