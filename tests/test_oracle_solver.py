"""Pins the oracle's full stage (stencils + WENO + HLL + RHS + RK + boundary
fill) by invariants the mathematics fixes:

  - the Lax-Liu configuration 3 is symmetric about x = y: with dx == dy the x
    and y code paths are mirror images, so rho(i,j) == rho(j,i) bit for bit;
  - periodic boundaries: the flux differences telescope, the totals stay
    constant to round-off;
  - a uniform flow with transmissive or matching-Dirichlet boundaries stays
    uniform;
  - the CFL reduction equals a brute-force evaluation of its definition;
  - with transmissive boundaries the summed right-hand side equals the net
    flux through the boundary faces (the interior fluxes telescope).
"""
import math

import numpy as np

import _oracle as O

cfs = O.cfs


def test_lax_liu_symmetry_bitwise():
    cfg = cfs.config_riemann(64)
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(cfs.riemann_quadrants(cfg))
    dt = 0.5 / o.cfl_smax()
    for _ in range(25):
        o.step(dt)
    V = o.get_state()
    assert np.all(np.isfinite(V))
    assert np.array_equal(V[0], V[0].T)
    assert np.array_equal(V[1], V[2].T)
    assert np.array_equal(V[3], V[3].T)


def test_periodic_conservation():
    cfg = cfs.make_config(40, 24, 1.0, 0.6, bc=(cfs.BC_PERIODIC,) * 4, rk=3)
    x, y = cfg.centres()
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y / 0.6)
    U = np.stack(cfs.conserved(rho, 0.4 + 0 * x, -0.3 + 0 * x, 1.0 + 0.2 * np.cos(2 * np.pi * x)))
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(U)
    c0 = o.checksums()
    dt = 0.5 / o.cfl_smax()
    for _ in range(40):
        o.step(dt)
    c1 = o.checksums()
    assert np.all(np.abs(c1 - c0) <= 1e-12 * np.abs(c0).max())
    # and the state did change
    assert np.abs(o.get_state() - U).max() > 1e-3


def test_uniform_flow_is_preserved():
    fs = (1.4, 3.0, 0.5, 1.0)
    cfg = cfs.config_cylinder(48, 32, 2.0, 1.0, freestream=fs, positivity=0, hll_bounds=0)
    U = cfs.cylinder_freestream(cfg, np.zeros(48 * 32, np.uint8), freestream=fs)
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(U)
    for _ in range(10):
        o.step(1e-3)
    assert np.abs(o.get_state() - U).max() <= 1e-13


def test_cfl_bound_brute_force():
    cfg = cfs.config_riemann(30)
    U = cfs.riemann_quadrants(cfg)
    o = O.Oracle(cfg)
    o.set_state(U)
    rho, mu, mv, E = U
    p = (1.4 - 1) * (E - 0.5 * (mu * mu + mv * mv) / rho)
    c = np.sqrt(1.4 * p / rho)
    ax = np.maximum(np.abs(mu / rho), np.abs(mu))
    ay = np.maximum(np.abs(mv / rho), np.abs(mv))
    s = ((ax + c) / cfg.dx + (ay + c) / cfg.dy).max()
    assert o.cfl_smax() == s


def test_stage_sequence_enforced():
    cfg = cfs.config_riemann(8)
    o = O.Oracle(cfg)
    o.set_state(cfs.riemann_quadrants(cfg))
    import pytest
    with pytest.raises(RuntimeError):
        o.stage(2, 1e-3)


def test_rhs_sum_equals_boundary_flux():
    # Cell.cpp:110-125: L = -(F_{i+1/2} - F_{i-1/2})/dx - (G_{j+1/2} - G_{j-1/2})/dy;
    # summed over the grid only the boundary faces remain.  A wrong sign, a
    # shifted face index or an unmatched interior flux breaks the identity.
    cfg = cfs.make_config(37, 29, 1.0, 0.8, rk=3)
    o = O.Oracle(cfg, nthreads=2)
    rng = np.random.default_rng(5)
    U = cfs.riemann_quadrants(cfs.config_riemann(64))[:, :37, :29].copy()
    U *= 1.0 + 0.05 * rng.random(U.shape)
    o.set_state(U)
    dt = 0.5 / o.cfl_smax()
    for stage in (1, 2, 3):
        o.stage(stage, dt)
        L = o.get_rhs()
        fx, fy = o.face_flux()
        dx, dy = 1.0 / 37, 0.8 / 29
        for q in range(4):
            lhs = math.fsum(L[q].ravel())
            rhs = -(math.fsum(fx[-1, :, q]) - math.fsum(fx[0, :, q])) / dx \
                  - (math.fsum(fy[:, -1, q]) - math.fsum(fy[:, 0, q])) / dy
            scale = np.abs(L[q]).sum()
            assert abs(lhs - rhs) <= 1e-12 * scale, (stage, q, lhs, rhs)
