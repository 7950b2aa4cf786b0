"""The oracle's Navier-Stokes terms, pinned by mathematics.

The reference's viscous module (source/gradient/FourthOrderCS.cpp) is never
called by its driver (Run.cpp:177,188), so no reference output exists for it.
The oracle's tables and flux are pinned here instead by:
  * the stencil tables, derived from their moment conditions in exact
    rationals (unique for the given nodes and evaluation points);
  * the viscous face flux of linear velocity/temperature fields, which the
    stencils reproduce exactly: Stokes stresses and Fourier heat flux;
  * decay rates of a shear wave (exp(-4 pi^2 nu t)) in both orientations
    and of an isobaric entropy mode (thermal diffusivity mu/(rho Pr));
  * discrete conservation on a periodic domain.
"""
from fractions import Fraction

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs


def _derivative_weights(nodes, x0, order):
    """Exact weights w with sum w_k f(nodes_k) = f^(order)(x0) for every
    polynomial of degree < len(nodes) (Vandermonde solve in rationals)."""
    n = len(nodes)
    A = [[Fraction(nd - x0) ** m for nd in nodes] for m in range(n)]
    b = [Fraction(0)] * n
    fact = 1
    for m in range(2, order + 1):
        fact *= m
    b[order] = Fraction(fact)
    # Gaussian elimination in exact arithmetic
    M = [row[:] + [b[m]] for m, row in enumerate(A)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [a - f * e for a, e in zip(M[r], M[c])]
    return [M[k][n] / M[k][k] for k in range(n)]


def test_stencil_tables_from_moment_conditions():
    thr, thc, zr = O.viscous_tables()
    half = Fraction(1, 2)
    # x face between cells i-1 and i at 0; stencil cells i-3..i+2 at -5/2..5/2
    cells6 = [Fraction(k) - 3 + half for k in range(6)]
    for r, x0 in enumerate((-1, 0, 1)):   # i-3/2, i-1/2, i+1/2
        w = _derivative_weights(cells6, Fraction(x0), 1)
        assert [float(a) for a in w] == list(thr[r]), "theta_r row %d" % r
        # 4-point interpolation from cells i-2..i+1 (the corrected i-3/2 row included)
        z = _derivative_weights(cells6[1:5], Fraction(x0), 0)
        assert [float(a) for a in z] == list(zr[r]), "zeta_r row %d" % r
    # 5-point central first derivative at the cell centre
    w = _derivative_weights([Fraction(k - 2) for k in range(5)], Fraction(0), 1)
    assert [float(a) for a in w] == list(thc)
    # the reference's i-3/2 interpolation row (FourthOrderCS.h:40) has zeroth moment 3/8, not 1
    assert sum(Fraction(a) for a in zr[0]) == 1


def _linear_state(cfg, a, b, c, d, e, f, u0=0.3, v0=-0.2, T0=1.5):
    """rho = 1; u = u0 + a x + b y; v = v0 + c x + d y; T = T0 + e x + f y."""
    x, y = cfg.centres()
    u = u0 + a * x + b * y
    v = v0 + c * x + d * y
    T = T0 + e * x + f * y
    rho = np.ones_like(x)
    p = rho * T / (cfg.gamma * cfg.ma_ref ** 2)     # T = gamma Ma^2 p / rho
    return np.stack(cfs.conserved(rho, u, v, p, gamma=cfg.gamma))


def _fluxes(cfg, U):
    o = O.Oracle(cfg)
    o.set_state(U)
    o.stage(1, 0.0)
    return o.face_flux()


@pytest.mark.parametrize("mu,pr,ma", [(0.03, 0.72, 1.0), (0.011, 0.5, 2.0)])
def test_viscous_flux_of_linear_fields(mu, pr, ma):
    nx, ny, lx, ly = 24, 20, 1.2, 0.8
    a, b, c, d, e, f = 0.7, -0.4, 0.25, -0.9, 0.6, -0.35
    kw = dict(bc=(cfs.BC_NEUMANN,) * 4, rk=2, mu=mu, prandtl=pr, ma_ref=ma)
    inv = cfs.make_config(nx, ny, lx, ly, viscous=0, **kw)
    vis = cfs.make_config(nx, ny, lx, ly, viscous=1, **kw)
    U = _linear_state(vis, a, b, c, d, e, f)
    fx0, fy0 = _fluxes(inv, U)
    fx1, fy1 = _fluxes(vis, U)
    gvx, gvy = fx0 - fx1, fy0 - fy1           # net flux = inviscid - viscous
    kappa = mu / ((vis.gamma - 1) * ma * ma * pr)
    dx, dy = lx / nx, ly / ny
    # interior faces whose stencils and cross-gradient stencils stay inside
    for i in range(4, nx - 3):
        for j in range(3, ny - 3):
            xf, yc = i * dx, (j + 0.5) * dy
            u, v = 0.3 + a * xf + b * yc, -0.2 + c * xf + d * yc
            txx = mu * (4.0 / 3.0 * a - 2.0 / 3.0 * d)
            txy = mu * (b + c)
            want = [0.0, txx, txy, u * txx + v * txy + kappa * e]
            np.testing.assert_allclose(gvx[i, j], want, rtol=0, atol=1e-11)
    for i in range(3, nx - 3):
        for j in range(4, ny - 3):
            xc, yf = (i + 0.5) * dx, j * dy
            u, v = 0.3 + a * xc + b * yf, -0.2 + c * xc + d * yf
            tyy = mu * (4.0 / 3.0 * d - 2.0 / 3.0 * a)
            txy = mu * (b + c)
            want = [0.0, txy, tyy, u * txy + v * tyy + kappa * f]
            np.testing.assert_allclose(gvy[i, j], want, rtol=0, atol=1e-11)


def _run(cfg, U, dt, nsteps):
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(U)
    for _ in range(nsteps):
        o.step(dt)
    return o.get_state()


def _mode_amplitude(f, k_axis):
    """Amplitude of the first Fourier sine/cosine mode of f along an axis."""
    F = np.fft.rfft(f, axis=k_axis)
    return np.abs(np.take(F, 1, axis=k_axis)).mean() * 2.0 / f.shape[k_axis]


@pytest.mark.parametrize("orientation", ["u(y)", "v(x)"])
def test_shear_wave_decay(orientation):
    n, mu, A, t_end = 32, 0.02, 1e-3, 0.25
    cfg = cfs.make_config(n, n, 1.0, 1.0, bc=(cfs.BC_PERIODIC,) * 4, rk=2, viscous=1, mu=mu)
    x, y = cfg.centres()
    rho = np.ones_like(x)
    p = np.full_like(x, 1.0 / cfg.gamma)   # sound speed 1
    if orientation == "u(y)":
        u, v, axis, comp = A * np.sin(2 * np.pi * y), 0 * x, 1, 1
    else:
        u, v, axis, comp = 0 * x, A * np.sin(2 * np.pi * x), 0, 2
    U0 = np.stack(cfs.conserved(rho, u, v, p, gamma=cfg.gamma))
    nsteps = 50
    U1 = _run(cfg, U0, t_end / nsteps, nsteps)
    ratio = _mode_amplitude(U1[comp], axis) / _mode_amplitude(U0[comp], axis)
    rate = -np.log(ratio) / t_end
    exact = 4 * np.pi ** 2 * mu        # nu = mu / rho, rho = 1
    print(orientation, "decay rate", rate, "exact", exact)
    assert abs(rate - exact) <= 0.01 * exact
    # the other velocity component stays at rest up to the motion driven by
    # viscous heating, which is second order in the amplitude
    other = U1[3 - comp]
    assert np.abs(other).max() <= 1e-4 * A


def test_entropy_mode_thermal_diffusion():
    # isobaric temperature wave at rest: T decays at the thermal diffusivity
    # alpha = kappa / (rho c_p) = mu / (rho Pr), independent of Ma
    n, mu, pr, eps, t_end = 32, 0.005, 0.5, 1e-3, 1.0
    cfg = cfs.make_config(n, 4, 1.0, 4.0 / n, bc=(cfs.BC_PERIODIC,) * 4, rk=2, viscous=1, mu=mu,
                          prandtl=pr, ma_ref=1.0)
    x, _ = cfg.centres()
    T = 1.0 + eps * np.sin(2 * np.pi * x)
    p = np.full_like(x, 1.0 / cfg.gamma)
    rho = cfg.gamma * p / T              # T = gamma Ma^2 p / rho
    U0 = np.stack(cfs.conserved(rho, 0 * x, 0 * x, p, gamma=cfg.gamma))
    nsteps = 200
    U1 = _run(cfg, U0, t_end / nsteps, nsteps)

    def temperature(U):
        r = U[0]
        pp = (cfg.gamma - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / r)
        return cfg.gamma * pp / r

    ratio = _mode_amplitude(temperature(U1), 0) / _mode_amplitude(temperature(U0), 0)
    rate = -np.log(ratio) / t_end
    exact = mu / pr * (2 * np.pi) ** 2   # rho ~ 1
    print("thermal decay rate", rate, "exact", exact)
    assert abs(rate - exact) <= 0.03 * exact


def test_periodic_viscous_conservation():
    cfg = cfs.config_kelvin_helmholtz(32, mu=1e-3, rk=3)
    U0 = cfs.kelvin_helmholtz(cfg)
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(U0)
    c0 = o.checksums()
    dt = 0.5 / o.cfl_smax()
    for _ in range(10):
        o.step(dt)
    c1 = o.checksums()
    np.testing.assert_allclose(c1, c0, rtol=0, atol=1e-12 * np.abs(c0).max())


def test_kelvin_helmholtz_initial_state():
    # SplitMix64 reference outputs for state 0: first two draws
    assert int(cfs.splitmix64(0)) == 0xE220A8397B1DCDAF
    assert int(cfs.splitmix64(0x9E3779B97F4A7C15)) == 0x6E789E6AA1B965F4
    cfg = cfs.config_kelvin_helmholtz(64)
    U = cfs.kelvin_helmholtz(cfg)
    x, y = cfg.centres()
    band = (y > 0.25) & (y < 0.75)
    assert np.all(U[0][band] == 2.0) and np.all(U[0][~band] == 1.0)
    np.testing.assert_array_equal(U[1] / U[0], np.where(band, 0.5, -0.5))
    noise = U[2] / U[0] - 0.01 * np.sin(4 * np.pi * x)
    assert np.abs(noise).max() <= 1e-3 and np.abs(noise).max() > 0.9e-3
    assert abs(noise.mean()) < 1e-4
    # same seed, same field; another seed, another field
    np.testing.assert_array_equal(U, cfs.kelvin_helmholtz(cfg))
    assert not np.array_equal(U, cfs.kelvin_helmholtz(cfg, seed=1))
