"""Pins the oracle's scalar building blocks by facts the mathematics fixes,
hand-derived known answers and the reference's documented quirks -- not by
re-typing its formulas.

  weno5           source/interpolation/Weno5th.cpp:4-28
  Euler flux      source/model/EulerMod.cpp:33-54
  HLL             source/flux/RiemmanSolverHLL.cpp:5-75
  face walk       source/meshes/CellInterface.cpp:135-195
  RK stages       source/timeadvance/TVD2rdRK.cpp:3-31, TVD3rdRK.cpp:3-48
  checksum tree   (no reference counterpart)
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import _oracle as O

G = 1.4


# ---- WENO5 ------------------------------------------------------------------
def test_weno_constant():
    for c in (1.0, -3.7, 1e-3, 12345.678):
        r = O.weno5([c] * 5)
        assert abs(r - c) <= 4 * np.spacing(abs(c)), (c, r)


@pytest.mark.parametrize("a,b,c", [(1.0, 0.0, 0.0), (0.3, -1.2, 2.0), (-2.0, 0.5, 1.0)])
def test_weno_reproduces_quadratics(a, b, c):
    # cell averages of p(x) = a x^2 + b x + c over unit cells centred at 0..4;
    # every candidate stencil is exact for quadratics, so whatever the nonlinear
    # weights are, the value at the face x = 2.5 is p(2.5)
    avg = [a * (k * k + 1.0 / 12.0) + b * k + c for k in range(5)]
    r = O.weno5(avg)
    exact = a * 2.5 ** 2 + b * 2.5 + c
    assert abs(r - exact) <= 1e-13 * max(1.0, abs(exact))


def test_weno_known_answer_pins_eps():
    # v = (0,0,0,0,1): beta0 = beta1 = 0, beta2 = 13/12 + 1/4 = 4/3;
    # alpha = (0.1/eps^2, 0.6/eps^2, 0.3/(eps+4/3)^2) with eps = 1/100;
    # q2 = -1/6, q0 = q1 = 0  =>  value = -w2/6 = -9/2273780 exactly
    exact = Fraction(-9, 2273780)
    r = O.weno5([0.0, 0.0, 0.0, 0.0, 1.0])
    assert abs(Fraction(r) - exact) <= abs(exact) * Fraction(1, 10 ** 14)
    # with the textbook eps = 1e-6 the value would be ~1e-18: eps = 1e-2 is pinned


def test_weno_non_oscillatory_at_a_step():
    # upwind side smooth: the face value stays next to the upwind plateau
    r = O.weno5([0.0, 0.0, 0.0, 1.0, 1.0])
    assert 0.0 <= r < 1e-3
    r = O.weno5([1.0, 1.0, 1.0, 0.0, 0.0])
    assert 1.0 - 1e-3 < r <= 1.0



def _weno5_exact(v):
    """Weno5th.cpp:4-28 evaluated in exact rational arithmetic on the FP64
    constants (eps = 1e-2, d = 0.1, 0.6, 0.3, 13/12 as doubles): the FP64
    evaluation may differ from it only by rounding."""
    F = Fraction
    v0, v1, v2, v3, v4 = (F(x) for x in v)
    c13, eps = F(13.0 / 12.0), F(1e-2)
    b0 = c13 * (v0 - 2 * v1 + v2) ** 2 + F(1, 4) * (v0 - 4 * v1 + 3 * v2) ** 2
    b1 = c13 * (v1 - 2 * v2 + v3) ** 2 + F(1, 4) * (v1 - v3) ** 2
    b2 = c13 * (v2 - 2 * v3 + v4) ** 2 + F(1, 4) * (3 * v2 - 4 * v3 + v4) ** 2
    a = [F(0.1) / (eps + b0) ** 2, F(0.6) / (eps + b1) ** 2, F(0.3) / (eps + b2) ** 2]
    q = [v0 / 3 - 7 * v1 / 6 + 11 * v2 / 6, -v1 / 6 + 5 * v2 / 6 + v3 / 3, v2 / 3 + 5 * v3 / 6 - v4 / 6]
    return sum(ai * qi for ai, qi in zip(a, q)) / sum(a)


def test_weno_within_rounding_of_exact_rational_value():
    # a dropped term, a wrong sign or a swapped coefficient moves the result by
    # far more than rounding: the bound is 64 ulp of the largest input
    rng = np.random.default_rng(11)
    for k in range(300):
        scale = 10.0 ** rng.integers(-3, 4)
        v = (rng.normal(size=5) * scale).tolist()
        if k % 3 == 0:                    # near-smooth data: small beta
            v = (v[0] + 1e-3 * scale * np.arange(5) ** 2).tolist()
        r = O.weno5(v)
        exact = _weno5_exact(v)
        tol = 64 * np.spacing(max(abs(x) for x in v))
        assert abs(Fraction(r) - exact) <= Fraction(tol), (v, r, float(exact))


def test_weno_fifth_order_on_smooth_data():
    # cell averages of sin(x) on cells of width h (2 sin(c) sin(h/2) / h, free
    # of cancellation); face value at the right edge of the middle cell; with
    # eps = 1e-2 >> beta ~ h^2 the weights approach the linear ones, so the
    # error falls as h^5 (ratio 32 per halving).  A scheme of order 3 (a wrong
    # linear weight) or 4 fails the lower bound.
    errs = []
    x0 = 0.3
    for h in (0.025, 0.0125, 0.00625):
        centres = x0 + h * (np.arange(5) - 2)
        avg = [2.0 * math.sin(c) * math.sin(h / 2) / h for c in centres]
        errs.append(abs(O.weno5(avg) - math.sin(x0 + h / 2)))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(4.8 < o < 5.5 for o in orders), (errs, orders)



def _weno5_source_order(v):
    """Weno5th.cpp:4-28 evaluated in Python floats (IEEE double, round to
    nearest, no fused operations) with the source's operations, association
    and order as written: an evaluation independent of the oracle's code."""
    eps = 1e-2
    beta0 = (13.0 / 12.0) * ((v[0] - 2 * v[1] + v[2]) * (v[0] - 2 * v[1] + v[2])) \
        + 0.25 * ((v[0] - 4 * v[1] + 3 * v[2]) * (v[0] - 4 * v[1] + 3 * v[2]))
    beta1 = (13.0 / 12.0) * ((v[1] - 2 * v[2] + v[3]) * (v[1] - 2 * v[2] + v[3])) \
        + 0.25 * ((v[1] - v[3]) * (v[1] - v[3]))
    beta2 = (13.0 / 12.0) * ((v[2] - 2 * v[3] + v[4]) * (v[2] - 2 * v[3] + v[4])) \
        + 0.25 * ((3 * v[2] - 4 * v[3] + v[4]) * (3 * v[2] - 4 * v[3] + v[4]))
    alpha0 = 0.1 / ((eps + beta0) * (eps + beta0))
    alpha1 = 0.6 / ((eps + beta1) * (eps + beta1))
    alpha2 = 0.3 / ((eps + beta2) * (eps + beta2))
    sum_alpha = alpha0 + alpha1 + alpha2
    omega0 = alpha0 / sum_alpha
    omega1 = alpha1 / sum_alpha
    omega2 = alpha2 / sum_alpha
    return (omega0 * (v[0] / 3.0 - 7.0 * v[1] / 6.0 + 11.0 * v[2] / 6.0)
            + omega1 * (-v[1] / 6.0 + 5.0 * v[2] / 6.0 + v[3] / 3.0)
            + omega2 * (v[2] / 3.0 + 5.0 * v[3] / 6.0 - v[4] / 6.0))


def test_weno_bitwise_equal_to_the_source_order():
    # a re-associated sum or a fused operation moves the result by an ulp or
    # more on some of these inputs, which the 64-ulp rational bound lets pass
    rng = np.random.default_rng(23)
    n_diff = 0
    for k in range(2000):
        scale = 10.0 ** rng.integers(-4, 5)
        v = (rng.normal(size=5) * scale).tolist()
        if k % 4 == 0:
            v = (v[0] + 1e-3 * scale * np.arange(5) ** 2).tolist()
        if k % 7 == 0:
            v[rng.integers(0, 5)] = 0.0
        a, b = O.weno5(v), _weno5_source_order(v)
        n_diff += math.copysign(1.0, a) != math.copysign(1.0, b) or a != b
    assert n_diff == 0, n_diff


# ---- Euler flux and HLL ------------------------------------------------------
def test_euler_flux_known_answer():
    S = [2.0, 1.0, 0.5, 5.0]           # rho, rho u, rho v, E
    p = 0.4 * (5.0 - 0.5 * (1.0 + 0.25) / 2.0)   # 1.875
    fx = O.euler_flux(G, S, True)
    fy = O.euler_flux(G, S, False)
    assert np.allclose(fx, [1.0, 0.5 + p, 0.25, (5.0 + p) * 0.5], rtol=1e-15)
    assert np.allclose(fy, [0.5, 0.25, 0.125 + p, (5.0 + p) * 0.25], rtol=1e-15)
    assert O.pressure(G, S) == pytest.approx(p, rel=1e-15)


def _cons(rho, u, v, p):
    return [rho, rho * u, rho * v, p / (G - 1) + 0.5 * (rho * u * rho * u + rho * v * rho * v) / rho]


def test_hll_consistency():
    # identical states: the HLL average collapses to the physical flux
    for st in (_cons(0.7, 0.3, -0.2, 1.1), _cons(1.3, -0.4, 0.9, 0.6)):
        for xdir in (True, False):
            f = O.hll(G, st, st, xdir)
            assert np.allclose(f, O.euler_flux(G, st, xdir), rtol=1e-14, atol=1e-14)


def test_hll_supersonic_upwind_is_exact():
    # rho = 1: momentum == velocity, u - c > 0 -> left flux, bit for bit
    L = _cons(1.0, 3.0, 0.1, 1.0)
    R = _cons(1.0, 2.5, 0.0, 0.8)
    assert np.array_equal(O.hll(G, L, R, True), O.euler_flux(G, L, True))


def test_hll_uses_momentum_in_outer_bounds():
    # Same velocities, density 0.2: with velocities SL = u - c > 0 (upwind
    # flux); the reference bounds use the momentum 0.2*u, so SL < 0 and the
    # HLL average is taken instead.
    rho = 0.2
    L = _cons(rho, 3.0, 0.0, 0.2 / 1.4)    # c = 1
    R = _cons(rho, 3.0, 0.0, 0.2 / 1.4)
    u, c, m = 3.0, 1.0, rho * 3.0
    assert u - c > 0 and m - c < 0
    R2 = _cons(rho, 3.2, 0.0, 0.2 / 1.4)
    f = O.hll(G, L, R2, True)
    assert not np.array_equal(f, O.euler_flux(G, L, True))
    # with rho = 1 the same velocity field is upwinded exactly
    L1, R1 = _cons(1.0, 3.0, 0.0, 1.0 / 1.4), _cons(1.0, 3.2, 0.0, 1.0 / 1.4)
    assert np.array_equal(O.hll(G, L1, R1, True), O.euler_flux(G, L1, True))
    del R


def test_hll_right_bound_takes_the_left_state():
    # Known answer for RiemmanSolverHLL.cpp:41-42, where the outer right bound
    # is SR = max(m_L + a_L, u* + a*_R): the LEFT state's momentum and sound
    # speed, where the textbook bound takes the right state's.  Hand
    # derivation, rho = 1 on both sides (momentum = velocity), p = 2, so
    # a = sqrt(1.4 * 2) = 1.67332; u_L = -2, u_R = -1.5:
    #   PVRS: u* = (u_L + u_R)/2 = -1.75 (equal pressures),
    #         p* = 2 - (u_R - u_L)/2 * rho * a = 1.58167 > 0,
    #         rho*_R = 1 - (u_R - u*) rho / a = 0.85060,
    #         a*_R = sqrt(1.4 p* / rho*_R) = 1.61347;
    #   u_L + a_L = -0.32668 and u* + a*_R = -0.13653: SR = -0.13653 <= 0, so
    #   the right state's upwind flux F(R) is taken, bit for bit;
    #   with the right state in SR, u_R + a_R = +0.17332 > 0 would select the
    #   HLL average instead (whose mass flux differs from m_R by ~0.03).
    a = (1.4 * 2.0) ** 0.5
    uL, uR = -2.0, -1.5
    ps = 2.0 - 0.5 * (uR - uL) * a
    rs = 1.0 - (uR - (uL + uR) / 2) / a
    asR = (1.4 * ps / rs) ** 0.5
    assert uL + a < 0 and (uL + uR) / 2 + asR < 0 < uR + a and ps > 0
    L, R = _cons(1.0, uL, 0.0, 2.0), _cons(1.0, uR, 0.0, 2.0)
    f = O.hll(G, L, R, True)
    FR = O.euler_flux(G, R, True)
    assert np.array_equal(f, FR)
    # the HLL average of the textbook bounds for comparison: SL from the same
    # formula, SR = u_R + a_R
    SL = min(uL - a, (uL + uR) / 2 - (1.4 * ps / (1.0 + (uL - (uL + uR) / 2) / a)) ** 0.5)
    SR = uR + a
    FL = O.euler_flux(G, L, True)
    h0 = (SR * FL[0] - SL * FR[0] + SL * SR * (R[0] - L[0])) / (SR - SL)
    assert abs(h0 - FR[0]) > 0.02
    # the same pair across a y-face (momentum components swapped)
    Ly, Ry = [L[0], L[2], L[1], L[3]], [R[0], R[2], R[1], R[3]]
    assert np.array_equal(O.hll(G, Ly, Ry, False), O.euler_flux(G, Ry, False))


def test_hll_x_y_mirror():
    # the y flux of a state equals the x flux of the state with its momentum
    # components swapped, components swapped back (bit for bit)
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = _cons(rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.2, 2))
        b = _cons(rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.2, 2))
        fy = O.hll(G, a, b, False)
        sa = [a[0], a[2], a[1], a[3]]
        sb = [b[0], b[2], b[1], b[3]]
        fx = O.hll(G, sa, sb, True)
        assert np.array_equal(fy, [fx[0], fx[2], fx[1], fx[3]])



def _hll_source_order(g, L, R, xdir):
    """EulerPhase::computePressurefromOther (EulerPhase.cpp:130-133),
    EulerMod::comFlux (EulerMod.cpp:33-54) and RiemmanSolverHLL::Solver
    (RiemmanSolverHLL.cpp:5-75) in Python floats, operations and order as
    written; the "velocity" of the source is the momentum (std::min / max as
    their ternaries)."""
    def pressure(s):
        rho, m0, m1, E = s
        return (g - 1) * (E - 0.5 * (m0 * m0 + m1 * m1) / rho)

    def flux(s, p):
        rho, u, v, E = s
        if xdir:
            return [u, u * u / rho + p, u * v / rho, (E + p) * (u / rho)]
        return [v, u * v / rho, v * v / rho + p, (E + p) * (v / rho)]

    def sqrt(x):   # std::sqrt: NaN below zero
        return math.sqrt(x) if x >= 0 else math.nan

    pL, pR = pressure(L), pressure(R)
    FL, FR = flux(L, pL), flux(R, pR)
    rhoL, uL, EL = L[0], (L[1], L[2]), L[3]
    rhoR, uR, ER = R[0], (R[1], R[2]), R[3]
    aL = sqrt(G * pL / rhoL)
    aR = sqrt(G * pR / rhoR)
    i = 0 if xdir else 1
    a_ave = 0.5 * (aL + aR)
    rho_ave = 0.5 * (rhoL + rhoR)
    p_star = 0.5 * (pL + pR) - 0.5 * (uR[i] / rhoR - uL[i] / rhoL) * rho_ave * a_ave
    u_star = 0.5 * (uR[i] / rhoR + uL[i] / rhoL) - 0.5 * (pR - pL) / (rho_ave * a_ave)
    rho_star_L = rhoL + (uL[i] / rhoL - u_star) * rho_ave / a_ave
    rho_star_R = rhoR - (uR[i] / rhoR - u_star) * rho_ave / a_ave
    a_star_L = sqrt(G * p_star / rho_star_L)
    a_star_R = sqrt(G * p_star / rho_star_R)
    x1, x2 = uL[i] - aL, u_star - a_star_L
    SL = x2 if x2 < x1 else x1
    y1, y2 = uL[i] + aL, u_star + a_star_R
    SR = y2 if y1 < y2 else y1
    if SL >= 0:
        return FL
    if SR <= 0:
        return FR
    inv = 1.0 / (SR - SL)
    return [(SR * FL[0] - SL * FR[0] + SL * SR * (rhoR - rhoL)) * inv,
            (SR * FL[1] - SL * FR[1] + SL * SR * (uR[0] - uL[0])) * inv,
            (SR * FL[2] - SL * FR[2] + SL * SR * (uR[1] - uL[1])) * inv,
            (SR * FL[3] - SL * FR[3] + SL * SR * (ER - EL)) * inv]


def test_hll_bitwise_equal_to_the_source_order():
    rng = np.random.default_rng(29)
    branches = set()
    for k in range(2000):
        st = []
        for _ in range(2):
            rho = float(rng.uniform(0.1, 2.0))
            u, v = (float(x) for x in rng.uniform(-3.0, 3.0, size=2))
            p = float(10.0 ** rng.uniform(-1, 1))
            st.append([rho, rho * u, rho * v, p / (G - 1) + 0.5 * rho * (u * u + v * v)])
        for xdir in (True, False):
            want = _hll_source_order(G, st[0], st[1], xdir)
            if not all(math.isfinite(x) for x in want):
                continue
            got = O.hll(G, st[0], st[1], xdir)
            assert [float(x) for x in got] == want, (st, xdir)
            fl = _hll_source_order(G, st[0], st[0], xdir)
            branches.add((want == fl, want == _hll_source_order(G, st[1], st[1], xdir)))
    assert len(branches) >= 2   # both upwind and HLL-average branches were taken


# ---- face walk ----------------------------------------------------------------
def test_face_stencils_n8():
    # cells 0..7, ghost -1 and 8; the walk stops at boundary faces 0 and 8
    table = {
        0: [-1, -1, -1, 0, 0, 0],
        1: [-1, -1, 0, 1, 2, 3],
        2: [-1, 0, 1, 2, 3, 4],
        3: [0, 1, 2, 3, 4, 5],
        5: [2, 3, 4, 5, 6, 7],
        6: [3, 4, 5, 6, 7, 8],
        7: [4, 5, 6, 7, 8, 8],
        8: [7, 7, 7, 8, 8, 8],
    }
    for f, want in table.items():
        assert O.collect_stencil(f, 8) == want, f
    assert O.collect_stencil(0, 8, periodic=True) == [5, 6, 7, 0, 1, 2]
    assert O.collect_stencil(8, 8, periodic=True) == [5, 6, 7, 0, 1, 2]


# ---- RK stage formulas --------------------------------------------------------
def test_rk2_constant_rhs():
    assert O.rk_apply(2, 1.0, [3.0, 3.0], 0.25) == 1.75


def test_rk3_as_written_constant_rhs_gives_four_thirds():
    # TVD3rdRK keeps the stage-1 snapshot in its last stage:
    # U^{n+1} = U0 + (4/3) dt L for constant L (a consistent RK3 gives U0 + dt L)
    r = O.rk_apply(3, 1.0, [3.0, 3.0, 3.0], 0.25)
    assert r == pytest.approx(1.0 + 4.0 / 3.0 * 0.25 * 3.0, rel=1e-15)


def test_rk3_as_written_ignores_second_stage_rhs():
    # U3 = U0/3 + 2/3 (U0 + dt L1) + 2/3 dt L3: the second stage's L2 only
    # reaches the result through the state it produces, never directly
    a = O.rk_apply(3, 1.0, [0.5, -7.0, 2.0], 0.1)
    b = O.rk_apply(3, 1.0, [0.5, 11.0, 2.0], 0.1)
    assert a == b
    assert a == pytest.approx(1.0 + 2.0 / 3.0 * 0.1 * 0.5 + 2.0 / 3.0 * 0.1 * 2.0, rel=1e-15)



def _rk_linear(rk, z, u0=1.0):
    """One step of the oracle's stage formulas on L(U) = lam U with dt = 1,
    lam = z: the right-hand side of each sub-stage is lam times the state
    the previous sub-stage produced (rk_apply replays the earlier ones)."""
    rhs, u = [], u0
    for k in range(rk):
        rhs.append(z * u)
        u = O.rk_apply(rk, u0, rhs, 1.0)
    return u


@pytest.mark.parametrize("z", [-0.9, -0.3, 0.05, 0.4])
def test_rk_amplification_factors_of_linear_problem(z):
    # TVD2rdRK: U1 = (1+z)U0, U = U0/2 + (U1 + z U1)/2  =>  R = 1 + z + z^2/2
    assert _rk_linear(2, z) == pytest.approx(1 + z + z * z / 2, rel=1e-14)
    # TVD3rdRK as written: U1 = (1+z)U0, U2 = 3/4 U0 + 1/4 (1+z)^2 U0,
    # U = U0/3 + 2/3 U1 + 2/3 z U2  =>  R = 1 + 4z/3 + z^2/3 + z^3/6
    # (the SSP scheme's R = 1 + z + z^2/2 + z^3/6)
    assert _rk_linear(3, z) == pytest.approx(1 + 4 * z / 3 + z * z / 3 + z ** 3 / 6, rel=1e-14)



def _rk_source_order(rk, u, rhs, dt):
    """TVD2rdRK::advance (TVD2rdRK.cpp:3-31) and TVD3rdRK::advance
    (TVD3rdRK.cpp:3-48) on one value over consecutive stage counters 1, 2, ...
    (stage = counter % rk mapped as the source does), coefficient tables of
    TVD2rdRK.h / TVD3rdRK.h, operations in the source's order."""
    if rk == 2:
        coef = [[0.0, 1.0, 1.0], [0.5, 0.5, 0.5]]
        phase = None
        for n, r in enumerate(rhs, start=1):
            stage = 1 if n % 2 == 0 else 0
            if stage == 0:
                phase = u
                u = coef[stage][1] * phase + coef[stage][2] * dt * r
            else:
                u = coef[stage][0] * phase + coef[stage][1] * u + coef[stage][2] * dt * r
        return u
    coef = [[0.0, 1.0, 1.0], [0.75, 0.25, 0.25], [1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0]]
    phase0 = phase1 = None
    for n, r in enumerate(rhs, start=1):
        stage = 2 if n % 3 == 0 else 0 if n % 3 == 1 else 1
        if stage == 0:
            phase0 = u
            u = coef[stage][1] * phase0 + coef[stage][2] * dt * r
        elif stage == 1:
            phase1 = u
            u = coef[stage][0] * phase0 + coef[stage][1] * phase1 + coef[stage][2] * dt * r
        else:
            u = coef[stage][0] * phase0 + coef[stage][1] * phase1 + coef[stage][2] * dt * r
    return u


def test_rk_stages_bitwise_equal_to_the_source_order():
    rng = np.random.default_rng(31)
    for _ in range(2000):
        u = float(rng.normal() * 10.0 ** rng.integers(-3, 4))
        dt = float(10.0 ** rng.uniform(-6, -1))
        rhs = (rng.normal(size=3) * 10.0 ** rng.integers(-2, 5)).tolist()
        for rk in (2, 3):
            for n in range(1, rk + 1):
                assert O.rk_apply(rk, u, rhs[:n], dt) == _rk_source_order(rk, u, rhs[:n], dt), (rk, n, u, rhs, dt)


# ---- checksum tree -----------------------------------------------------------
def test_tree_sum_order():
    # the halving tree adds a0 + a2 first: (1e16 + -1e16) + 1 = 1, while a
    # left-to-right sum would give 0
    assert O.tree_sum([1e16, 1.0, -1e16]) == 1.0
    rng = np.random.default_rng(3)
    a = rng.normal(size=100_003)
    assert abs(O.tree_sum(a) - math.fsum(a)) <= 1e-12 * np.abs(a).sum()
    assert O.tree_sum([]) == 0.0
