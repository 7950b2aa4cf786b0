"""The exact rewrites the FP64 WENO5 kernel uses in place of the reference's
operations, checked against exact rational arithmetic (Python's int / Fraction
divisions round correctly):

* x/3 and x/6 as one multiply and one fma with 1/d split into two doubles,
  for 2^-900 <= |x| <= 2^1000 (Weno5th.cpp:4-28 divides by 3.0 and 6.0);
* the smoothness indicator 13/12 s^2 + 1/4 t^2 with the quarter folded into
  an fma, which changes eps + beta (its only use) for no input.

The kernel's results are compared with the CPU restatement bit for bit by the
GPU parity tests; these pin the arithmetic claims on their own."""
import math
import random
from fractions import Fraction

import pytest

Z3, Z3L = float.fromhex("0x1.5555555555555p-2"), float.fromhex("0x1.5555555555555p-56")
Z6, Z6L = float.fromhex("0x1.5555555555555p-3"), float.fromhex("0x1.5555555555555p-57")
EPS = 1e-2
C1312 = 13.0 / 12.0


def fma(a, b, c):
    r = Fraction(a) * Fraction(b) + Fraction(c)
    if r == 0:   # IEEE zero sums: -0 only from two negative zeros
        neg_prod = math.copysign(1.0, a) * math.copysign(1.0, b) < 0
        return -0.0 if (a == 0 or b == 0) and c == 0 and neg_prod and math.copysign(1.0, c) < 0 else 0.0
    return float(r)


def div_split(x, z, zl):
    return fma(x, z, x * zl)


def samples(n, lo, hi, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        m = rng.getrandbits(52) | (1 << 52)
        e = rng.randint(lo, hi)
        out.append((-1) ** rng.getrandbits(1) * float(Fraction(m, 1 << 52) * Fraction(2) ** e))
    # mantissas that are multiples of 3 (exact quotients), all ones, and the
    # binade ends
    for e in (lo, -1, 0, 1, 52, hi - 1):
        for m in (1 << 52, (1 << 53) - 1, 3 << 50, 3 * ((1 << 51) - 1), 0x15555555555555, 0x1AAAAAAAAAAAAB):
            out.append(float(Fraction(m, 1 << 52) * Fraction(2) ** e))
    return out


def test_split_constants():
    assert Z3 == 1.0 / 3.0 and Z6 == 1.0 / 6.0
    assert Z3L == float(Fraction(1, 3) - Fraction(Z3))
    assert Z6L == float(Fraction(1, 6) - Fraction(Z6))


@pytest.mark.parametrize("d, z, zl", [(3, Z3, Z3L), (6, Z6, Z6L)])
def test_constant_division_correctly_rounded(d, z, zl):
    for x in samples(20000, -900, 999, seed=d):
        assert div_split(x, z, zl) == float(Fraction(x) / d), x.hex()
    for x in (0.0, -0.0):
        q = div_split(x, z, zl)
        assert q == 0.0 and str(q) == str(x)          # the sign of the zero


def test_constant_division_plain_product_is_not_enough():
    # a single product x * RN(1/3) misrounds some quotients: the split is needed
    xs = samples(2000, -10, 10, seed=7)
    assert any(x * Z3 != float(Fraction(x) / 3) for x in xs)


def beta_reference(s, t):
    return C1312 * (s * s) + 0.25 * (t * t)


def beta_fma(s, t):
    return fma(t * t, 0.25, C1312 * (s * s))


def test_beta_fma_equal_where_quarter_is_exact():
    rng = random.Random(11)
    for _ in range(20000):
        s = rng.uniform(-1, 1) * 2.0 ** rng.randint(-400, 400)
        t = rng.uniform(-1, 1) * 2.0 ** rng.randint(-400, 400)
        assert beta_fma(s, t) == beta_reference(s, t)


def test_beta_fma_same_eps_plus_beta_near_underflow():
    # t^2/4 subnormal: beta may differ in its last place, eps + beta may not
    rng = random.Random(12)
    differ = 0
    for _ in range(20000):
        t = rng.uniform(-1, 1) * 2.0 ** rng.randint(-545, -505)
        s = rng.uniform(-1, 1) * 2.0 ** rng.randint(-545, -480)
        b0, b1 = beta_reference(s, t), beta_fma(s, t)
        differ += b0 != b1
        assert EPS + b0 == EPS + b1
        if b0 != b1:
            assert max(b0, b1) < 2.0 ** -967
    assert differ > 0          # the case the argument covers does occur
