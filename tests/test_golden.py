"""The committed fixtures of tests/golden (written by make_golden.py there):
the oracle's WENO5 against exact rational values of Weno5th.cpp:4-28, and
the oracle's Lax-Liu regression anchor.  The reference itself holds no test
vectors, so these are the fixed points the oracle is held to across changes."""
import hashlib
import json
import os
from fractions import Fraction

import numpy as np

import _oracle as O

cfs = O.cfs
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_weno5_fixtures_within_rounding_of_exact():
    vecs = _load("weno5_exact.json")["vectors"]
    assert len(vecs) >= 16
    for row in vecs:
        v = [float.fromhex(x) for x in row["v"]]
        exact = Fraction(int(row["exact_num"]), int(row["exact_den"]))
        assert float.fromhex(row["nearest"]) == float(exact)
        r = O.weno5(v)
        tol = Fraction(64 * float(np.spacing(max(abs(x) for x in v))))
        assert abs(Fraction(r) - exact) <= tol, (v, r, float(exact))


def laxliu_anchor():
    g = _load("laxliu3_32x32_rk3_dt1e-3_20steps.json")
    cfg = cfs.config_riemann(g["n"], rk=g["rk"])
    return g, cfg, cfs.riemann_quadrants(cfg)


def digest(U):
    return hashlib.sha256(np.ascontiguousarray(U, dtype="<f8").tobytes()).hexdigest()


def test_oracle_reproduces_the_laxliu_anchor():
    g, cfg, U0 = laxliu_anchor()
    o = O.Oracle(cfg, nthreads=2)
    o.set_state(U0)
    for _ in range(g["steps"]):
        o.step(g["dt"])
    U = o.get_state()
    o.close()
    assert digest(U) == g["sha256"]
    assert float(U[0].min()) == float.fromhex(g["rho_min"])
    assert float(U[0].max()) == float.fromhex(g["rho_max"])
