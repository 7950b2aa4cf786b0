"""Writes the fixtures of this folder.

weno5_exact.json -- Weno5th.cpp:4-28 (Jiang-Shu indicators, eps = 1e-2,
    ideal weights 0.1 / 0.6 / 0.3, the constants taken as the FP64 values the
    source writes) evaluated in exact rational arithmetic on fixed inputs:
    smooth, oscillatory, step, constant, tiny, huge and mixed-sign data.  The
    values are independent of the oracle and of the kernels (Python
    fractions only); a strict-IEEE FP64 evaluation lies within rounding of
    them.

laxliu3_32x32_rk3_dt1e-3_20steps.json -- a regression anchor, not an
    independent check: the SHA-256 of the state bytes (float64, [4][32][32],
    little-endian) after 20 TVD-RK3 steps of dt = 1e-3 on the Lax-Liu
    configuration 3 with transmissive boundaries, as the oracle computes it.
    The GPU path must reproduce it bit for bit, and so must any later oracle.

Usage: python tests/golden/make_golden.py   (from the repository root)
"""
import hashlib
import json
import os
import struct
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))                      # tests/ (the oracle binding)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))     # the repository root (the package)


def weno5_exact(v):
    F = Fraction
    v0, v1, v2, v3, v4 = (F(x) for x in v)
    c13, eps = F(13.0 / 12.0), F(1e-2)
    b0 = c13 * (v0 - 2 * v1 + v2) ** 2 + F(1, 4) * (v0 - 4 * v1 + 3 * v2) ** 2
    b1 = c13 * (v1 - 2 * v2 + v3) ** 2 + F(1, 4) * (v1 - v3) ** 2
    b2 = c13 * (v2 - 2 * v3 + v4) ** 2 + F(1, 4) * (3 * v2 - 4 * v3 + v4) ** 2
    a = [F(0.1) / (eps + b0) ** 2, F(0.6) / (eps + b1) ** 2, F(0.3) / (eps + b2) ** 2]
    q = [v0 / 3 - 7 * v1 / 6 + 11 * v2 / 6, -v1 / 6 + 5 * v2 / 6 + v3 / 3, v2 / 3 + 5 * v3 / 6 - v4 / 6]
    return sum(ai * qi for ai, qi in zip(a, q)) / sum(a)


INPUTS = [
    [1.0, 1.0, 1.0, 1.0, 1.0],
    [0.0, 0.0, 0.0, 0.0, 1.0],
    [0.0, 0.0, 0.0, 1.0, 1.0],
    [1.0, 1.0, 1.0, 0.0, 0.0],
    [0.1, 0.2, 0.3, 0.4, 0.5],
    [1.0, -1.0, 1.0, -1.0, 1.0],
    [0.125, 0.5, 1.125, 2.0, 3.125],
    [1.4, 1.4, 0.5323, 0.138, 0.138],
    [1.5, 0.5323, 0.138, 0.5323, 1.5],
    [2.5e-3, 2.6e-3, 2.4e-3, 2.9e-3, 2.2e-3],
    [1e70, 2e70, -3e70, 4e70, 5e70],
    [1e-150, 3e-150, 2e-150, -1e-150, 4e-150],
    [3.0, 2.9999999999, 3.0000000001, 2.9999999998, 3.0000000002],
    [-7.25, 11.5, -0.003, 1e3, -2.5e2],
    [0.9, 0.3, 0.7, 0.2, 0.8],
    [12345.678, 12345.679, 12345.677, 12345.68, 12345.676],
]


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def write_weno():
    rows = []
    for v in INPUTS:
        e = weno5_exact(v)
        rows.append({"v": [float.hex(x) for x in v], "exact_num": str(e.numerator), "exact_den": str(e.denominator),
                     "nearest": float.hex(float(e))})
    with open(os.path.join(HERE, "weno5_exact.json"), "w") as f:
        json.dump({"source": "Weno5th.cpp:4-28 in exact rationals (make_golden.py)", "vectors": rows}, f, indent=1)


LAXLIU = dict(n=32, rk=3, dt=1e-3, steps=20)


def laxliu_digest(U):
    import numpy as np
    return hashlib.sha256(np.ascontiguousarray(U, dtype="<f8").tobytes()).hexdigest()


def write_laxliu():
    import _oracle as O
    cfs = O.cfs
    cfg = cfs.config_riemann(LAXLIU["n"], rk=LAXLIU["rk"])
    o = O.Oracle(cfg, nthreads=1)
    o.set_state(cfs.riemann_quadrants(cfg))
    for _ in range(LAXLIU["steps"]):
        o.step(LAXLIU["dt"])
    U = o.get_state()
    o.close()
    with open(os.path.join(HERE, "laxliu3_32x32_rk3_dt1e-3_20steps.json"), "w") as f:
        json.dump(dict(LAXLIU, sha256=laxliu_digest(U),
                       rho_min=float.hex(float(U[0].min())), rho_max=float.hex(float(U[0].max()))), f, indent=1)


if __name__ == "__main__":
    write_weno()
    write_laxliu()
    print("wrote", sorted(p for p in os.listdir(HERE) if p.endswith(".json")))
