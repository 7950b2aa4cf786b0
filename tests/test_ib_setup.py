"""Immersed-boundary set-up: the library's host code (cfs_circle_setup, runs
without a GPU) against the oracle's independent restatement of
source/object/Circle.cpp, and both against closed forms:

  - solid iff |centre - c| <= r                    (Circle.cpp:43-51)
  - immersed ghosts = non-fluid cells reachable from a fluid cell in <= 3
    steps along a grid line through non-fluid cells (Circle.cpp:53-103)
  - image point = mirror of the ghost centre through the nearest polygon
    segment line: |x_ip - x_g| = 2 d              (Circle.cpp:105-138)
  - donors = the 4 nearest fluid cells of the +-5 window (Circle.cpp:160-215)
  - reference interpolation solves sym(tril(M)) w = f (Eigen LDLT reads the
    lower triangle only, ImagePoint.cpp:22-51)
  - positivity-preserving weights are a convex combination that reproduces
    bilinear functions (bilinear branch) or equals inverse-distance weights.
"""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs

CASES = [
    (128, 64, 4.0, 2.0, 1.0, 1.0, 0.25),
    (200, 200, 2.0, 2.0, 0.5, 1.0, 0.1),
    (333, 170, 4.0, 2.0, 1.3, 0.9, 0.137),
    # bodies cut by the domain edge: the set-up scans only the body's
    # bounding box, which is then clipped to the grid
    (160, 80, 4.0, 2.0, 0.1, 1.0, 0.25),
    (96, 96, 1.0, 1.0, 0.95, 0.95, 0.2),
    (2048, 1024, 4.0, 2.0, 1.0, 1.0, 0.125),
]


@pytest.mark.parametrize("case", CASES)
def test_library_setup_equals_oracle(case):
    nx, ny, lx, ly, cx, cy, r = case
    cfg = cfs.config_cylinder(nx, ny, lx, ly)
    ib = cfs.circle_setup(cfg, cx, cy, r, interp=cfs.IB_REFERENCE)
    o = O.Oracle(cfg)
    o.set_circle(cx, cy, r)
    ob = o.ib_tables()
    for key in ("cell_type", "ig_cells", "donors", "image_xy"):
        assert np.array_equal(ib[key], ob[key]), key
    # reference factors equal the oracle's factorisation bit for bit
    coef = ib["coef"].reshape(-1, 10)
    tr = ib["transp"].reshape(-1, 4)
    for k in range(ib["n_ig"]):
        d = ib["donors"][4 * k:4 * k + 4]
        x = (d // ny + 0.5) * cfg.dx
        y = (d % ny + 0.5) * cfg.dy
        M = np.stack([x * y, x, y, np.ones(4)], 1)
        A, t = O.ldlt_factor(M)
        want = [A[1, 0], A[2, 0], A[2, 1], A[3, 0], A[3, 1], A[3, 2], A[0, 0], A[1, 1], A[2, 2], A[3, 3]]
        assert np.array_equal(coef[k], want)
        assert np.array_equal(tr[k], t)


@pytest.mark.parametrize("case", CASES)
def test_geometry_closed_forms(case):
    nx, ny, lx, ly, cx, cy, r = case
    cfg = cfs.config_cylinder(nx, ny, lx, ly)
    ib = cfs.circle_setup(cfg, cx, cy, r, interp=cfs.IB_REFERENCE)
    T = ib["cell_type"].reshape(nx, ny)
    x, y = cfg.centres()
    dist = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    assert np.array_equal(T != cfs.CELL_FLUID, dist <= r)
    # ghost layers by brute force
    nonfluid = T != cfs.CELL_FLUID
    ghost = np.zeros_like(nonfluid)
    for i, j in np.argwhere(~nonfluid):
        for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            for k in range(1, 4):
                a, b = i + k * di, j + k * dj
                if not (0 <= a < nx and 0 <= b < ny) or not nonfluid[a, b]:
                    break
                ghost[a, b] = True
    assert np.array_equal(T == cfs.CELL_IMMERSED_GHOST, ghost)
    # image points: mirror through the boundary, distance 2 d (polygon error ~ r(1-cos(pi/400)))
    cells = ib["ig_cells"]
    gx, gy = x.reshape(-1)[cells], y.reshape(-1)[cells]
    ip = ib["image_xy"].reshape(-1, 2)
    d_ip = np.hypot(ip[:, 0] - gx, ip[:, 1] - gy)
    d_wall = r - np.hypot(gx - cx, gy - cy)
    assert np.allclose(d_ip, 2 * d_wall, atol=2 * r * (1 - np.cos(np.pi / 400)) + 1e-12)
    # donors: the 4 nearest fluid cells of the window
    origin = (0.5 * cfg.dx, 0.5 * cfg.dy)
    for k in range(ib["n_ig"]):
        ic = int((ip[k, 0] - origin[0]) / cfg.dx)
        jc = int((ip[k, 1] - origin[1]) / cfg.dy)
        cand = []
        for a in range(ic - 5, ic + 6):
            for b in range(jc - 5, jc + 6):
                if 0 <= a < nx and 0 <= b < ny and T[a, b] == cfs.CELL_FLUID:
                    cand.append((x[a, b] - ip[k, 0]) ** 2 + (y[a, b] - ip[k, 1]) ** 2)
        cand = np.sort(cand)
        d = ib["donors"][4 * k:4 * k + 4]
        dd = (x.reshape(-1)[d] - ip[k, 0]) ** 2 + (y.reshape(-1)[d] - ip[k, 1]) ** 2
        assert np.all(T.reshape(-1)[d] == cfs.CELL_FLUID)
        assert np.allclose(np.sort(dd), cand[:4], rtol=0, atol=1e-15)


def test_reference_solve_is_lower_triangle_system():
    cfg = cfs.config_cylinder(128, 64, 4.0, 2.0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.25, interp=cfs.IB_REFERENCE)
    rng = np.random.default_rng(0)
    for k in range(ib["n_ig"]):
        d = ib["donors"][4 * k:4 * k + 4]
        x = (d // 64 + 0.5) * cfg.dx
        y = (d % 64 + 0.5) * cfg.dy
        M = np.stack([x * y, x, y, np.ones(4)], 1)
        S = np.tril(M) + np.tril(M, -1).T
        A, t = O.ldlt_factor(M)
        f = rng.normal(size=4)
        w = O.ldlt_solve(A, t, f)
        assert np.linalg.norm(S @ w - f) <= 1e-9 * np.linalg.cond(S) * np.linalg.norm(f) + 1e-12


def test_positive_weights():
    for nx, ny, lx, ly, cx, cy, r in CASES:
        cfg = cfs.config_cylinder(nx, ny, lx, ly)
        ib = cfs.circle_setup(cfg, cx, cy, r, interp=cfs.IB_BILINEAR_POSITIVE)
        ref = cfs.circle_setup(cfg, cx, cy, r, interp=cfs.IB_REFERENCE)
        assert np.array_equal(ib["donors"], ref["donors"])
        w = ib["coef"].reshape(-1, 10)
        assert np.all(w[:, 4:] == 0)
        g = w[:, :4]
        assert np.all(g >= 0)
        assert np.allclose(g.sum(1), 1.0, atol=1e-14)
        ip = ib["image_xy"].reshape(-1, 2)
        n_bilinear = 0
        for k in range(ib["n_ig"]):
            d = ib["donors"][4 * k:4 * k + 4]
            x = (d // ny + 0.5) * cfg.dx
            y = (d % ny + 0.5) * cfg.dy
            f = [np.ones(4), x, y, x * y]
            want = [1.0, ip[k, 0], ip[k, 1], ip[k, 0] * ip[k, 1]]
            if all(abs(g[k] @ fi - wi) <= 1e-11 * max(1, abs(wi)) for fi, wi in zip(f, want)):
                n_bilinear += 1
                continue
            d2 = (x - ip[k, 0]) ** 2 + (y - ip[k, 1]) ** 2
            idw = (1 / d2) / (1 / d2).sum()
            assert np.allclose(g[k], idw, rtol=1e-14, atol=0), k
        assert n_bilinear > 0
