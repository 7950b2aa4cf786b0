"""Immersed-boundary set-up on the GPU (cfs_circle_setup_device) against the
host set-up (cfs_circle_setup, itself bit-identical to the oracle's
restatement of Circle.cpp:12-215): the same cell types, ghost list, image
points, donors and coefficients, bit for bit, including ghosts whose donor
candidates tie in distance (finished on the host by the same std::sort)."""
import time

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu

KEYS = ("cell_type", "ig_cells", "donors", "image_xy", "coef", "transp")


def _both(cfg, cx, cy, r, npts=400, layers=3, interp=cfs.IB_BILINEAR_POSITIVE):
    t0 = time.perf_counter()
    h = cfs.circle_setup(cfg, cx, cy, r, npts, layers, interp=interp)
    t1 = time.perf_counter()
    d = cfs.circle_setup(cfg, cx, cy, r, npts, layers, interp=interp, device=True)
    t2 = time.perf_counter()
    assert h["n_ig"] == d["n_ig"]
    for k in KEYS:
        a, b = np.asarray(h[k]), np.asarray(d[k])
        assert a.shape == b.shape, k
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
    print("%dx%d r=%g npts=%d layers=%d interp=%d: %d ghosts, %d finished on the host; host %.2f ms, device %.2f ms"
          % (cfg.nx, cfg.ny, r, npts, layers, interp, d["n_ig"], d["n_host"], 1e3 * (t1 - t0), 1e3 * (t2 - t1)))
    return h, d


@pytest.mark.parametrize("interp", [cfs.IB_REFERENCE, cfs.IB_BILINEAR_POSITIVE])
def test_headline_cylinder(interp):
    cfg = cfs.config_cylinder(2048, 1024, 4.0, 2.0)
    _, d = _both(cfg, 1.0, 1.0, 0.125, interp=interp)
    assert d["n_ig"] > 1000


def test_tied_candidates_are_finished_on_the_host():
    # an octagon centred on a cell corner: the image points of 8 ghosts have
    # equally distant donor candidates, whose order std::sort decides
    cfg = cfs.make_config(64, 64, 1.0, 1.0)
    _, d = _both(cfg, 0.5, 0.5, 0.25, npts=8, interp=cfs.IB_REFERENCE)
    assert d["n_host"] > 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_served_cylinders(seed):
    rng = np.random.default_rng(seed)
    ny = int(rng.choice([128, 192, 256, 384, 512]))
    cfg = cfs.config_cylinder(2 * ny, ny, 4.0, 2.0)
    _both(cfg, 1.0, 1.0, float(rng.uniform(0.05, 0.2)))


@pytest.mark.parametrize("layers", [1, 2])
def test_layers_origin_and_coarse_polygon(layers):
    cfg = cfs.make_config(200, 150, 2.0, 1.5, x0=-0.3, y0=0.2)
    _both(cfg, 0.8, 0.9, 0.31, npts=40, layers=layers, interp=cfs.IB_REFERENCE)


def test_same_error_for_a_body_on_the_boundary():
    # a body across the domain edge leaves ghosts with short donor windows or
    # none at all: both paths give the same answer, error or tables
    cfg = cfs.make_config(96, 96, 1.0, 1.0)
    out = []
    for device in (False, True):
        try:
            out.append(cfs.circle_setup(cfg, 0.02, 0.5, 0.2, device=device))
        except cfs.CfsError as e:
            out.append(e.status)
    if isinstance(out[0], dict):
        for k in KEYS:
            assert np.array_equal(np.asarray(out[0][k]).view(np.uint8), np.asarray(out[1][k]).view(np.uint8))
    else:
        assert out[0] == out[1]


def test_setup_device_tables_run_the_cylinder():
    # the device tables drive a solve exactly as the host tables do
    cfg = cfs.config_cylinder(256, 128, 4.0, 2.0, freestream=(1.4, 3.0, 0.0, 1.0), positivity=2, hll_bounds=0)
    ibd = cfs.circle_setup(cfg, 1.0, 1.0, 0.125, device=True)
    ibh = cfs.circle_setup(cfg, 1.0, 1.0, 0.125)
    outs = []
    for ib in (ibh, ibd):
        s = cfs.Solver(cfg)
        s.set_immersed(ib)
        s.set_state(cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=(1.4, 3.0, 0.0, 1.0)))
        s.advance(20, dt=0.0, cfl=0.5)
        outs.append(s.get_state())
        s.close()
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
