"""The C++ driver on the GPU: its snapshots (reference field layout, one .npy
per dataset) hold the same bits as the Python path through the same C-ABI,
and the FluxRho/FluxMu/FluxMv/FluxEnergy datasets equal the oracle's cell
right-hand sides of the last stage."""
import json
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu
RUN = os.path.join(os.path.dirname(cfs.LIB_PATH), "cfs_run")


NAMES = ("Mu", "Mv", "Rho", "Energy", "Pressure", "FluxMu", "FluxMv", "FluxRho", "FluxEnergy", "Udir", "Vdir",
         "Tdir")


def _load(d):
    """The snapshot's .npy datasets; where an HDF5 library is present, its
    snapshot.h5 must hold the same datasets (the reference's names, [nx][ny]
    native double) with the same bits."""
    out = {k: np.load(os.path.join(d, k + ".npy")) for k in NAMES}
    import _h5
    meta = json.load(open(os.path.join(d, "meta.json")))
    if _h5.lib() is not None:
        assert meta["hdf5"] == "snapshot.h5"
        h5 = _h5.read(os.path.join(d, "snapshot.h5"), NAMES)
        for k in NAMES:
            assert h5[k] is not None and np.array_equal(h5[k].view(np.uint64), out[k].view(np.uint64)), k
    return out


def test_driver_lax_liu_snapshots(tmp_path):
    cfg = cfs.config_riemann(40, rk=2)
    U = cfs.riemann_quadrants(cfg)
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(U)
    dt = 0.5 / o.cfl_smax()
    out = str(tmp_path / "run")
    log = str(tmp_path / "run.jsonl")
    r = subprocess.run([RUN, "--case", "lax-liu", "--nx", "40", "--ny", "40", "--rk", "2", "--stages", "101",
                        "--dt", repr(dt), "--every", "50", "--out", out, "--log", log], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    # run log: one record per output interval, the checksums of the state then
    records = [json.loads(line) for line in open(log)]
    assert [rec["stage"] for rec in records] == [50, 100]
    s = cfs.Solver(cfg)
    s.set_state(U)
    for i in range(1, 101):
        s.stage(i, dt)
        o.stage(i, dt)
        if i % 50:
            continue
        rec = records[i // 50 - 1]
        assert rec["step"] == i // 2 and rec["interval_stages"] == 50 and rec["finite"] is True
        assert rec["t"] == dt * i / 2 and rec["mcell_updates_per_s"] > 0 and rec["interval_s"] > 0
        assert np.array_equal(np.array(rec["checksums"]), s.checksums()), "run-log checksums at stage %d" % i
        snap = _load(os.path.join(out, "snapshot_%06d" % i))
        meta = json.load(open(os.path.join(out, "snapshot_%06d" % i, "meta.json")))
        assert meta["stage"] == i and meta["nx"] == 40 and meta["t"] == rec["t"]
        V, R = s.get_state(), s.last_rhs()
        for q, (name, flux) in enumerate((("Rho", "FluxRho"), ("Mu", "FluxMu"), ("Mv", "FluxMv"),
                                          ("Energy", "FluxEnergy"))):
            assert np.array_equal(snap[name], V[q]), name
            assert np.array_equal(snap[flux], R[q]), flux
        np.testing.assert_array_equal(R, o.get_rhs())
        np.testing.assert_array_equal(V, o.get_state())
        p = (cfg.gamma - 1) * (V[3] - 0.5 * (V[1] * V[1] + V[2] * V[2]) / V[0])
        assert np.array_equal(snap["Pressure"], p)
        assert not snap["Udir"].any() and not snap["Tdir"].any()


def test_driver_cylinder_default_geometry(tmp_path):
    # the reference's own set-up on a small mesh: circle r = 0.1 at (0.5, 1) on
    # the 2 x 2 square, inflow rho = u = p = 1, TVD-RK2, fixed dt
    n, stages, dt = 96, 31, 1e-3
    out = str(tmp_path / "cyl")
    r = subprocess.run([RUN, "--case", "cylinder", "--nx", str(n), "--ny", str(n), "--stages", str(stages),
                        "--dt", repr(dt), "--every", "30", "--out", out], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    cfg = cfs.config_cylinder(n, n, 2.0, 2.0, freestream=(1.0, 1.0, 0.0, 1.0), rk=2, positivity=0, hll_bounds=0)
    ib = cfs.circle_setup(cfg, 0.5, 1.0, 0.1, 400, 3, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=(1.0, 1.0, 0.0, 1.0))
    s = cfs.Solver(cfg)
    s.set_immersed(ib)
    s.set_state(U)
    for i in range(1, 31):
        s.stage(i, dt)
    snap = _load(os.path.join(out, "snapshot_%06d" % 30))
    V, R = s.get_state(), s.last_rhs()
    assert np.array_equal(snap["Rho"], V[0]) and np.array_equal(snap["Energy"], V[3])
    assert np.array_equal(snap["FluxMu"], R[1])
    solid = ib["cell_type"].reshape(n, n) == cfs.CELL_SOLID
    assert not snap["Pressure"][solid].any() and not R[:, solid].any()


def test_driver_cfl_run_log(tmp_path):
    # CFL steps to an end time in chunks of --every steps: one run-log record
    # per chunk, whose time, steps and checksums are those of the same chunked
    # advance through the Python path
    log = str(tmp_path / "cfl.jsonl")
    r = subprocess.run([RUN, "--case", "lax-liu", "--nx", "40", "--ny", "40", "--rk", "2", "--cfl", "0.5",
                        "--t-end", "0.05", "--every", "10", "--log", log], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    records = [json.loads(line) for line in open(log)]
    cfg = cfs.config_riemann(40, rk=2)
    s = cfs.Solver(cfg)
    s.set_state(cfs.riemann_quadrants(cfg))
    t, steps = 0.0, 0
    for rec in records:
        t, done = s.advance(10, dt=0.0, cfl=0.5, t_end=0.05, t=t)
        steps += done
        assert rec["step"] == steps and rec["stage"] == 2 * steps and rec["t"] == t and rec["finite"] is True
        assert np.array_equal(np.array(rec["checksums"]), s.checksums())
    assert len(records) >= 2 and abs(records[-1]["t"] - 0.05) <= 1e-15
    assert json.loads(r.stdout)["t"] == records[-1]["t"]
