"""Checkpoint and resume: a solve stopped at a step boundary, its state taken
out with cfs_get_state and put into a fresh handle with cfs_set_state,
continues exactly as the uninterrupted solve does -- 100 CFL steps equal 50
steps, a save and a restore, then 50 more, bit for bit, with the same step
count and time.  The reference only writes snapshots (Run.cpp:131,207,235,
Output.cpp:20); the state is the whole of what a restart needs because the
CFL bound of a stored state is recomputed from it with the same function the
stages fold into their last sub-stage."""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu


def _cylinder(precision):
    cfg = cfs.config_cylinder(256, 128, 4.0, 2.0, precision=precision, positivity=2, hll_bounds=0, rk=3)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.2, interp=cfs.IB_BILINEAR_POSITIVE)
    return cfg, ib, cfs.cylinder_freestream(cfg, ib["cell_type"])


def _riemann(precision):
    cfg, U = cfs.riemann_batch_problems(5, count=1, sizes=(96,), precision=precision, rk=2)[0]
    return cfg, None, U


def _solver(cfg, ib, U):
    s = cfs.Solver(cfg)
    if ib is not None:
        s.set_immersed(ib)
    s.set_state(U)
    return s


@pytest.mark.parametrize("case", [_cylinder, _riemann])
@pytest.mark.parametrize("precision", [cfs.PREC_FP64, cfs.PREC_FP32_BULK])
def test_resume_equals_uninterrupted(case, precision):
    cfg, ib, U = case(precision)
    s = _solver(cfg, ib, U)
    t_all, n_all = s.advance(100, dt=0.0, cfl=0.5)
    ref = s.get_state()
    s.close()
    s = _solver(cfg, ib, U)
    t_half, n1 = s.advance(50, dt=0.0, cfl=0.5)
    saved = s.get_state()            # the checkpoint
    s.close()
    r = _solver(cfg, ib, saved)      # a fresh handle resumes from it
    t_end, n2 = r.advance(50, dt=0.0, cfl=0.5, t=t_half)
    got = r.get_state()
    r.close()
    print("%s %s: 100 steps to t=%.6g; 50 + 50 to t=%.6g" % (case.__name__, precision, t_all, t_end))
    assert (n1, n2, n_all) == (50, 50, 100)
    assert t_end == t_all
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_resume_with_fixed_steps_equals_the_oracle():
    # fixed dt, stage counter restarted at 1 by the fresh handle: the RK
    # sub-stage sequence begins again at a step boundary, as it must
    cfg = cfs.config_riemann(64, rk=3)
    U = cfs.riemann_quadrants(cfg)
    o = O.Oracle(cfg, nthreads=4)
    o.set_state(U)
    dt = 0.4 / o.cfl_smax()
    for _ in range(20):
        o.step(dt)
    s = _solver(cfg, None, U)
    for _ in range(10):
        s.step(dt)
    r = _solver(cfg, None, s.get_state())
    s.close()
    for _ in range(10):
        r.step(dt)
    got = r.get_state()
    r.close()
    assert np.array_equal(got.view(np.uint64), o.get_state().view(np.uint64))
