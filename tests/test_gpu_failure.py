"""Failure detection forced from the host: states the reference would carry on
with silently (it has no NaN checks) are reported through the C-ABI's status
codes, and the handle stays usable afterwards."""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu


def test_nan_state_checksums_nonfinite():
    cfg = cfs.config_riemann(40, rk=3)
    U = cfs.riemann_quadrants(cfg)
    s = cfs.Solver(cfg)
    s.set_state(U)
    good = s.checksums()
    assert np.all(np.isfinite(good))
    V = U.copy()
    V[2, 17, 23] = np.nan                      # one NaN in the y-momentum plane
    s.set_state(V)
    with pytest.raises(cfs.CfsError) as e:
        s.checksums()
    assert e.value.status == cfs.CFS_E_NONFINITE
    sums = s.checksums(check_finite=False)
    assert np.isnan(sums[2]) and np.all(np.isfinite(sums[[0, 1, 3]]))
    assert np.array_equal(sums[[0, 1, 3]], good[[0, 1, 3]])
    # the handle recovers with a good state
    s.set_state(U)
    assert np.array_equal(s.checksums(), good)


def test_negative_density_time_step_nonfinite():
    cfg = cfs.config_riemann(40, rk=3)
    U = cfs.riemann_quadrants(cfg)
    V = U.copy()
    V[0, 5, 5] = -V[0, 5, 5]                   # sound speed sqrt(gamma p / rho) of a negative density
    s = cfs.Solver(cfg)
    s.set_state(V)
    assert np.all(np.isfinite(s.checksums()))  # finite, so the sums cannot tell
    with pytest.raises(cfs.CfsError) as e:
        s.compute_dt(0.5)
    assert e.value.status == cfs.CFS_E_NONFINITE
    with pytest.raises(cfs.CfsError) as e:
        s.advance(10, dt=0.0, cfl=0.5)         # a CFL advance stops at the non-finite bound
    assert e.value.status == cfs.CFS_E_NONFINITE
    s.set_state(U)
    assert s.compute_dt(0.5) > 0.0
