"""Batched solves on the GPU: every problem's result in a batch is
bit-identical to solving it alone with the single-solve entry points (whose
parity with the oracle the other GPU tests establish), whatever else is in
the batch and whenever it was admitted."""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu

SUBSONIC = (1.0, 1.0, 0.0, 1.0)


def _problems(precision, rk, positivity=1):
    probs = []
    for k, (cfg, U) in enumerate(cfs.riemann_batch_problems(11, count=6, sizes=(37, 48, 64, 75, 96),
                                                            precision=precision, rk=rk, positivity=positivity)):
        probs.append((cfg, U, None, 0.5, 0.01 + 0.004 * k))
    fs = SUBSONIC if positivity < 2 else (1.4, 3.0, 0.0, 1.0)
    cyl = cfs.config_cylinder(96, 48, 4.0, 2.0, freestream=fs, positivity=positivity if positivity == 2 else 0,
                              hll_bounds=0, rk=rk, precision=precision)
    ib = cfs.circle_setup(cyl, 1.0, 1.0, 0.25, interp=cfs.IB_BILINEAR_POSITIVE)
    probs.append((cyl, cfs.cylinder_freestream(cyl, ib["cell_type"], freestream=fs), ib, 0.45, 0.15))
    return probs


def _alone(cfg, U, ib, cfl, t_end):
    s = cfs.Solver(cfg)
    if ib is not None:
        s.set_immersed(ib)
    s.set_state(U)
    t, done = s.advance(10000, dt=0.0, cfl=cfl, t_end=t_end)
    out = s.get_state()
    s.close()
    return out, t, done


def _run_batch(b, probs, first):
    """Admit `first` problems, step a few times, admit the rest, run all to
    their end times; returns {index: (U, t, steps)}."""
    slots = {}
    for k in range(first):
        cfg, U, ib, cfl, t_end = probs[k]
        slots[b.admit(cfg, U, cfl=cfl, t_end=t_end, ib=ib)] = k
    b.step(5)
    for k in range(first, len(probs)):
        cfg, U, ib, cfl, t_end = probs[k]
        slots[b.admit(cfg, U, cfl=cfl, t_end=t_end, ib=ib)] = k
    out = {}
    for _ in range(400):
        b.step(3)
        for slot, (state, steps, t) in b.poll().items():
            assert state != cfs.SLOT_FAILED, (slot, probs[slots[slot]][0].nx)
            if state == cfs.SLOT_DONE:
                out[slots[slot]] = (b.retire(slot), t, steps)
        if len(out) == len(probs):
            break
    assert len(out) == len(probs)
    return out


@pytest.mark.parametrize("precision,rk,positivity", [(cfs.PREC_FP64, 3, 1), (cfs.PREC_FP32_BULK, 3, 1),
                                                     (cfs.PREC_FP64, 2, 1), (cfs.PREC_FP64, 3, 2),
                                                     (cfs.PREC_FP32_BULK, 2, 2)])
def test_batch_equals_alone(precision, rk, positivity):
    probs = _problems(precision, rk, positivity)
    with cfs.Batch(rk=rk, precision=precision, max_problems=16, arena_bytes=256 << 20,
                   limiter=positivity == 2) as b:
        got = _run_batch(b, probs, first=3)
    for k, (cfg, U, ib, cfl, t_end) in enumerate(probs):
        ref, t, done = _alone(cfg, U, ib, cfl, t_end)
        Ub, tb, steps = got[k]
        print("problem %d (%dx%d): %d steps, t=%g" % (k, cfg.nx, cfg.ny, steps, tb))
        assert steps == done and tb == t and abs(t - t_end) <= 1e-15
        assert np.array_equal(Ub.view(np.uint64), ref.view(np.uint64)), "problem %d differs" % k


@pytest.mark.parametrize("workgroups", [8, 24, 40])
def test_batch_small_persistent_grid(workgroups):
    """A persistent grid of 8, 24 or 40 workgroups (one, three or five per XCD
    range: fewer than the parts a range is cut into, an odd count, and a
    count that leaves parts with different numbers of workgroups) still
    takes every tile of every problem once: the results equal solving alone."""
    probs = _problems(cfs.PREC_FP64, 2, 2)
    with cfs.Batch(rk=2, precision=cfs.PREC_FP64, max_problems=16, arena_bytes=256 << 20, limiter=True,
                   workgroups=workgroups) as b:
        got = _run_batch(b, probs, first=2)
    for k, (cfg, U, ib, cfl, t_end) in enumerate(probs):
        ref, t, done = _alone(cfg, U, ib, cfl, t_end)
        Ub, tb, steps = got[k]
        assert steps == done and tb == t
        assert np.array_equal(Ub.view(np.uint64), ref.view(np.uint64)), "problem %d differs" % k


def test_batch_arena_full_then_reuse():
    cfg = cfs.make_config(64, 64, 1.0, 1.0, rk=3)
    U = cfs.riemann_quadrants(cfg)
    fp = cfs.Batch.footprint(cfg)
    with cfs.Batch(rk=3, precision=cfs.PREC_FP64, max_problems=8, arena_bytes=3 * fp + fp // 2) as b:
        slots = [b.admit(cfg, U, cfl=0.5, t_end=0.05) for _ in range(3)]
        assert None not in slots
        assert b.admit(cfg, U, cfl=0.5, t_end=0.05) is None          # arena full
        b.step(4)
        b.retire(slots[1], copy=False)
        assert b.free_bytes() >= fp
        s = b.admit(cfg, U, cfl=0.5, t_end=0.05)                    # the freed space again
        assert s is not None
        for _ in range(100):
            b.step(2)
            if all(st == cfs.SLOT_DONE for st, _, _ in b.poll().values()):
                break
        ref, t, done = _alone(cfg, U, None, 0.5, 0.05)
        info = b.poll()
        for slot in (slots[0], slots[2], s):
            assert info[slot][0] == cfs.SLOT_DONE and info[slot][1] == done
            assert np.array_equal(b.retire(slot), ref)


def test_batch_slot_states():
    cfg = cfs.make_config(40, 40, 1.0, 1.0, rk=3)
    U = cfs.riemann_quadrants(cfg)
    bad = U.copy()
    bad[3, 5, 5] = -1.0                                                # negative pressure: non-finite bound
    with cfs.Batch(rk=3, precision=cfs.PREC_FP64, max_problems=2, arena_bytes=16 << 20) as b:
        a = b.admit(cfg, U, cfl=0.5, t_end=0.0)                          # no end time
        c = b.admit(cfg, bad, cfl=0.5, t_end=0.1)
        assert b.admit(cfg, U, cfl=0.5) is None                          # slot table full
        b.step(7)
        info = b.poll()
        assert info[a][0] == cfs.SLOT_RUNNING and info[a][1] == 7 and info[a][2] > 0
        assert info[c][0] == cfs.SLOT_FAILED and info[c][1] == 0
        with pytest.raises(cfs.CfsError):
            b.admit(cfs.make_config(40, 40, 1.0, 1.0, rk=2), U, cfl=0.5)   # rk differs from the batch's
