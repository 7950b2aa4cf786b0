"""Handles used from several threads at once on the GPU: worker threads create
solvers (allocations, an immersed boundary's tables), advance them (their
first CFL advance captures a graph), read them back and destroy them, while
the main thread steps a batch (its own graph capture, admission, retirement).
Every result equals the same solve run alone, bit for bit, and no call fails:
the library makes its captures, allocations and frees one at a time across
threads."""
import threading

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu


def _requests():
    reqs = cfs.serving_requests(23, 10, rate=1e9, precision=cfs.PREC_FP64, riemann_sizes=(48, 64, 80),
                                cylinder_ny=(32, 48), t_end_riemann=(0.01, 0.02), t_end_cylinder=(0.03, 0.05))
    return reqs


def _solve(rq):
    s = cfs.Solver(rq["cfg"])
    if rq["ib"] is not None:
        s.set_immersed(rq["ib"])
    s.set_state(rq["U"])
    t, done = s.advance(100000, dt=0.0, cfl=rq["cfl"], t_end=rq["t_end"])
    out = s.get_state()
    s.close()
    return out, t, done


def test_solvers_and_batch_from_several_threads():
    reqs = _requests()
    alone = [_solve(rq) for rq in reqs]
    results = {}
    errors = []

    def worker(w):
        try:
            for rnd in range(3):
                for k in range(w, len(reqs), 3):
                    results[(w, rnd, k)] = _solve(reqs[k])
        except Exception as e:   # reported below, with the worker that raised it
            errors.append((w, repr(e)))

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(3)]
    for th in threads:
        th.start()
    # the main thread meanwhile batches the same problems, admitting and retiring
    batched = {}
    for rnd in range(3):
        with cfs.Batch(rk=2, precision=cfs.PREC_FP64, max_problems=16, arena_bytes=64 << 20, limiter=True) as b:
            slots = {b.admit(rq["cfg"], rq["U"], cfl=rq["cfl"], t_end=rq["t_end"], ib=rq["ib"]): k
                     for k, rq in enumerate(reqs)}
            for _ in range(2000):
                b.step(2)
                for slot, (state, steps, t) in b.poll().items():
                    assert state != cfs.SLOT_FAILED
                    if state == cfs.SLOT_DONE:
                        batched[(rnd, slots[slot])] = (b.retire(slot), t, steps)
                if sum(1 for key in batched if key[0] == rnd) == len(reqs):
                    break
    for th in threads:
        th.join()
    assert not errors, errors
    assert len(results) == 3 * len(reqs) and len(batched) == 3 * len(reqs)
    for (w, rnd, k), (U, t, done) in results.items():
        assert done == alone[k][2] and t == alone[k][1]
        assert np.array_equal(U.view(np.uint64), alone[k][0].view(np.uint64)), (w, rnd, k)
    for (rnd, k), (U, t, steps) in batched.items():
        assert steps == alone[k][2] and t == alone[k][1]
        assert np.array_equal(U.view(np.uint64), alone[k][0].view(np.uint64)), (rnd, k)
