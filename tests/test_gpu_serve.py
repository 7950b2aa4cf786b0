"""The serving loop on the GPU: requests arriving over time are admitted by
footprint into a small arena, batched per RK stage with whatever else is
resident, and each result equals the same solve run alone, bit for bit."""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu


def _alone(rq):
    s = cfs.Solver(rq["cfg"])
    if rq["ib"] is not None:
        s.set_immersed(rq["ib"])
    s.set_state(rq["U"])
    try:
        t, done = s.advance(100000, dt=0.0, cfl=rq["cfl"], t_end=rq["t_end"])
        ok = True
    except cfs.CfsError as e:
        assert e.status == cfs.CFS_E_NONFINITE
        t, done, ok = None, None, False
    out = s.get_state()
    s.close()
    return out, t, done, ok


@pytest.mark.parametrize("precision", [cfs.PREC_FP64, cfs.PREC_FP32_BULK])
def test_serve_results_equal_alone(precision):
    reqs = cfs.serving_requests(5, 12, rate=400.0, precision=precision, riemann_sizes=(40, 64, 75),
                                cylinder_ny=(32, 48), t_end_riemann=(0.01, 0.03), t_end_cylinder=(0.05, 0.1))
    for rq in reqs[:8]:
        rq["arrival_s"] = 0.0   # a burst: more than the arena holds
    fps = [cfs.Batch.footprint(r["cfg"], r["ib"]["n_ig"] if r["ib"] is not None else 0) for r in reqs]
    # room for about three problems at a time: later arrivals wait for space
    with cfs.Batch(rk=2, precision=precision, max_problems=8, arena_bytes=3 * max(fps), limiter=True) as b:
        res = b.serve(reqs, steps_per_poll=3)
    waited = 0
    for k, (rq, r) in enumerate(zip(reqs, res)):
        assert r["admitted_s"] >= rq["arrival_s"] - 1e-9
        assert r["finished_s"] >= r["admitted_s"]
        waited += r["admitted_s"] - rq["arrival_s"] > 1e-3
        ref, t, done, ok = _alone(rq)
        print("request %d %dx%d ib=%s: %s steps %d, latency %.4f s"
              % (k, rq["cfg"].nx, rq["cfg"].ny, rq["ib"] is not None, r["state"], r["steps"], r["latency_s"]))
        if not ok:
            assert r["state"] == cfs.SLOT_FAILED
            continue
        assert r["state"] == cfs.SLOT_DONE and r["steps"] == done and r["t"] == t
        assert np.array_equal(r["U"].view(np.uint64), ref.view(np.uint64)), "request %d differs" % k
    assert waited > 0   # the arena bound made some requests queue


def test_serve_rejects_a_request_larger_than_the_arena():
    cfg = cfs.make_config(64, 64, 1.0, 1.0, rk=3)
    U = cfs.riemann_quadrants(cfg)
    with cfs.Batch(rk=3, precision=cfs.PREC_FP64, max_problems=2, arena_bytes=cfs.Batch.footprint(cfg) // 2) as b:
        with pytest.raises(cfs.CfsError) as e:
            b.serve([dict(cfg=cfg, U=U, ib=None, cfl=0.5, t_end=0.01, arrival_s=0.0)])
        assert e.value.status == cfs.CFS_E_FULL


@pytest.mark.parametrize("precision", [cfs.PREC_FP64, cfs.PREC_FP32_BULK])
def test_serve_solo_lane_for_requests_that_fill_the_gpu(precision):
    # a 1024^2 request (4,416 tiles: more than the GPU holds workgroups at
    # once) goes to the solo lane: it is served although its footprint is
    # larger than the whole arena, and its result equals the solve alone, as
    # do those of the small requests batched meanwhile
    reqs = cfs.serving_requests(9, 6, rate=400.0, precision=precision, riemann_sizes=(40, 64),
                                cylinder_ny=(32, 48), t_end_riemann=(0.01, 0.02), t_end_cylinder=(0.05, 0.08))
    big_cfg, big_U = cfs.riemann_batch_problems(17, count=1, sizes=(1024,), precision=precision, rk=2)[0]
    reqs.insert(1, dict(cfg=big_cfg, U=big_U, ib=None, cfl=0.5, t_end=0.002, arrival_s=0.0))
    # three more of the same configuration (other states and end times): a
    # solo-lane worker keeps its solver handle for them
    for k, (c2, U2) in enumerate(cfs.riemann_batch_problems(18, count=3, sizes=(1024,), precision=precision,
                                                            rk=2)):
        reqs.append(dict(cfg=big_cfg, U=U2, ib=None, cfl=0.5, t_end=0.001 * (k + 1), arrival_s=0.0))
    fps = [cfs.Batch.footprint(r["cfg"], r["ib"]["n_ig"] if r["ib"] is not None else 0) for r in reqs]
    arena = 3 * max(fps[:1] + fps[2:-3])
    assert fps[1] > arena
    with cfs.Batch(rk=2, precision=precision, max_problems=8, arena_bytes=arena, limiter=True) as b:
        res = b.serve(reqs, steps_per_poll=3)
    for k, (rq, r) in enumerate(zip(reqs, res)):
        ref, t, done, ok = _alone(rq)
        print("request %d %dx%d: %s steps %d, latency %.4f s"
              % (k, rq["cfg"].nx, rq["cfg"].ny, r["state"], r["steps"], r["latency_s"]))
        if not ok:
            assert r["state"] == cfs.SLOT_FAILED
            continue
        assert r["state"] == cfs.SLOT_DONE and r["steps"] == done and r["t"] == t
        assert np.array_equal(r["U"].view(np.uint64), ref.view(np.uint64)), "request %d differs" % k
    assert res[1]["state"] == cfs.SLOT_DONE and res[1]["steps"] > 0
