"""Batched and served solves at the sizes of the batched and serving
workloads, against the oracle and against solving alone.

Each batched or served problem must reproduce the reference's own stage loop
(Run.cpp:153-204, one problem per run, main.cpp:4-13) for that problem:

* the 256-problem batch (seeded four-quadrant Riemann problems, N drawn from
  {64, 96, ..., 512}, TVD-RK2, CFL 0.5, t_end 0.2, positivity limiter) in
  FP64 and FP32-bulk: every problem's final state, step count and time equal
  the single-solve path's bit for bit;
* six FP64 problems of that batch, two of them N = 512, after 10 batched CFL
  steps equal the oracle in its multi-rank form bit for bit;
* a seeded request stream at the serving workload's sizes (Riemann 128^2 to
  1024^2, cylinders 256x128 to 1024x512, Mach 1.5 to 4, t_end 0.1 to 0.5),
  served through both lanes (batched and solo), equals solving each request
  alone, in FP64 and FP32-bulk; and the FP64 stream, with every end time cut
  to 4.5 CFL time steps of its initial state, equals the oracle (which takes
  the same CFL steps, the last one clipped to the end time).
"""
import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu

BATCH_SEED = 2024      # the batched workload's seed
SERVE_SEED = 7         # the serving workload's seed
CFL = 0.5


def _alone(cfg, U, ib, cfl, t_end):
    s = cfs.Solver(cfg)
    if ib is not None:
        s.set_immersed(ib)
    s.set_state(U)
    try:
        t, done = s.advance(1 << 30, dt=0.0, cfl=cfl, t_end=t_end)
        ok = True
    except cfs.CfsError as e:
        assert e.status == cfs.CFS_E_NONFINITE
        t, done, ok = None, None, False
    out = s.get_state()
    s.close()
    return out, t, done, ok


def _same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def _oracle_cfl_run(cfg, U, ib, cfl, t_end=0.0, nsteps=None, nthreads=8):
    """The oracle in its multi-rank form taking CFL steps the way the library
    does: dt = cfl / s_max of the entering state, clipped so that the time
    does not pass t_end (t_end > 0), until t_end or for nsteps steps."""
    c64 = cfs.Config.from_buffer_copy(cfg)
    c64.precision = cfs.PREC_FP64
    o = O.Oracle(c64, nthreads=nthreads)
    o.set_ranks(min(nthreads, cfg.nx * cfg.ny // 4096 or 1))
    if ib is not None:
        o.set_ib_table(ib)
    o.set_state(U)
    t, steps = 0.0, 0
    while True:
        if nsteps is not None and steps == nsteps:
            break
        if t_end > 0.0 and not t < t_end:
            break
        d = cfl / o.cfl_smax()
        assert 0.0 < d < 1e300
        if t_end > 0.0 and not d < t_end - t:
            d = max(t_end - t, 0.0)
        o.step(d)
        t = t + d
        steps += 1
    out = o.get_state()
    o.close()
    return out, t, steps


def _batch_arena(problems):
    return int(sum(cfs.Batch.footprint(c) for c, _ in problems) * 1.05) + (64 << 20)


@pytest.mark.parametrize("precision", [cfs.PREC_FP32_BULK, cfs.PREC_FP64])
def test_batch_of_256_equals_alone(precision):
    probs = cfs.riemann_batch_problems(BATCH_SEED, 256, precision=precision, rk=2)
    assert len(probs) == 256 and {c.nx for c, _ in probs} == set(range(64, 513, 32))
    t_end = 0.2
    got = {}
    with cfs.Batch(rk=2, precision=precision, max_problems=256, arena_bytes=_batch_arena(probs),
                   limiter=True) as b:
        slots = {b.admit(cfg, U, cfl=CFL, t_end=t_end): k for k, (cfg, U) in enumerate(probs)}
        assert None not in slots and len(slots) == 256
        for _ in range(100000):
            b.step(50)
            for slot, (state, steps, t) in b.poll().items():
                assert state != cfs.SLOT_FAILED, "problem %d failed" % slots[slot]
                if state == cfs.SLOT_DONE:
                    got[slots[slot]] = (b.retire(slot), t, steps)
            if len(got) == 256:
                break
    assert len(got) == 256
    bad = []
    for k, (cfg, U) in enumerate(probs):
        ref, t, done, ok = _alone(cfg, U, None, CFL, t_end)
        Ub, tb, steps = got.pop(k)
        assert ok, "problem %d failed alone" % k
        if not (steps == done and tb == t and abs(t - t_end) <= 1e-12 and _same_bits(Ub, ref)):
            bad.append((k, cfg.nx, steps, done, tb, t))
    print("256 problems, N 64..512: batched == alone for %d" % (256 - len(bad)))
    assert not bad, bad


def test_batch_fp64_problems_equal_the_oracle():
    probs = cfs.riemann_batch_problems(BATCH_SEED, 256, precision=cfs.PREC_FP64, rk=2)
    sizes = [c.nx for c, _ in probs]
    n512 = [k for k, n in enumerate(sizes) if n == 512][:2]
    picks = sorted(set(n512 + [sizes.index(64), sizes.index(160), sizes.index(288), sizes.index(416)]))
    assert len(picks) == 6
    nsteps = 10
    with cfs.Batch(rk=2, precision=cfs.PREC_FP64, max_problems=256, arena_bytes=_batch_arena(probs),
                   limiter=True) as b:
        slots = [b.admit(cfg, U, cfl=CFL, t_end=0.0) for cfg, U in probs]   # no end time
        b.step(nsteps)
        info = b.poll()
        got = {}
        for k in picks:
            state, steps, t = info[slots[k]]
            assert state == cfs.SLOT_RUNNING and steps == nsteps
            got[k] = (b.retire(slots[k]), t)
    for k in picks:
        cfg, U = probs[k]
        ref, t, steps = _oracle_cfl_run(cfg, U, None, CFL, nsteps=nsteps)
        Ub, tb = got[k]
        print("problem %d (N=%d): %d CFL steps, t=%.6g, batched == oracle: %s"
              % (k, cfg.nx, steps, tb, _same_bits(Ub, ref)))
        assert tb == t
        assert _same_bits(Ub, ref), "problem %d (N=%d) differs from the oracle" % (k, cfg.nx)


def _stream(precision, count=12):
    """The serving workload's seeded stream (its first `count` requests):
    both lanes, Poisson arrivals at 50 requests/s."""
    reqs = cfs.serving_requests(SERVE_SEED, count, rate=50.0, precision=precision)
    tiles = [((r["cfg"].nx + 14) // 15) * ((r["cfg"].ny + 15) // 16) for r in reqs]
    # at least one request larger than any GPU holds workgroups of (solo lane)
    # and several that a batch takes; both bodies and grids of every kind
    assert max(tiles) >= 2048 and sum(t < 700 for t in tiles) >= 4
    assert any(r["ib"] is not None for r in reqs) and any(r["ib"] is None for r in reqs)
    return reqs


def _serve(reqs, precision):
    fps = [cfs.Batch.footprint(r["cfg"], r["ib"]["n_ig"] if r["ib"] is not None else 0) for r in reqs]
    # requests of fewer tiles than 2048 may be batched (the GPU holds at most
    # 2048 workgroups of the batched kernel at once); room for about three of
    # the largest of them: later arrivals queue
    small = [f for f, r in zip(fps, reqs) if (r["cfg"].nx + 14) // 15 * ((r["cfg"].ny + 15) // 16) < 2048]
    with cfs.Batch(rk=2, precision=precision, max_problems=16, arena_bytes=3 * max(small), limiter=True) as b:
        return b.serve(reqs, steps_per_poll=5)


@pytest.mark.parametrize("precision", [cfs.PREC_FP32_BULK, cfs.PREC_FP64])
def test_served_stream_equals_alone(precision):
    reqs = _stream(precision)
    res = _serve(reqs, precision)
    bad = []
    for k, (rq, r) in enumerate(zip(reqs, res)):
        ref, t, done, ok = _alone(rq["cfg"], rq["U"], rq["ib"], rq["cfl"], rq["t_end"])
        print("request %d %dx%d ib=%s: %s, %d steps, latency %.3f s"
              % (k, rq["cfg"].nx, rq["cfg"].ny, rq["ib"] is not None, r["state"], r["steps"], r["latency_s"]))
        assert r["admitted_s"] >= rq["arrival_s"] - 1e-9 and r["finished_s"] >= r["admitted_s"]
        assert ok and r["state"] == cfs.SLOT_DONE, "request %d failed" % k
        if not (r["steps"] == done and r["t"] == t and _same_bits(r["U"], ref)):
            bad.append(k)
    assert not bad, "served != alone for requests %s" % bad


def test_served_fp64_stream_equals_the_oracle():
    reqs = _stream(cfs.PREC_FP64)
    for rq in reqs:   # about five CFL steps each, the last clipped to the end time
        s = cfs.Solver(rq["cfg"])
        if rq["ib"] is not None:
            s.set_immersed(rq["ib"])
        s.set_state(rq["U"])
        rq["t_end"] = 4.5 * s.compute_dt(CFL)
        s.close()
    res = _serve(reqs, cfs.PREC_FP64)
    for k, (rq, r) in enumerate(zip(reqs, res)):
        ref, t, steps = _oracle_cfl_run(rq["cfg"], rq["U"], rq["ib"], rq["cfl"], t_end=rq["t_end"])
        print("request %d %dx%d ib=%s: %d steps (oracle %d), t=%.6g"
              % (k, rq["cfg"].nx, rq["cfg"].ny, rq["ib"] is not None, r["steps"], steps, r["t"]))
        assert r["state"] == cfs.SLOT_DONE and r["steps"] == steps and r["t"] == t
        assert abs(t - rq["t_end"]) <= 1e-12
        assert _same_bits(r["U"], ref), "request %d differs from the oracle" % k
