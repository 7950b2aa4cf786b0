"""Benchmark of the per-RK-stage hot path on one MI355X.

Workload (the headline configuration, one GPU): Mach-3 inviscid flow past a
circular cylinder with the ghost-cell immersed boundary, 2048 x 1024 cells on
[0,4] x [0,2], r = 0.125 at (1,1), 400 Lagrangian points, 3 ghost layers;
freestream rho = 1.4, u = 3, v = 0, p = 1 (c = 1); Dirichlet (supersonic)
inflow on the left, extrapolation (Neumann) elsewhere; gamma = 1.4, TVD-RK3,
CFL 0.5, 2000 timed RK3 steps (after 20 warm-up steps), FP64, then again in
FP32-bulk.  The reference's scheme goes non-finite within ~10 steps on this
flow (a near-vacuum wake behind the impulsively started body), so the
positivity-preserving flux limiter is on (positivity = 2: WENO face-state
fallback + Lax-Friedrichs blending per face); HLL keeps the reference's
bounds.  The image-point interpolation is the positivity-preserving bilinear
one (the reference's LDLT-on-lower-triangle interpolation gives negative
ghost densities).  One bench "step" = one RK3 step: 3 fused stage kernels
(IB fill in the tile load, WENO5/HLL/limiter/RK update, CFL bound in the last
stage, dt in the first), two steps per HIP graph replay.

Metric: Mcell-updates/s per RK stage = cells * stages / wall / 1e6, with the
state resident in HBM.  Beside it: the FP32-bulk form (FP64 state, FP32
reconstruction / flux / RHS) on the same workload, and the reference's scheme
unchanged (no limiter) on the same grid and body with the reference's own
subsonic inflow rho = u = p = 1 (Cylinder.cpp:5-16), 500 steps.

roofline: the stage kernel, algorithmic bytes per launch = cells * 97 B
(RK3 mean 96 B/cell-stage of FP64 state planes + 1 B cell-type mask) over its
mean launch duration measured with HIP events on the library's stream.
cpu_baseline: the strict-IEEE CPU restatement of the reference (oracle/) on
the same workload in the reference's MPI form (2D block decomposition, one
rank per usable host core -- the CPUs this process may run on, capped by the
control group's CPU quota -- 3-cell ghost frames exchanged every stage), a
bounded number of steps; the line names the host's cores and CPU model.

Beside the headline the line carries the rest of the metric:
  fp32_bulk    FP32-bulk throughput, and its L1(rho) against FP64 at the
               workload's own horizon (the time FP64 reaches in 2000 steps,
               t = 0.108) whatever --steps is, beside FP64 control runs from
               a density perturbed by 1e-7 and by 1e-12 and the 1e-4 target;
  kh           20 steps of the 4096^2 viscous shear layer in FP64 and
               FP32-bulk: state planes of 537 MB, beyond the Infinity Cache,
               so its stage kernel streams from HBM (roofline object);
  laxliu       the Lax-Liu 400^2 solve to t = 0.3 (RK3, CFL 0.5, FP64) on
               the GPU, with its rooflines (its bitwise comparison with the
               CPU restatement is --workload riemann's);
  serving      the seeded request stream served with continuous batching:
               a 48-request burst and Poisson arrivals at 2 and 4 requests/s
               (32 requests), p50/p99 latency and solves/s, against the burst
               solved one request after another in FP64.

The batch and shear-layer lines carry the same two objects: the batched
kernel's roofline over the batched leg's wall time (a lower bound on the
kernel's rate; its launches shrink as problems retire), the viscous kernel's
over its launch duration by HIP events; their CPU baselines run the
restatement on the same problems for 20 steps each within about 12 s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload batch | serve    (batched / serving measurements)
       python bench.py --workload kh               (viscous shear layer 4096^2 to t = 1, FP64 and FP32-bulk)
       python bench.py --workload riemann          (Lax-Liu 400^2 to t = 0.3, GPU vs CPU restatement, bitwise)
       python bench.py --workload ibsetup          (immersed-boundary set-up latency, host vs device)
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
cfs = importlib.import_module("2d_compressible_flow_solver_amd")

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_T = 39.3      # lane-ops/s / 1e12: 78.6 TFLOP/s FP64 vector = 39.3 T FMA lanes/s
NX, NY, LX, LY = 2048, 1024, 4.0, 2.0
CX, CY, RADIUS = 1.0, 1.0, 0.125
CFL = 0.5
FREESTREAM = (1.4, 3.0, 0.0, 1.0)           # Mach 3 at c = 1
REFERENCE_INFLOW = (1.0, 1.0, 0.0, 1.0)     # Cylinder.cpp:5-16
BYTES_PER_CELL_STAGE = 97.0
BYTES_PER_CELL_STAGE_RK3 = 96.0   # RK3 mean without the cell-type byte (no immersed body)


DUAL_FACE_SHAPE = 3   # shape 0's tile, each thread computing its x-face and its y-face in one pass


def kernel_name(prec, limiter, shape):
    ti, tj, nt = cfs.tile_shape_info(shape)
    return "stage_kernel<%s, false, %s, TileShape<%d, %d, %d%s>>" % (
        "double" if prec == cfs.PREC_FP64 else "float", "true" if limiter else "false", ti, tj, nt,
        ", true" if shape == DUAL_FACE_SHAPE else "")


def shape_report(r):
    ti, tj, nt = cfs.tile_shape_info(r["shape"])
    out = {"chosen": r["shape"], "tile": [ti, tj], "threads": nt}
    if r["shape_ms"] is not None:
        out["autotune_ms_per_launch"] = [round(x, 5) for x in r["shape_ms"]]
    return out


def profile_figure(key, shape=0):
    """A figure of the FP64 stage kernel of tile shape `shape` measured with
    rocprofv3 --pmc, from the committed summary: traffic_bytes_per_launch
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) or
    valu_lane_ops_per_cell_stage (SQ_INSTS_VALU x 64 / cells), keyed
    `<key>_shape<n>` for shapes other than 0; None when absent."""
    p = os.path.join(ROOT, "profiles", "stage_kernel_limiter_fp64_cylinder_mach3_2048x1024_counters.txt")
    if not os.path.exists(p):
        return None
    name = key if shape == 0 else "%s_shape%d" % (key, shape)
    for line in open(p):
        if line.startswith(name + ":"):
            return float(line.split(":", 1)[1])
    return None


def profile_row(name, label, nth=0, col=2):
    """Cell `col` (2: the second) of the nth table row whose first cell starts
    with `label` in the committed summary profiles/<name>, as a float (its
    first number); None when absent."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    rows = [l for l in open(p) if l.startswith("| " + label)]
    if len(rows) <= nth:
        return None
    cell = rows[nth].split("|")[col].strip()
    try:
        return float(cell.split()[0].replace(",", ""))
    except ValueError:
        return None


def host_cpu():
    """The host's processors as this process may use them: the CPUs it may run
    on (affinity), capped by the control group's CPU quota when one is set
    (cpu.max: quota and period in microseconds), plus what the machine reports
    (logical CPUs, lscpu model name).  `threads` is the worker count of every
    CPU baseline: one rank per usable core."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    threads = affinity if quota is None else max(1, min(affinity, int(quota)))
    return dict(threads=threads, logical_cpus=os.cpu_count(), affinity=affinity, cgroup_cpu_quota=quota,
                model=model)


HOST = None


def cpu_threads():
    global HOST
    if HOST is None:
        HOST = host_cpu()
    return HOST["threads"]


def cpu_window(problems, steps_each, budget_s=12.0):
    """The strict-IEEE CPU restatement (oracle/) in the reference's MPI form on
    a bounded sample: each (cfg, U, ib) of `problems` in turn for `steps_each`
    CFL steps (FP64, the problem's own RK order), until budget_s is spent
    (at least one step).  Returns (cell-stages, wall seconds, steps, threads,
    decompositions used)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    nthreads = cpu_threads()
    cell_stages, steps, dims_seen, wall = 0, 0, set(), 0.0
    for cfg, U, ib in problems:
        c64 = cfs.Config.from_buffer_copy(cfg)
        c64.precision = cfs.PREC_FP64
        o = O.Oracle(c64, nthreads=nthreads)
        dims_seen.add(tuple(o.set_ranks(min(nthreads, c64.nx * c64.ny // 64))))
        if ib is not None:
            o.set_ib_table(ib)
        o.set_state(U)
        t0 = time.perf_counter()
        for _ in range(steps_each):
            o.step(CFL / o.cfl_smax())
            steps += 1
            cell_stages += c64.nx * c64.ny * c64.rk
            if wall + time.perf_counter() - t0 > budget_s:
                break
        wall += time.perf_counter() - t0
        o.close()
        if wall > budget_s:
            break
    return cell_stages, wall, steps, nthreads, sorted(dims_seen)


def setup(precision, reference_scheme=False):
    fs = REFERENCE_INFLOW if reference_scheme else FREESTREAM
    cfg = cfs.config_cylinder(NX, NY, LX, LY, freestream=fs, precision=precision,
                              positivity=0 if reference_scheme else 2, hll_bounds=0)
    ib = cfs.circle_setup(cfg, CX, CY, RADIUS, 400, 3, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=fs)
    return cfg, ib, U


def run_gpu(precision, steps, warmup, reference_scheme=False, tile_shape=None):
    """tile_shape None: the autotuner picks it on the initial state."""
    cfg, ib, U = setup(precision, reference_scheme)
    s = cfs.Solver(cfg)
    s.set_immersed(ib)
    s.set_state(U)
    if tile_shape is None:
        shape, shape_ms = s.autotune(reps=20)
    else:
        s.set_tile_shape(tile_shape)
        shape, shape_ms = tile_shape, None
    tw = 0.0
    if warmup:
        tw, _ = s.advance(warmup, dt=0.0, cfl=CFL)
    t0 = time.perf_counter()
    t, done = s.advance(steps, dt=0.0, cfl=CFL, t=tw)
    wall = time.perf_counter() - t0
    # stage-kernel duration: a separate pass with HIP events around every
    # stage kernel (the events add queue latency, so not in the timed window)
    s.set_kernel_timing(True)
    t_end, _ = s.advance(min(steps, 40), dt=0.0, cfl=CFL, t=t)
    kms = s.last_kernel_ms()
    V = s.get_state()
    ok = bool(np.all(np.isfinite(V)))
    cells = NX * NY
    stages = 3 * done
    s.close()
    fluid = ib["cell_type"].reshape(NX, NY) == cfs.CELL_FLUID
    return dict(wall=wall, steps=done, mcups=cells * stages / wall / 1e6, kernel_ms=kms,
                ms_per_step=1e3 * wall / max(done, 1), finite=ok, t=t, t_end=t_end,
                total_steps=warmup + done + min(steps, 40), rho=V[0][fluid], shape=shape, shape_ms=shape_ms)


def density_at_time(precision, t_end, chunk=500, perturb=0.0):
    """Density on fluid cells of the headline workload advanced with CFL steps
    to exactly t_end (the last step clipped to it): accuracy comparisons at one
    physical time, free of the offset between runs that take the CFL steps of
    their own states.  perturb > 0 multiplies the initial density by
    1 + perturb * sin(k) over the cells k (the flow's own sensitivity)."""
    cfg, ib, U = setup(precision)
    if perturb:
        U = U.copy()
        U[0] = U[0] * (1.0 + perturb * np.sin(np.arange(U[0].size, dtype=np.float64)).reshape(U[0].shape))
    s = cfs.Solver(cfg)
    s.set_immersed(ib)
    s.set_state(U)
    t, steps = 0.0, 0
    while t < t_end:
        t, done = s.advance(chunk, dt=0.0, cfl=CFL, t_end=t_end, t=t)
        steps += done
        if done == 0:
            break
    V = s.get_state()
    s.close()
    fluid = ib["cell_type"].reshape(NX, NY) == cfs.CELL_FLUID
    return V[0][fluid], t, steps


CONTRACT_STEPS = 2000   # the headline workload's length: 2000 RK3 steps at CFL 0.5 (t = 0.108)
TARGET_L1 = 1e-4        # FP32-bulk target: L1(rho) against FP64


def density_after_steps(precision, nsteps):
    """Density on fluid cells of the headline workload after nsteps CFL steps
    from its initial state, and the time reached."""
    cfg, ib, U = setup(precision)
    s = cfs.Solver(cfg)
    s.set_immersed(ib)
    s.set_state(U)
    t, done = s.advance(nsteps, dt=0.0, cfl=CFL)
    V = s.get_state()
    s.close()
    fluid = ib["cell_type"].reshape(NX, NY) == cfs.CELL_FLUID
    return V[0][fluid], t, done


def run_accuracy():
    """FP32-bulk against FP64 on the headline flow at the workload's own
    horizon, whatever the timed window's length: FP64 takes CONTRACT_STEPS CFL
    steps from the initial state; FP32-bulk, and as controls FP64 from a
    density perturbed at the FP32 rounding level (1e-7) and far below it
    (1e-12), advance to exactly the time FP64 reached (their own CFL steps,
    the last one clipped).  The controls measure the flow's own sensitivity:
    what any perturbation of that size grows to by the horizon."""
    rho64, t64, n64 = density_after_steps(cfs.PREC_FP64, CONTRACT_STEPS)
    rho32, _, n32 = density_at_time(cfs.PREC_FP32_BULK, t64)
    rhoc, _, _ = density_at_time(cfs.PREC_FP64, t64, perturb=1e-7)
    rhoc12, _, _ = density_at_time(cfs.PREC_FP64, t64, perturb=1e-12)
    l1, l1c = float(np.abs(rho32 - rho64).mean()), float(np.abs(rhoc - rho64).mean())
    l1c12 = float(np.abs(rhoc12 - rho64).mean())
    print("fp32-bulk vs fp64 at t = %.5f: L1(rho) %.3e (controls: 1e-7 %.3e, 1e-12 %.3e)" % (t64, l1, l1c, l1c12),
          file=sys.stderr, flush=True)
    return {"compared_at_time": t64, "steps_fp64": n64, "steps_fp32_bulk": n32,
            "l1_density_vs_fp64": l1, "linf_density_vs_fp64": float(np.abs(rho32 - rho64).max()),
            "control_fp64_density_perturbed_1e-7": {"l1_density_vs_fp64": l1c,
                                                    "linf_density_vs_fp64": float(np.abs(rhoc - rho64).max())},
            "control_fp64_density_perturbed_1e-12": {"l1_density_vs_fp64": l1c12,
                                                     "linf_density_vs_fp64": float(np.abs(rhoc12 - rho64).max())},
            "target_l1": TARGET_L1, "meets_target": l1 <= TARGET_L1,
            "l1_over_control": round(l1 / l1c, 3) if l1c > 0 else None}


def run_kh_short(n=4096, steps=20):
    """A few CFL steps of the viscous shear layer at n x n (its planes far
    exceed the 256 MB Infinity Cache: the stage kernels stream from HBM), in
    FP64 and FP32-bulk, from its seeded initial state after 2 warm-up steps;
    stage-kernel durations by HIP events in a further 3 steps."""
    U = None
    out = {}
    for prec in (cfs.PREC_FP64, cfs.PREC_FP32_BULK):
        cfg = cfs.config_kelvin_helmholtz(n, precision=prec)
        if U is None:
            U = cfs.kelvin_helmholtz(cfg)
        s = cfs.Solver(cfg)
        s.set_state(U)
        t, _ = s.advance(2, dt=0.0, cfl=CFL)
        t0 = time.perf_counter()
        t, done = s.advance(steps, dt=0.0, cfl=CFL, t=t)
        wall = time.perf_counter() - t0
        s.set_kernel_timing(True)
        s.advance(3, dt=0.0, cfl=CFL, t=t)
        kms = s.last_kernel_ms()
        s.close()
        key = "fp64" if prec == cfs.PREC_FP64 else "fp32_bulk"
        out[key] = {"value": round(n * n * 3 * done / wall / 1e6, 2), "unit": "Mcell-updates/s", "steps": done,
                    "wall_s": round(wall, 4), "kernel_ms": round(kms, 5)}
        print("kh %s: %.1f Mcell-updates/s, stage kernel %.4f ms" % (key, out[key]["value"], kms),
              file=sys.stderr, flush=True)
    bytes_per_launch = n * n * BYTES_PER_CELL_STAGE_RK3
    kms = out["fp64"]["kernel_ms"]
    achieved = bytes_per_launch / (kms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profile_row(KH_PROFILE, "traffic (sum)") if n == 4096 else None,
            "kernel": "stage_kernel<double, true, false, TileShape<15, 16, 256>>", "kernel_ms": kms,
            "algorithmic_bytes_per_launch": bytes_per_launch}
    lane_ops = profile_row(KH_PROFILE, "SQ_INSTS_VALU", col=3) if n == 4096 else None
    if lane_ops:
        v = lane_ops * n * n / (kms * 1e-3) / 1e12
        roof["valu"] = {"achieved": round(v, 2), "peak": FP64_VALU_PEAK_T, "unit": "T lane-ops/s",
                        "frac": round(v / FP64_VALU_PEAK_T, 4), "lane_ops_per_cell_stage": lane_ops}
    k32 = out["fp32_bulk"]["kernel_ms"]
    return {"workload": "kelvin_helmholtz_%dx%d_viscous_re1e5_pr0.72_rk3_cfl0.5" % (n, n),
            "data": "synthetic (seeded shear layer, noise seed %d)" % cfs.KH_SEED,
            "fp64": out["fp64"], "fp32_bulk": out["fp32_bulk"],
            "fp32_bulk_kernel_speedup": round(kms / k32, 3) if k32 > 0 else None,
            "roofline": roof}


def run_cpu_baseline(budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    cfg, ib, U = setup(cfs.PREC_FP64)
    nthreads = cpu_threads()
    o = O.Oracle(cfg, nthreads=nthreads)
    dims = o.set_ranks(nthreads)   # the reference's MPI decomposition, one rank per host thread
    o.set_ib_table(ib)
    o.set_state(U)
    steps = 0
    t0 = time.perf_counter()
    while True:
        dt = CFL / o.cfl_smax()
        o.step(dt)
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    wall = time.perf_counter() - t0
    o.close()
    return dict(value=NX * NY * 3 * steps / wall / 1e6, unit="Mcell-updates/s", cores=nthreads,
                kind="port", host=HOST,
                sample="%d RK3 steps (%d stages incl. per-step CFL) of the same 2048x1024 cylinder "
                       "workload, %.1f s, strict-IEEE CPU restatement of the reference in its MPI form: "
                       "%dx%d blocks, one rank per thread, 3-cell ghost frames exchanged every stage"
                       % (steps, 3 * steps, wall, dims[0], dims[1]))


BATCH_SEED = 2024
BATCH_T_END = 0.2
BATCH_BYTES_PER_CELL_STAGE = 80    # TVD-RK2: stage 0 reads 32 B, stage 1 reads 64 B, each writes 32 B
BATCH_PROFILE = "batch_stage_kernel_riemann_batch_256.md"
KH_PROFILE = "stage_kernel_viscous_shear_layer_4096x4096.md"
RIEMANN_PROFILE = "stage_kernel_laxliu_400x400.md"


def run_batch(precision, count, poll_every=20, seed=BATCH_SEED, max_steps=20000, batch_only=False):
    """The batched workload: `count` seeded random Riemann problems of uneven
    size (N ~ U{64, 96, ..., 512}, transmissive boundaries, t_end = 0.2,
    CFL 0.5, TVD-RK2, positivity-preserving limiter) advanced together through the batch entry points until all
    reach t_end; then the same problems solved one after another with the
    single-solve entry points.  Inputs are uploaded at admission (outside the
    compute but inside the wall time of both legs)."""
    probs = cfs.riemann_batch_problems(seed, count=count, precision=precision, rk=2)
    total_cells = sum(c.nx * c.ny for c, _ in probs)
    arena = int(sum(cfs.Batch.footprint(c) for c, _ in probs) * 1.05) + (64 << 20)
    res = {}
    # tile shape: the autotuner on a sample of the workload (its first 64 problems)
    sample = probs[:64]
    shape, shape_ms = cfs.batch_autotune(sample, rk=2, precision=precision, limiter=True, steps=6,
                                         arena_bytes=int(sum(cfs.Batch.footprint(c) for c, _ in sample) * 1.05)
                                         + (64 << 20))
    print("batch autotune: shape %d, ms per step %s" % (shape, shape_ms), file=sys.stderr, flush=True)
    with cfs.Batch(rk=2, precision=precision, max_problems=count, arena_bytes=arena, limiter=True,
                   tile_shape=shape) as b:
        t0 = time.perf_counter()
        slot_of = {}
        for k, (cfg, U) in enumerate(probs):
            slot_of[b.admit(cfg, U, cfl=CFL, t_end=BATCH_T_END)] = k
        t_admit = time.perf_counter() - t0
        polls = 0
        while slot_of:
            b.step(poll_every)
            polls += 1
            for slot, (state, steps, t) in b.poll().items():
                if state != cfs.SLOT_RUNNING or steps >= max_steps:
                    res[slot_of.pop(slot)] = (state, steps, t)
                    b.retire(slot, copy=False)
            if polls % 25 == 0:
                print("batch: %d of %d problems left after %.1f s" % (len(slot_of), count, time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
    failed = sum(1 for st, _, _ in res.values() if st == cfs.SLOT_FAILED)
    # accuracy of the batched FP32-bulk results: the first 16 problems again in
    # FP64 through the single-solve path, L1(rho) at the same end time
    l1 = []
    if precision == cfs.PREC_FP32_BULK and not batch_only:
        with cfs.Batch(rk=2, precision=precision, max_problems=16, limiter=True, tile_shape=shape,
                       arena_bytes=int(sum(cfs.Batch.footprint(c) for c, _ in probs[:16]) * 1.05) + (64 << 20)) as b2:
            slots = {b2.admit(c, u, cfl=CFL, t_end=BATCH_T_END): k for k, (c, u) in enumerate(probs[:16])}
            got = {}
            while len(got) < len(slots):
                b2.step(poll_every)
                for slot, (state, steps, t) in b2.poll().items():
                    if state != cfs.SLOT_RUNNING and slots[slot] not in got:
                        got[slots[slot]] = b2.retire(slot)
        for k, (c, u) in enumerate(probs[:16]):
            c64 = cfs.Config.from_buffer_copy(c)
            c64.precision = cfs.PREC_FP64
            s = cfs.Solver(c64)
            s.set_state(u)
            s.advance(max_steps, dt=0.0, cfl=CFL, t_end=BATCH_T_END)
            l1.append(float(np.abs(got[k][0] - s.get_state()[0]).mean()))
            s.close()
    stalled = sum(1 for st, _, _ in res.values() if st == cfs.SLOT_RUNNING)
    # stages per step: the problems' RK order (TVD-RK2: 2)
    cell_stages = sum(probs[k][0].rk * res[k][1] * probs[k][0].nx * probs[k][0].ny for k in range(count))
    # the same problems one after another, each alone
    t0 = time.perf_counter()
    seq_steps = 0
    for k, (cfg, U) in enumerate([] if batch_only else probs):
        s = cfs.Solver(cfg)
        s.set_state(U)
        try:
            _, done = s.advance(max_steps, dt=0.0, cfl=CFL, t_end=BATCH_T_END)
        except cfs.CfsError:
            done = 0
        seq_steps += done
        s.close()
        if k % 32 == 31:
            print("one after another: %d of %d solved after %.1f s" % (k + 1, count, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    seq_wall = time.perf_counter() - t0
    return dict(problems=count, cells=total_cells, cell_stages=cell_stages, wall=wall, admit_s=t_admit,
                mcups=cell_stages / wall / 1e6, solves_per_s=count / wall, failed=failed, stalled=stalled,
                seq_wall=seq_wall, seq_mcups=cell_stages / seq_wall / 1e6, seq_solves_per_s=count / seq_wall,
                steps_min=min(r[1] for r in res.values()), steps_max=max(r[1] for r in res.values()),
                shape=shape, shape_ms=shape_ms, l1=l1)


def main_batch(args):
    prec = cfs.PREC_FP64 if args.batch_fp64 else cfs.PREC_FP32_BULK
    r = run_batch(prec, args.batch_count, batch_only=args.batch_only)
    print("batch: %d problems, %.2f s batched (%.1f Mcell-updates/s)" % (r["problems"], r["wall"], r["mcups"])
          + ("" if args.batch_only else ", %.2f s one after another (%.1f)" % (r["seq_wall"], r["seq_mcups"])),
          file=sys.stderr, flush=True)
    out = {
        "metric": "Mcell-updates/s per RK stage (batched solves)",
        "value": round(r["mcups"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64" if prec == cfs.PREC_FP64 else "f32 (FP64 state)",
        "data": "synthetic (seeded random four-quadrant Riemann states, seed %d)" % BATCH_SEED,
        "config": {"workload": "riemann_batch_%d_uneven_N64-512_rk2_cfl0.5_t0.2" % r["problems"],
                   "precision": "fp64" if prec == cfs.PREC_FP64 else "fp32_bulk",
                   "total_cells": r["cells"], "positivity_limiter": True, "hll_bounds": "velocity",
                   "tile_shape": shape_report(r)},
        "batched": {"wall_s": round(r["wall"], 3), "solves_per_s": round(r["solves_per_s"], 2),
                    "admission_s": round(r["admit_s"], 3), "failed": r["failed"], "stalled": r["stalled"],
                    "steps_range": [r["steps_min"], r["steps_max"]]},
        "one_after_another": None if args.batch_only else {
            "wall_s": round(r["seq_wall"], 3), "value": round(r["seq_mcups"], 2),
            "solves_per_s": round(r["seq_solves_per_s"], 2)},
        "speedup_vs_one_after_another": None if args.batch_only else round(r["seq_wall"] / r["wall"], 3),
    }
    if r["l1"]:
        out["l1_density_vs_fp64"] = {"problems": len(r["l1"]), "median": float(np.median(r["l1"])),
                                     "max": float(np.max(r["l1"]))}
    # roofline of the batched stage kernel over the batched leg's wall time
    # (admission, polls and retirements included: a lower bound on the
    # kernel's own rate; its launches shrink as problems retire)
    alg = BATCH_BYTES_PER_CELL_STAGE * r["cell_stages"]
    achieved = alg / r["wall"] / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "timed_over": "batched leg wall time",
                       "kernel": "batch_stage_kernel<%s, true, TileShape<%d, %d, %d>>" % (
                           ("double" if prec == cfs.PREC_FP64 else "float",) + tuple(cfs.tile_shape_info(r["shape"]))),
                       "algorithmic_bytes_per_cell_stage": BATCH_BYTES_PER_CELL_STAGE}
    if prec == cfs.PREC_FP32_BULK and r["problems"] == 256:
        ratio = profile_row(BATCH_PROFILE, "traffic / algorithmic")
        lane_ops = profile_row(BATCH_PROFILE, "VALU lane-ops per cell-stage")
        if ratio:
            out["roofline"]["traffic_over_algorithmic"] = ratio
        if lane_ops:
            # FP32 without packed math issues at the FP64 lane rate
            v = lane_ops * r["cell_stages"] / r["wall"] / 1e12
            out["roofline"]["valu"] = {"achieved": round(v, 2), "peak": FP64_VALU_PEAK_T, "unit": "T lane-ops/s",
                                       "frac": round(v / FP64_VALU_PEAK_T, 4), "lane_ops_per_cell_stage": lane_ops}
    out["cpu_baseline"] = None
    if not args.no_cpu and not args.batch_only:
        probs = cfs.riemann_batch_problems(BATCH_SEED, count=args.batch_count, precision=cfs.PREC_FP64, rk=2)
        cs, wall, steps, nthreads, dims = cpu_window([(c, u, None) for c, u in probs], 20)
        out["cpu_baseline"] = {
            "value": round(cs / wall / 1e6, 3), "unit": "Mcell-updates/s", "cores": nthreads, "kind": "port", "host": HOST,
            "sample": "%d RK2 steps of the batch's problems in order, 20 each (%.1f s), FP64, strict-IEEE CPU "
                      "restatement of the reference in its MPI form, one rank per thread, blocks %s"
                      % (steps, wall, ", ".join("%dx%d" % d for d in dims))}
        print("cpu baseline: %.2f Mcell-updates/s on %d threads" % (out["cpu_baseline"]["value"], nthreads),
              file=sys.stderr, flush=True)
    print(json.dumps(out))


def run_laxliu():
    """The Lax-Liu configuration 3 solve on 400 x 400 (RK3, CFL 0.5, FP64, to
    t = 0.3) on the GPU alone, timed after a warm-up that captures the CFL
    graph it replays; Mcell-updates/s and the rooflines over its wall time per
    stage (the bitwise comparison with the CPU restatement is
    --workload riemann's)."""
    cfg = cfs.config_riemann(400, rk=3)
    U = cfs.riemann_quadrants(cfg)
    s = cfs.Solver(cfg)
    s.set_state(U)
    s.advance(32, dt=0.0, cfl=CFL, t_end=0.3)
    s.set_state(U)
    t0 = time.perf_counter()
    t, steps = s.advance(100000, dt=0.0, cfl=CFL, t_end=0.3)
    wall = time.perf_counter() - t0
    shape = s.tile_shape
    s.close()
    cells = cfg.nx * cfg.ny
    out = {"workload": "riemann_laxliu3_400x400_rk3_cfl0.5_t0.3_fp64", "value": round(cells * 3 * steps / wall / 1e6, 2),
           "unit": "Mcell-updates/s", "steps": steps, "t": t, "wall_s": round(wall, 4), "tile_shape": shape}
    out["roofline"] = laxliu_roofline(cells, wall / (3 * steps))
    return out


def laxliu_roofline(cells, stage_s):
    """Rooflines of the Lax-Liu solve over its wall time per stage (launch
    gaps, halt polls and launches replayed after t = 0.3 included):
    algorithmic bytes 96 B per cell-stage (RK3 mean), VALU lane-ops per
    cell-stage from the committed counters of its kernel."""
    gbs = BYTES_PER_CELL_STAGE_RK3 * cells / stage_s / 1e9
    roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "timed_over": "solve wall time",
            "kernel": kernel_name(cfs.PREC_FP64, False, DUAL_FACE_SHAPE)}
    lane_ops = profile_row(RIEMANN_PROFILE, "SQ_INSTS_VALU", col=3)
    if lane_ops is not None:
        v = lane_ops * cells / stage_s / 1e12
        roof["valu"] = {"achieved": round(v, 2), "peak": FP64_VALU_PEAK_T, "unit": "T lane-ops/s",
                        "frac": round(v / FP64_VALU_PEAK_T, 4), "lane_ops_per_cell_stage": lane_ops}
    return roof


def main_riemann(args):
    """The reference's own CPU-runnable case end to end: the Lax-Liu
    four-quadrant problem on 400 x 400, WENO5 + HLL + TVD-RK3, CFL 0.5, FP64,
    transmissive boundaries, to t = 0.3; on the GPU (CFL steps on the device),
    then the CPU restatement on all host threads with the same time steps
    (dt = 0.5 / s_max of its own state, bit-identical to the device's), and the
    final states compared bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    cfg = cfs.config_riemann(400, rk=3)
    U = cfs.riemann_quadrants(cfg)
    s = cfs.Solver(cfg)
    s.set_state(U)
    # warm-up (not the measured solve): module load, first launches, and the
    # capture of the 8-step CFL graph the measured advance replays (same CFL,
    # end time and parities: an even number of steps of at least 32)
    s.advance(32, dt=0.0, cfl=CFL, t_end=0.3)
    s.set_state(U)
    t0 = time.perf_counter()
    t, steps = s.advance(100000, dt=0.0, cfl=CFL, t_end=0.3)
    gpu_wall = time.perf_counter() - t0
    G = s.get_state()
    s.close()
    # FP32-bulk on the same problem to the same end time.  The reference's
    # TVD3rdRK as written is not stable at CFL 0.5 here: rounding noise grows
    # until the state is non-finite (an FP64 run from a density perturbed by
    # 1e-12 shows it), so FP32-bulk is compared with FP64 on the reference
    # driver's TVD2rdRK (Run.cpp:47), stable at CFL 0.5.
    def solve(precision, rk, perturb=0.0):
        c = cfs.config_riemann(400, rk=rk, precision=precision)
        s = cfs.Solver(c)
        V0 = U
        if perturb:
            V0 = U.copy()
            V0[0] = V0[0] * (1.0 + perturb * np.sin(np.arange(V0[0].size, dtype=np.float64)).reshape(V0[0].shape))
        s.set_state(V0)
        t0 = time.perf_counter()
        t, n, status = 0.0, 0, "ok"
        while t < 0.3:
            try:
                t, done = s.advance(10, dt=0.0, cfl=CFL, t_end=0.3, t=t)
            except cfs.CfsError:
                status = "non-finite within steps %d-%d (t = %.4f at step %d)" % (n + 1, n + 10, t, n)
                break
            n += done
            if done == 0:
                break
        wall = time.perf_counter() - t0
        V = s.get_state()
        s.close()
        return dict(steps=n, t=round(t, 6), wall_s=round(wall, 4), status=status), V
    rk3_32, _ = solve(cfs.PREC_FP32_BULK, 3)
    rk3_pert, _ = solve(cfs.PREC_FP64, 3, perturb=1e-12)
    r2_64, H64 = solve(cfs.PREC_FP64, 2)
    r2_32, H32 = solve(cfs.PREC_FP32_BULK, 2)
    rk2 = {"fp64": r2_64, "fp32_bulk": r2_32}
    if r2_64["status"] == "ok" and r2_32["status"] == "ok":
        rk2["l1_density_fp32_bulk_vs_fp64"] = float(np.abs(H32[0] - H64[0]).mean())
        rk2["linf_density_fp32_bulk_vs_fp64"] = float(np.abs(H32[0] - H64[0]).max())

    nthreads = cpu_threads()
    o = O.Oracle(cfg, nthreads=nthreads)
    o.set_ranks(nthreads)
    o.set_state(U)
    tc, n = 0.0, 0
    t0 = time.perf_counter()
    while tc < 0.3:
        dt = CFL / o.cfl_smax()
        rem = 0.3 - tc                # the device's end-time clip
        if not (dt < rem):
            dt = rem
        o.step(dt)
        tc += dt
        n += 1
    cpu_wall = time.perf_counter() - t0
    C = o.get_state()
    o.close()
    cells = cfg.nx * cfg.ny
    out = {
        "metric": "Mcell-updates/s per RK stage (Lax-Liu 400x400 to t = 0.3)",
        "value": round(cells * 3 * steps / gpu_wall / 1e6, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "Lax-Liu configuration 3 quadrant states (no randomness)",
        "config": {"workload": "riemann_laxliu3_400x400_rk3_cfl0.5_t0.3_fp64"},
        "gpu": {"steps": steps, "t": t, "wall_s": round(gpu_wall, 4)},
        "rk3_fp32_bulk": rk3_32,
        "rk3_fp64_density_perturbed_1e-12": rk3_pert,
        "rk2_fp32_bulk_vs_fp64": rk2,
        "cpu_baseline": {"value": round(cells * 3 * n / cpu_wall / 1e6, 3), "unit": "Mcell-updates/s",
                         "cores": nthreads, "kind": "port", "steps": n, "wall_s": round(cpu_wall, 3),
                         "sample": "the whole solve to t = 0.3"},
        "bit_identical_at_t_end": bool(np.array_equal(G.view(np.uint64), C.view(np.uint64)) and n == steps),
        "max_abs_diff": float(np.abs(G - C).max()),
    }
    out["roofline"] = laxliu_roofline(cells, gpu_wall / (3 * steps))
    print(json.dumps(out))


def run_kh(precision, n, t_end, chunk=1000, checkpoints=(), perturb=0.0):
    """The viscous shear layer from its seeded initial state to t_end with CFL
    steps (progress every chunk steps), stopping exactly at each checkpoint
    time to record the density.  perturb > 0 multiplies the initial density
    by (1 + perturb) (a control for the flow's own sensitivity)."""
    cfg = cfs.config_kelvin_helmholtz(n, precision=precision)
    s = cfs.Solver(cfg)
    U = cfs.kelvin_helmholtz(cfg)
    if perturb:
        U = U.copy()
        U[0] *= 1.0 + perturb
    s.set_state(U)
    t, steps, status = 0.0, 0, "ok"
    snaps = {}
    t0 = time.perf_counter()
    for stop in sorted(set(list(checkpoints) + [t_end])):
        while t < stop:
            try:
                t, done = s.advance(chunk, dt=0.0, cfl=CFL, t_end=stop, t=t)
            except cfs.CfsError as e:
                status = str(e)
                break
            steps += done
            print("kh %s%s: %d steps, t = %.4f, %.1f s" % (
                "fp64" if precision == cfs.PREC_FP64 else "fp32-bulk", " (perturbed)" if perturb else "", steps,
                t, time.perf_counter() - t0), file=sys.stderr, flush=True)
            if done == 0:
                break
        if status != "ok":
            break
        snaps[stop] = s.get_state()[0]
    wall = time.perf_counter() - t0
    rho = snaps.get(t_end, s.get_state()[0])
    s.set_kernel_timing(True)
    s.advance(5, dt=0.0, cfl=CFL, t_end=0.0, t=t)
    kms = s.last_kernel_ms()
    s.close()
    return dict(steps=steps, t=t, wall=wall, mcups=n * n * 3 * steps / wall / 1e6, kernel_ms=kms, rho=rho,
                snaps=snaps, status=status, finite=bool(np.all(np.isfinite(rho))))


def main_kh(args):
    n, t_end = args.kh_n, args.kh_t_end
    cps = (0.1, 0.25, 0.5) if args.kh_control else ()
    r64 = run_kh(cfs.PREC_FP64, n, t_end, checkpoints=cps)
    r32 = run_kh(cfs.PREC_FP32_BULK, n, t_end, checkpoints=cps)
    l1 = float(np.abs(r32["rho"] - r64["rho"]).mean())
    strip = lambda r: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in r.items()
                       if k not in ("rho", "snaps")}
    out = {
        "metric": "Mcell-updates/s per RK stage (viscous shear layer)",
        "value": round(r64["mcups"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded shear layer, noise seed %d)" % cfs.KH_SEED,
        "config": {"workload": "kelvin_helmholtz_%dx%d_viscous_re1e5_pr0.72_rk3_cfl0.5_t%g" % (n, n, t_end),
                   "periodic": True},
        "fp64": strip(r64),
        "fp32_bulk": strip(r32),
        "fp32_bulk_speedup": round(r32["mcups"] / r64["mcups"], 3),
        "l1_density_fp32_bulk_vs_fp64": l1,
        "linf_density_fp32_bulk_vs_fp64": float(np.abs(r32["rho"] - r64["rho"]).max()),
    }
    # roofline of the FP64 viscous stage kernel (launch duration by HIP events
    # at t_end; RK3 mean 96 B per cell-stage)
    bytes_per_launch = n * n * BYTES_PER_CELL_STAGE_RK3
    if r64["kernel_ms"] > 0:
        achieved = bytes_per_launch / (r64["kernel_ms"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4),
                           "traffic": profile_row(KH_PROFILE, "traffic (sum)") if n == 4096 else None,
                           "kernel": "stage_kernel<double, true, false, ...>",
                           "kernel_ms": round(r64["kernel_ms"], 5), "algorithmic_bytes_per_launch": bytes_per_launch}
        lane_ops = profile_row(KH_PROFILE, "SQ_INSTS_VALU", col=3) if n == 4096 else None
        if lane_ops:
            v = lane_ops * n * n / (r64["kernel_ms"] * 1e-3) / 1e12
            out["roofline"]["valu"] = {"achieved": round(v, 2), "peak": FP64_VALU_PEAK_T, "unit": "T lane-ops/s",
                                       "frac": round(v / FP64_VALU_PEAK_T, 4), "lane_ops_per_cell_stage": lane_ops}
    out["cpu_baseline"] = None
    if not args.no_cpu:
        cfg = cfs.config_kelvin_helmholtz(n, precision=cfs.PREC_FP64)
        cs, wall, steps, nthreads, dims = cpu_window([(cfg, cfs.kelvin_helmholtz(cfg), None)], 20)
        out["cpu_baseline"] = {
            "value": round(cs / wall / 1e6, 3), "unit": "Mcell-updates/s", "cores": nthreads, "kind": "port", "host": HOST,
            "sample": "%d RK3 steps of the same %dx%d shear layer from its initial state (%.1f s), strict-IEEE CPU "
                      "restatement of the reference in its MPI form, %dx%d blocks, one rank per thread"
                      % (steps, n, n, wall, dims[0][0], dims[0][1])}
    if args.kh_control:
        # the flow's own sensitivity: FP64 from a density perturbed by 1e-14
        # (relative), compared with the FP64 run at the same times
        rp = run_kh(cfs.PREC_FP64, n, t_end, checkpoints=cps, perturb=1e-14)
        out["l1_density_vs_time"] = [
            {"t": tc, "fp32_bulk_vs_fp64": float(np.abs(r32["snaps"][tc] - r64["snaps"][tc]).mean()),
             "fp64_perturbed_1e-14_vs_fp64": float(np.abs(rp["snaps"][tc] - r64["snaps"][tc]).mean())}
            for tc in sorted(r64["snaps"]) if tc in r32["snaps"] and tc in rp["snaps"]]
    print(json.dumps(out))


SERVE_SEED = 7


def run_sequential(reqs, precision):
    """The requests one after another, each alone through the single-solve
    entry points, in the given precision; returns (wall seconds, failed)."""
    t0 = time.perf_counter()
    failed = 0
    for rq in reqs:
        cfg = cfs.Config.from_buffer_copy(rq["cfg"])
        cfg.precision = precision
        s = cfs.Solver(cfg)
        if rq["ib"] is not None:
            s.set_immersed(rq["ib"])
        s.set_state(rq["U"])
        try:
            s.advance(100000, dt=0.0, cfl=rq["cfl"], t_end=rq["t_end"])
        except cfs.CfsError:
            failed += 1
        s.close()
    return time.perf_counter() - t0, failed


def run_serving(prec, count, rates, rate_count=None, arena_gb=4, slots=64, steps_per_poll=5):
    """The serving workload: the seeded request stream served with continuous
    batching, once per offered rate (0: a burst, every request at t = 0; the
    burst uses `count` requests, the Poisson rates the first `rate_count`),
    then the burst's requests one after another in FP64 through the
    single-solve path.  Returns the measurement points and the ratio of the
    burst's solves/s to the one-after-another FP64 solves/s."""
    arena = arena_gb << 30
    rate_count = rate_count or count
    points = []
    sample = cfs.serving_requests(SERVE_SEED + 1, 16, rate=1e9, precision=prec)
    shape, shape_ms = cfs.batch_autotune(sample, rk=2, precision=prec, limiter=True, steps=6, arena_bytes=arena,
                                         max_problems=slots)
    print("serve autotune: shape %d, ms per step %s" % (shape, shape_ms), file=sys.stderr, flush=True)
    for rate in rates:
        burst = rate <= 0
        n = count if burst else rate_count
        reqs = cfs.serving_requests(SERVE_SEED, n, rate=1e9 if burst else rate, precision=prec)
        if burst:
            for rq in reqs:
                rq["arrival_s"] = 0.0
        with cfs.Batch(rk=2, precision=prec, max_problems=slots, arena_bytes=arena, limiter=True,
                       tile_shape=shape) as b:
            res = b.serve(reqs, steps_per_poll=steps_per_poll, keep_states=False)
        lat = np.array([r["latency_s"] for r in res])
        span = max(r["finished_s"] for r in res) - min(rq["arrival_s"] for rq in reqs)
        failed = sum(r["state"] == cfs.SLOT_FAILED for r in res)
        pt = {"offered_per_s": "burst" if burst else rate, "requests": n, "solves_per_s": round(n / span, 3),
              "p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
              "makespan_s": round(span, 3), "failed": int(failed)}
        points.append(pt)
        print("serve: %s" % json.dumps(pt), file=sys.stderr, flush=True)
    # the burst's requests one after another in FP64
    reqs = cfs.serving_requests(SERVE_SEED, count, rate=1e9, precision=prec)
    seq64, f64 = run_sequential(reqs, cfs.PREC_FP64)
    print("one after another, fp64: %.2f s (%d failed)" % (seq64, f64), file=sys.stderr, flush=True)
    burst = [p for p in points if p["offered_per_s"] == "burst"]
    cap = burst[0]["solves_per_s"] if burst else max(p["solves_per_s"] for p in points)
    return {
        "workload": "serving_riemann128-1024_cylinder256x128-1024x512_mach1.5-4_t0.1-0.5_rk2_cfl0.5",
        "precision": "fp64" if prec == cfs.PREC_FP64 else "fp32_bulk",
        "data": "synthetic (seeded request stream, seed %d)" % SERVE_SEED,
        "arena_gb": arena_gb, "slots": slots, "steps_per_poll": steps_per_poll,
        "tile_shape": shape_report(dict(shape=shape, shape_ms=shape_ms)),
        "solves_per_s": cap,
        "points": points,
        "sequential_fp64": {"requests": count, "wall_s": round(seq64, 3), "solves_per_s": round(count / seq64, 3),
                            "failed": f64},
        "vs_sequential_fp64": round(cap / (count / seq64), 3),
        "target_vs_sequential_fp64": 3.0,
    }


def main_serve(args):
    prec = cfs.PREC_FP64 if args.batch_fp64 else cfs.PREC_FP32_BULK
    r = run_serving(prec, args.serve_count, [float(x) for x in args.serve_rates.split(",")],
                    arena_gb=args.serve_arena_gb, slots=args.serve_slots, steps_per_poll=args.steps_per_poll)
    out = {
        "metric": "solves/s (serving, continuous batching)",
        "value": r["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64" if prec == cfs.PREC_FP64 else "f32 (FP64 state)",
        "data": r["data"],
        "config": {k: r[k] for k in ("workload", "arena_gb", "slots", "steps_per_poll", "tile_shape")},
        "points": r["points"],
        "sequential_fp64": r["sequential_fp64"],
        "vs_sequential_fp64": r["vs_sequential_fp64"],
    }
    print(json.dumps(out))


def main_ibsetup(args):
    """Immersed-boundary set-up latency, host (cfs_circle_setup) against device
    (cfs_circle_setup_device), median of 15 calls after a warm-up call, on
    the headline body and on served cylinder sizes; the tables are compared
    bit for bit on every geometry."""
    cases = [("cylinder_2048x1024_r0.125", cfs.config_cylinder(2048, 1024, 4.0, 2.0), 0.125),
             ("cylinder_1024x512_r0.1", cfs.config_cylinder(1024, 512, 4.0, 2.0), 0.1),
             ("cylinder_512x256_r0.2", cfs.config_cylinder(512, 256, 4.0, 2.0), 0.2),
             ("cylinder_4096x2048_r0.125", cfs.config_cylinder(4096, 2048, 4.0, 2.0), 0.125)]
    rows = []
    for name, cfg, r in cases:
        cfs.circle_setup(cfg, 1.0, 1.0, r, device=True)   # warm-up: module load, first launches
        th, td = [], []
        for _ in range(15):
            t0 = time.perf_counter()
            h = cfs.circle_setup(cfg, 1.0, 1.0, r)
            t1 = time.perf_counter()
            d = cfs.circle_setup(cfg, 1.0, 1.0, r, device=True)
            t2 = time.perf_counter()
            th.append(t1 - t0)
            td.append(t2 - t1)
        same = all(np.array_equal(np.asarray(h[k]).view(np.uint8), np.asarray(d[k]).view(np.uint8))
                   for k in ("cell_type", "ig_cells", "donors", "image_xy", "coef", "transp"))
        rows.append({"workload": name, "ghosts": d["n_ig"], "finished_on_host": d["n_host"],
                     "host_ms": round(1e3 * float(np.median(th)), 3), "device_ms": round(1e3 * float(np.median(td)), 3),
                     "identical": bool(same)})
        print("ibsetup: %s" % json.dumps(rows[-1]), file=sys.stderr, flush=True)
    print(json.dumps({"metric": "immersed-boundary set-up latency (ms, median of 15)", "unit": "ms",
                      "higher_is_better": False, "n_gpus": 1, "data": "circle bodies (no randomness)",
                      "rows": rows}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-bulk leg")
    ap.add_argument("--no-reference-scheme", action="store_true", help="skip the no-limiter leg")
    ap.add_argument("--fp32-only", action="store_true", help="time the FP32-bulk leg only (profiling)")
    ap.add_argument("--no-horizon", action="store_true",
                    help="skip the FP32-bulk accuracy comparison at the workload's 2000-step horizon")
    ap.add_argument("--no-kh", action="store_true", help="skip the 4096^2 shear-layer leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-laxliu", action="store_true", help="skip the Lax-Liu 400^2 leg")
    ap.add_argument("--workload", choices=["cylinder", "batch", "serve", "kh", "riemann", "ibsetup"],
                    default="cylinder",
                    help="cylinder: the headline line; batch / serve: the batched-solves and serving measurements")
    ap.add_argument("--kh-n", type=int, default=4096, help="shear-layer grid side (--workload kh)")
    ap.add_argument("--kh-t-end", type=float, default=1.0, help="shear-layer end time (--workload kh)")
    ap.add_argument("--kh-control", action="store_true",
                    help="also record L1(rho) at t = 0.1, 0.25, 0.5 and run FP64 from a 1e-14-perturbed density")
    ap.add_argument("--serve-count", type=int, default=48)
    ap.add_argument("--serve-rates", default="1,2,4,0", help="offered request rates per second (0: burst)")
    ap.add_argument("--serve-arena-gb", type=int, default=4)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--steps-per-poll", type=int, default=5)
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-fp64", action="store_true", help="batched workload in FP64 instead of FP32-bulk")
    ap.add_argument("--batch-only", action="store_true",
                    help="batched leg only (no one-after-another leg, no FP64 comparison; profiling)")
    ap.add_argument("--tile-shape", type=int, default=None, choices=range(cfs.NUM_TILE_SHAPES),
                    help="stage-kernel tile shape (default: picked by the autotuner)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run N plain copy-kernel launches over one state plane-set (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the multi-GPU design is a 2D block decomposition with xGMI halo exchange, not built)",
              file=sys.stderr)
        sys.exit(2)
    if cfs.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batch":
        main_batch(args)
        return
    if args.workload == "serve":
        main_serve(args)
        return
    if args.workload == "kh":
        main_kh(args)
        return
    if args.workload == "riemann":
        main_riemann(args)
        return
    if args.workload == "ibsetup":
        main_ibsetup(args)
        return
    if args.copy_calibration:
        cfg, ib, U = setup(cfs.PREC_FP64)
        s = cfs.Solver(cfg)
        s.set_state(U)
        rc = cfs.library().cfs_calibrate_copy(s._h, args.copy_calibration)
        print(json.dumps({"copy_calibration": {"launches": args.copy_calibration, "rc": rc,
                                               "bytes_read_per_launch": 32 * NX * NY,
                                               "bytes_written_per_launch": 32 * NX * NY}}))
        return
    if args.fp32_only:
        r = run_gpu(cfs.PREC_FP32_BULK, args.steps, args.warmup, tile_shape=args.tile_shape)
        r.pop("rho")
        print(json.dumps({"fp32_bulk_only": r}))
        return

    r64 = run_gpu(cfs.PREC_FP64, args.steps, args.warmup, tile_shape=args.tile_shape)
    if not r64["finite"]:
        print("FP64 state not finite at the end of the timed window", file=sys.stderr)
    print("fp64: %d steps, %.3f s, %.1f Mcell-updates/s, stage kernel %.4f ms, tile shape %d (autotune ms %s)"
          % (r64["steps"], r64["wall"], r64["mcups"], r64["kernel_ms"], r64["shape"], r64["shape_ms"]),
          file=sys.stderr, flush=True)
    r32 = None
    acc = None
    if not args.no_fp32:
        r32 = run_gpu(cfs.PREC_FP32_BULK, args.steps, args.warmup, tile_shape=args.tile_shape)
        print("fp32-bulk: %.1f Mcell-updates/s" % r32["mcups"], file=sys.stderr, flush=True)
        # accuracy at the workload's horizon (outside every timed window)
        if not args.no_horizon:
            acc = run_accuracy()
    rref = None
    if not args.no_reference_scheme:
        rref = run_gpu(cfs.PREC_FP64, min(args.steps, 500), args.warmup, reference_scheme=True,
                       tile_shape=args.tile_shape)
        print("reference scheme (no limiter, subsonic inflow): %.1f Mcell-updates/s, stage kernel %.4f ms"
              % (rref["mcups"], rref["kernel_ms"]), file=sys.stderr, flush=True)
    kh = None
    if not args.no_kh:
        kh = run_kh_short()
    laxliu = None
    if not args.no_laxliu:
        laxliu = run_laxliu()
        print("lax-liu 400^2 to t = 0.3: %.1f Mcell-updates/s (%d steps, %.4f s)"
              % (laxliu["value"], laxliu["steps"], laxliu["wall_s"]), file=sys.stderr, flush=True)
    serving = None
    if not args.no_serve:
        serving = run_serving(cfs.PREC_FP32_BULK, 48, [0.0, 2.0, 4.0], rate_count=32)
    cpu = None
    if not args.no_cpu:
        cpu = run_cpu_baseline()
        print("cpu baseline: %.2f Mcell-updates/s on %d threads" % (cpu["value"], cpu["cores"]),
              file=sys.stderr, flush=True)

    cells = NX * NY
    bytes_per_launch = cells * BYTES_PER_CELL_STAGE
    achieved = bytes_per_launch / (r64["kernel_ms"] * 1e-3) / 1e9 if r64["kernel_ms"] > 0 else 0.0
    out = {
        "metric": "Mcell-updates/s per RK stage",
        "value": round(r64["mcups"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform Mach-3 freestream rho=1.4, u=3, p=1 on fluid cells; no seed needed)",
        "config": {"workload": "cylinder_ib_mach3_2048x1024_rk3_cfl0.5_fp64", "nx": NX, "ny": NY,
                   "inflow": {"rho": FREESTREAM[0], "u": FREESTREAM[1], "v": FREESTREAM[2], "p": FREESTREAM[3]},
                   "positivity_limiter": True, "hll_bounds": "reference",
                   "domain": [LX, LY], "cylinder": {"centre": [CX, CY], "radius": RADIUS,
                                                    "lagrangian_points": 400, "ghost_layers": 3},
                   "rk": 3, "cfl": CFL, "precision": "fp64", "stages_per_step": 3,
                   "ib_interpolation": "bilinear_positive", "finite_at_end": r64["finite"],
                   "tile_shape": shape_report(r64)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profile_figure("traffic_bytes_per_launch", r64["shape"]),
                     "kernel": kernel_name(cfs.PREC_FP64, True, r64["shape"]),
                     "kernel_ms": round(r64["kernel_ms"], 5),
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    lane_ops = profile_figure("valu_lane_ops_per_cell_stage", r64["shape"])
    if lane_ops and r64["kernel_ms"] > 0:
        # the bound that binds: FP64 VALU issue (the kernel's VALU lane-ops per
        # cell-stage from the committed counters, at the measured launch duration)
        v = lane_ops * cells / (r64["kernel_ms"] * 1e-3) / 1e12
        out["roofline"]["valu"] = {"achieved": round(v, 2), "peak": FP64_VALU_PEAK_T, "unit": "T lane-ops/s",
                                   "frac": round(v / FP64_VALU_PEAK_T, 4), "lane_ops_per_cell_stage": lane_ops}
    if rref is not None:
        out["reference_scheme"] = {"value": round(rref["mcups"], 2), "unit": "Mcell-updates/s",
                                   "kernel": kernel_name(cfs.PREC_FP64, False, rref["shape"]),
                                   "kernel_ms": round(rref["kernel_ms"], 5),
                                   "workload": "same grid and body, inflow rho=u=p=1, the reference's WENO5 + HLL + "
                                               "TVD-RK3 with no limiter, %d steps" % rref["steps"],
                                   "ib_interpolation": "bilinear_positive (the reference's LDLT interpolation turns "
                                                       "this flow non-finite within a few steps)"}
    if r32 is not None:
        out["fp32_bulk"] = {"value": round(r32["mcups"], 2), "unit": "Mcell-updates/s",
                            "ms_per_step": round(r32["ms_per_step"], 4),
                            "kernel_ms": round(r32["kernel_ms"], 5),
                            "kernel": kernel_name(cfs.PREC_FP32_BULK, True, r32["shape"]),
                            "tile_shape": shape_report(r32),
                            "speedup_vs_fp64": round(r32["mcups"] / r64["mcups"], 3)}
        if acc is not None:
            # at the time FP64 reaches after the workload's 2000 steps; the
            # FP32-bulk run takes the CFL steps of its own state, the last one
            # clipped to that time.  The control shows what a perturbation of
            # FP32 rounding size does to the FP64 solution of this flow.
            out["fp32_bulk"].update(acc)
    if kh is not None:
        out["kh"] = kh
    if laxliu is not None:
        out["laxliu"] = laxliu
    if serving is not None:
        out["serving"] = serving
    print(json.dumps(out))


if __name__ == "__main__":
    main()
