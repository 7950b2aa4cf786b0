"""Benchmark of the per-RK-stage hot path on one MI355X.

Workload (the headline configuration that fits one GPU): inviscid flow past
a circular cylinder with the ghost-cell immersed boundary, 2048 x 1024 cells
on [0,4] x [0,2], r = 0.125 at (1,1), 400 Lagrangian points, 3 ghost layers,
Dirichlet inflow on the left and extrapolation (Neumann) elsewhere,
gamma = 1.4, TVD-RK3, CFL 0.5, FP64, the reference's scheme unchanged
(reference HLL bounds, no positivity fallback).  Inflow: the reference's own
cylinder freestream rho = 1, u = 1, p = 1 (Cylinder.cpp:5-16): with the
reference's scheme a Mach-3 inflow turns non-finite within the first 100
steps at this resolution, and this inflow within ~800, so the timed window is
520 steps (20 warm-up + 500).  The image-point interpolation is the
positivity-preserving bilinear one (the reference's LDLT-on-lower-triangle
interpolation gives negative ghost densities on this grid at step 1).
One bench "step" = one RK3 step (3 stages: IB fill + fused
WENO5/HLL/RK stage kernel each, plus the device CFL reduction).

Metric: Mcell-updates/s per RK stage = cells * stages / wall / 1e6, with the
state resident in HBM.  The FP32-bulk form (FP64 state, FP32 reconstruction /
flux / RHS) is timed on the same workload and reported beside it.

roofline: the stage kernel, algorithmic bytes per launch = cells * 97 B
(RK3 mean 96 B/cell-stage of FP64 state planes + 1 B cell-type mask) over its
mean launch duration measured with HIP events on the library's stream.
cpu_baseline: the strict-IEEE CPU restatement of the reference (oracle/) on
the same workload, all host threads (up to 16), a bounded number of steps.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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
NX, NY, LX, LY = 2048, 1024, 4.0, 2.0
CX, CY, RADIUS = 1.0, 1.0, 0.125
CFL = 0.5
FREESTREAM = (1.0, 1.0, 0.0, 1.0)
BYTES_PER_CELL_STAGE = 97.0


def traffic_from_profiles():
    """HBM bytes per stage-kernel launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), from the
    committed summary; None when absent."""
    p = os.path.join(ROOT, "profiles", "stage_kernel_fp64_cylinder_2048x1024_traffic.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def setup(precision):
    cfg = cfs.config_cylinder(NX, NY, LX, LY, freestream=FREESTREAM, precision=precision, positivity=0,
                              hll_bounds=0)
    ib = cfs.circle_setup(cfg, CX, CY, RADIUS, 400, 3, interp=cfs.IB_BILINEAR_POSITIVE)
    U = cfs.cylinder_freestream(cfg, ib["cell_type"], freestream=FREESTREAM)
    return cfg, ib, U


def run_gpu(precision, steps, warmup):
    cfg, ib, U = setup(precision)
    s = cfs.Solver(cfg)
    s.set_immersed(ib)
    s.set_state(U)
    if warmup:
        s.advance(warmup, dt=0.0, cfl=CFL)
    t0 = time.perf_counter()
    t, done = s.advance(steps, dt=0.0, cfl=CFL)
    wall = time.perf_counter() - t0
    # stage-kernel duration: a separate pass with HIP events around every
    # stage kernel (the events add queue latency, so not in the timed window)
    s.set_kernel_timing(True)
    s.advance(min(steps, 40), dt=0.0, cfl=CFL)
    kms = s.last_kernel_ms()
    V = s.get_state()
    ok = bool(np.all(np.isfinite(V)))
    cells = NX * NY
    stages = 3 * done
    s.close()
    return dict(wall=wall, steps=done, mcups=cells * stages / wall / 1e6, kernel_ms=kms,
                ms_per_step=1e3 * wall / max(done, 1), finite=ok, t=t)


def run_cpu_baseline(budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    cfg, ib, U = setup(cfs.PREC_FP64)
    nthreads = max(1, min(16, os.cpu_count() or 1))
    o = O.Oracle(cfg, nthreads=nthreads)
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
                kind="port",
                sample="%d RK3 steps (%d stages incl. per-step CFL) of the same 2048x1024 cylinder "
                       "workload, %.1f s, strict-IEEE CPU restatement of the reference" % (steps, 3 * steps, wall))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-bulk leg")
    ap.add_argument("--fp32-only", action="store_true", help="time the FP32-bulk leg only (profiling)")
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
        r = run_gpu(cfs.PREC_FP32_BULK, args.steps, args.warmup)
        print(json.dumps({"fp32_bulk_only": r}))
        return

    r64 = run_gpu(cfs.PREC_FP64, args.steps, args.warmup)
    if not r64["finite"]:
        print("FP64 state not finite at the end of the timed window", file=sys.stderr)
    print("fp64: %d steps, %.3f s, %.1f Mcell-updates/s, stage kernel %.4f ms"
          % (r64["steps"], r64["wall"], r64["mcups"], r64["kernel_ms"]), file=sys.stderr, flush=True)
    r32 = None
    if not args.no_fp32:
        r32 = run_gpu(cfs.PREC_FP32_BULK, args.steps, args.warmup)
        print("fp32-bulk: %.1f Mcell-updates/s" % r32["mcups"], file=sys.stderr, flush=True)
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
        "data": "synthetic (uniform freestream rho=1, u=1, p=1 on fluid cells; no seed needed)",
        "config": {"workload": "cylinder_ib_2048x1024_rk3_cfl0.5_fp64", "nx": NX, "ny": NY,
                   "inflow": {"rho": 1.0, "u": 1.0, "v": 0.0, "p": 1.0},
                   "domain": [LX, LY], "cylinder": {"centre": [CX, CY], "radius": RADIUS,
                                                    "lagrangian_points": 400, "ghost_layers": 3},
                   "rk": 3, "cfl": CFL, "precision": "fp64", "stages_per_step": 3,
                   "ib_interpolation": "bilinear_positive", "finite_at_end": r64["finite"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "kernel": "stage_kernel<double>", "kernel_ms": round(r64["kernel_ms"], 5),
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    if r32 is not None:
        out["fp32_bulk"] = {"value": round(r32["mcups"], 2), "unit": "Mcell-updates/s",
                            "ms_per_step": round(r32["ms_per_step"], 4),
                            "kernel_ms": round(r32["kernel_ms"], 5),
                            "speedup_vs_fp64": round(r32["mcups"] / r64["mcups"], 3)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
