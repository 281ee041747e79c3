#!/usr/bin/env python3
"""RK4 time-step throughput of the axisymmetric jet on one MI355X.

Workload: the plain FP64 jet at 4400 x 640 (z x r), Mach 1.5, Re 1e4, Pr 0.7, seeded
hyperbolic-tangent top-hat initial condition, SAT inflow / NSCBC top and outflow / axis
boundaries, CFL 0.5.  A step is one {ComputeTimeStep; RK4} iteration (ModSolver.f90:122-125)
over the whole grid, with the state resident in HBM.  Prints one JSON line.

  value      grid-point updates/s = nz * nr * steps / time of the timed region
  roofline   the four RK-stage kernel launches of a step against the HBM peak: the
             algorithmic traffic is 640 B per grid-point update (5 FP64 variables; stage 1
             reads q^n and writes q1, A; stage 2 reads q1, q^n and writes q2, S; stage 3 reads
             q2, q^n, S and writes q3, S; stage 4 reads q3, A, S and writes q^{n+1}), divided by
             the summed average durations of the four stage launches measured with HIP events
  cpu_baseline  the C restatement (oracle) of the same step on the host cores, OpenMP over
             grid lines, on a bounded sample of steps of the same grid
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_UPDATE = 640
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(case, seconds: float, threads: int):
    from oracle import pyoracle as po
    o = po.Oracle(case, threads=threads)
    o.run(1)  # warm-up (first touch of the ~1.6 GB of halo fields)
    n, t0 = 0, time.perf_counter()
    while True:
        o.run(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    o.close()
    return case.nz * case.nr * n / el, n, el


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nz", type=int, default=4400)
    ap.add_argument("--nr", type=int, default=640)
    ap.add_argument("--stage-steps", type=int, default=20, help="steps timed launch by launch for the roofline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device copy of this many MiB (a known-size kernel that scales "
                         "hardware byte counters)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (use --gpus 1)",
              file=sys.stderr)
        return 2

    import torch
    from axisymmetricjet_amd import capi
    from axisymmetricjet_amd import jetcase as jc

    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written")
        del src, dst

    nz, nr = args.nz, args.nr
    npts = nz * nr
    log(f"building case {nz}x{nr}")
    case = jc.make_jet(nz, nr, seed=1234)
    ctx = capi.Context(case)
    stream = torch.cuda.ExternalStream(ctx.stream())
    log(f"warm-up {args.warmup} steps")
    ctx.run(args.warmup)
    torch.cuda.synchronize()

    log(f"timed region: {args.steps} steps")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    ctx.run(args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    dts, t = ctx.dt_log()
    if not (dts[-1] > 0 and dts[-1] < 1.0):
        raise RuntimeError(f"implausible time step {dts[-1]}")
    value = npts * args.steps / (ms / 1e3)
    log(f"{ms / args.steps:.4f} ms/step, {value:.4e} updates/s, dt[-1]={dts[-1]:.6e}")

    # the four stage kernels, each launch bracketed by events on the library's stream
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(4)]
          for _ in range(args.stage_steps)]
    for k in range(args.stage_steps):
        ctx.compute_dt(sync=False)
        for s in range(4):
            ev[k][s][0].record(stream)
            ctx.rk4_stage(s + 1)
            ev[k][s][1].record(stream)
    torch.cuda.synchronize()
    stage_ms = [sum(ev[k][s][0].elapsed_time(ev[k][s][1]) for k in range(args.stage_steps)) / args.stage_steps
                for s in range(4)]
    step_kernel_ms = sum(stage_ms)
    achieved = BYTES_PER_UPDATE * npts / (step_kernel_ms / 1e3) / 1e9
    log("stage kernel avg ms: " + ", ".join(f"{x:.4f}" for x in stage_ms))

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle, {threads} threads, ~{args.cpu_seconds:.0f} s")
        v, n, el = cpu_baseline(case, args.cpu_seconds, threads)
        cpu = {"value": v, "unit": "grid-point updates/s", "cores": threads, "kind": "port",
               "sample": f"{n} RK4 steps of the same {nz}x{nr} jet (after 1 warm-up step), {el:.1f} s, "
                         f"C restatement at -O2 without FMA, OpenMP over grid lines"}
        log(f"cpu baseline {v:.4e} updates/s")

    out = {
        "metric": "RK4 grid-pt updates/s at 4400x640 FP64, %HBM roofline; learned-LAD vs plain",
        "value": value,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 axisymmetric jet RK4 step, {nz}x{nr} (z x r), Mach 1.5, "
                               f"Re 1e4, Pr 0.7, CFL 0.5, SAT/NSCBC/axis boundaries",
                   "nz": nz, "nr": nr, "points": npts, "method": "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "stage_kernel<1..4> (one launch per RK stage)",
                     "stage_ms": stage_ms, "bytes_per_update": BYTES_PER_UPDATE},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
