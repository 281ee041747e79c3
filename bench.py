#!/usr/bin/env python3
"""RK4 time-step throughput of the axisymmetric jet on one MI355X.

Workload: the FP64 jet at 4400 x 640 (z x r), Mach 1.5, Re 1e4, Pr 0.7, seeded
hyperbolic-tangent top-hat initial condition, SAT inflow / NSCBC top and outflow / axis
boundaries, CFL 0.5, with shock capturing and filtering on (ShockCapturing = 1: the localized
artificial diffusivity with C_mu, C_beta, C_kappa = jetcase.LAD_C, the TG filters in z and r).
A step is one {ComputeTimeStep; RK4} iteration (ModSolver.f90:122-125) over the whole grid,
with the state resident in HBM.  The plain step (ShockCapturing = 0) is timed beside it.
Prints one JSON line.

  value      grid-point updates/s = nz * nr * steps / time of the timed region
  roofline   the four RK stages of a step (each: the artificial-diffusivity launches, then the
             stage kernel) against the HBM peak: the algorithmic traffic is 640 B per grid-point
             update (5 FP64 variables; stage 1 reads q^n and writes q1, A; stage 2 reads q1, q^n
             and writes q2, S; stage 3 reads q2, q^n, S and writes q3, S; stage 4 reads q3, A, S
             and writes q^{n+1}), divided by the summed average durations of the four stages
             measured with HIP events on the library's stream.  traffic is the HBM bytes per
             step measured with rocprofv3 PMC counters (FETCH_SIZE, WRITE_SIZE) for the same
             command, kept in profiles/traffic.json (null when the sources changed since)
  cpu_baseline  the C restatement (oracle) of the same step on the host cores, OpenMP over
             grid lines, on a bounded sample of steps of the same grid
  learned    the same jet with the learned LAD coefficient (ajet_set_learned_lad, the weights in
             axisymmetricjet_amd/data): updates/s and the speed-up over the reference's LAD, the
             inference kernel against the BF16 MFMA peak (92.7 kFLOP per point evaluated:
             2 (104 x 128 + 2 x 128 x 128 + 128 x 2)), and a scoring run from the same initial
             condition -- far-field OASPL and spectrum on an arc of probes, compression peaks on
             the axis and lip line -- against the reference-LAD run and against the learned mode
             without the network (zero coefficients outside the sensor's blocks)
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
BF16_PEAK_TFLOPS = 2500.0   # dense MFMA
MLP_FLOP_PER_POINT = 2 * (104 * 128 + 2 * 128 * 128 + 128 * 2)

# HBM bytes per grid-point update of a whole step, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE over the same workload (bench.py --traffic-from), kept in profiles/traffic.json
# with a digest of the library sources they were measured on; a digest that no longer matches
# the sources gives traffic = null rather than a stale number.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
CALIB_MIB = 1024


def sources_digest():
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "axisymmetricjet_amd", "csrc")
    for f in sorted(os.listdir(d)) + ["../build.py"]:
        with open(os.path.join(d, f), "rb") as fh:
            h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()


def measured_traffic(method):
    """HBM bytes per grid-point update of `method` from profiles/traffic.json, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    if doc.get("sources_sha256") != sources_digest():
        return None
    return doc.get("bytes_per_update", {}).get(method)


def traffic_from(root, methods, points):
    """profiles/traffic.json from the rocprofv3 counter runs under root: for each method,
    root/pmc_<method>_FETCH_SIZE and root/pmc_<method>_WRITE_SIZE hold the counter CSVs of
    `bench.py --method <method> ... --calib-copy-mib 1024`.  Each counter is scaled by what it
    reads for that run's 1 GiB device copy (the dispatch whose name holds "copy" with the largest
    count), and the per-step bytes are the step's kernels summed over the dispatches of the run
    divided by its stage-4 launches (one per step)."""
    import csv
    out = {"sources_sha256": sources_digest(), "points": points, "bytes_per_update": {}, "scale": {}}
    for m in methods:
        per = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            path = os.path.join(root, f"pmc_{m}_{counter}", "run_counter_collection.csv")
            rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
            calib = max((r for r in rows if "copy" in r["Kernel_Name"].lower()), key=lambda r: float(r["Counter_Value"]))
            scale = CALIB_MIB * 1024.0 / float(calib["Counter_Value"])
            step = [r for r in rows if any(k in r["Kernel_Name"] for k in ("stage_", "lad_", "mlp_", "cfl_"))]
            nsteps = sum(1 for r in step if "<4," in r["Kernel_Name"] and ("true" in r["Kernel_Name"] or "mixed" in r["Kernel_Name"]))
            kib = sum(float(r["Counter_Value"]) for r in step)
            per[counter] = kib * 1024.0 * scale / max(nsteps, 1) / points
            out["scale"][f"{m}_{counter}"] = scale
        out["bytes_per_update"][m] = per["FETCH_SIZE"] + per["WRITE_SIZE"]
        out.setdefault("split", {})[m] = per
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(out, f, indent=1)
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, seconds: float, threads: int, lad_c):
    from oracle import pyoracle as po
    c = lad_c or (0.0, 0.0, 0.0)
    o = po.Oracle(case, shock=1 if lad_c else 0, c_mu=c[0], c_beta=c[1], c_kappa=c[2], threads=threads)
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
    ap.add_argument("--method", choices=["lad", "plain", "learned"], default="lad",
                    help="lad: ShockCapturing = 1 with the LAD constants; plain: ShockCapturing = 0; learned: "
                         "lad with the learned coefficient (the --learned-* options)")
    ap.add_argument("--other-steps", type=int, default=200, help="steps of the other method timed beside")
    ap.add_argument("--stage-steps", type=int, default=20, help="steps timed launch by launch for the roofline")
    ap.add_argument("--ensemble-runs", type=int, default=64, help="runs of the batched ensemble (0: skip it)")
    ap.add_argument("--ensemble-steps", type=int, default=400,
                    help="the ensemble's runs diverge after several hundred steps under the reference's "
                         "numerics, so 400 is the default rather than 2000")
    ap.add_argument("--learned-steps", type=int, default=1000, help="steps of the learned-LAD mode timed (0: skip)")
    ap.add_argument("--learned-refresh", type=int, default=12, help="the network runs every this many steps")
    ap.add_argument("--sensor-threshold", type=float, default=0.1)
    ap.add_argument("--ramp-min", type=float, default=1e-8)
    ap.add_argument("--score-steps", type=int, default=4000,
                    help="steps of the learned-LAD scoring runs of the ideally expanded jet (0: skip)")
    ap.add_argument("--score-warm", type=int, default=1500,
                    help="steps under the reference LAD before the scoring record (the jet develops; under the "
                         "reference's numerics the runs leave the finite range near step 5800 from the initial "
                         "condition, so warm-up + record stay below that)")
    ap.add_argument("--shocked-score-steps", type=int, default=1400,
                    help="steps of the scoring runs of the under-expanded jet, npr 1.5 (0: skip; finite to ~2200 "
                         "steps from its initial condition)")
    ap.add_argument("--shocked-warm", type=int, default=750)
    ap.add_argument("--shocked-timing-steps", type=int, default=400,
                    help="steps of the learned and the reference-LAD mode timed on the under-expanded jet (npr 1.5) "
                         "from its state after --shocked-warm steps, where the sensor flags blocks (0: skip)")
    ap.add_argument("--combined-long-steps", type=int, default=0,
                    help="one long run of the combined batch (the configured 10000 steps): learned and reference-LAD "
                         "batches with probes every 8 steps, ms/step, the step each run leaves the finite range, "
                         "OASPL difference over the record all runs share (0: skip)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "axisymmetricjet_amd", "data", "learned_lad_4400x640.npz"))
    ap.add_argument("--combined-runs", type=int, default=32,
                    help="runs of the combined form: learned-LAD jets at the bench grid with a forcing sweep (0: skip)")
    ap.add_argument("--combined-steps", type=int, default=50, help="steps of the combined batch timed")
    ap.add_argument("--combined-score-steps", type=int, default=400,
                    help="steps of the combined batch's OASPL comparison with the same batch under the reference LAD")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-from", default=None,
                    help="write profiles/traffic.json from the rocprofv3 --pmc runs under this folder and exit")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device copy of this many MiB (a known-size kernel that scales "
                         "hardware byte counters)")
    args = ap.parse_args(argv)
    if args.traffic_from:
        print(json.dumps(traffic_from(args.traffic_from, ("plain", "lad", "learned"), args.nz * args.nr)))
        return 0
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

    def context(method):
        if method in ("lad", "learned"):
            c = jc.LAD_C
            ctx = capi.Context(case, shock_capturing=1, C_mu=c[0], C_beta=c[1], C_kappa=c[2])
            if method == "learned":
                from axisymmetricjet_amd import learned_lad as ll
                ctx.set_learned_lad(ll.load(args.weights), args.sensor_threshold, args.ramp_min, args.learned_refresh)
            return ctx
        return capi.Context(case)

    def timed_run(ctx, steps, warmup):
        stream = torch.cuda.ExternalStream(ctx.stream())
        ctx.run(warmup)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.run(steps)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        dts, _ = ctx.dt_log()
        if not (dts[-1] > 0 and dts[-1] < 1.0):
            raise RuntimeError(f"implausible time step {dts[-1]}")
        return ms, dts[-1]

    other = "plain" if args.method != "plain" else "lad"
    others = {}
    if args.other_steps > 0:
        octx = context(other)
        oms, _ = timed_run(octx, args.other_steps, args.warmup)
        others[other] = {"value": npts * args.other_steps / (oms / 1e3), "ms_per_step": oms / args.other_steps,
                         "steps": args.other_steps}
        log(f"{other}: {oms / args.other_steps:.4f} ms/step")
        octx.close()

    ctx = context(args.method)
    stream = torch.cuda.ExternalStream(ctx.stream())
    log(f"{args.method}: warm-up {args.warmup} steps, timed region {args.steps} steps")
    ms, dt_last = timed_run(ctx, args.steps, args.warmup)
    value = npts * args.steps / (ms / 1e3)
    log(f"{ms / args.steps:.4f} ms/step, {value:.4e} updates/s, dt[-1]={dt_last:.6e}")

    # the four stages, each bracketed by events on the library's stream
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
    log("stage avg ms: " + ", ".join(f"{x:.4f}" for x in stage_ms))

    learned = None
    lad_ms = ms / args.steps if args.method == "lad" else others.get("lad", {}).get("ms_per_step")
    if args.learned_steps > 0 and os.path.exists(args.weights):
        from axisymmetricjet_amd import learned_lad as ll
        model = ll.load(args.weights)
        lctx = context("lad")
        lctx.set_learned_lad(model, args.sensor_threshold, args.ramp_min, args.learned_refresh)
        lms, _ = timed_run(lctx, args.learned_steps, args.warmup)
        frac = lctx.fallback_fraction()
        lctx.close()

        # the inference kernel: stage 1 with the network (refresh 1) minus stage 1 without it
        def stage1_ms(refresh, n=10):
            c = context("lad")
            c.set_learned_lad(model, args.sensor_threshold, args.ramp_min, refresh)
            c.run(1)
            st = torch.cuda.ExternalStream(c.stream())
            tot = 0.0
            for _ in range(n):
                c.compute_dt(sync=False)
                a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a0.record(st)
                c.rk4_stage(1)
                a1.record(st)
                for s in (2, 3, 4):
                    c.rk4_stage(s)
                torch.cuda.synchronize()
                tot += a0.elapsed_time(a1)
            c.close()
            return tot / n
        mlp_ms = stage1_ms(1) - stage1_ms(1 << 30)
        import numpy as np
        zc = case.z[:, 0]
        ramp = (1.0 + np.tanh(5.0 * (zc - 10.0))) / 2.0
        last = np.minimum(np.arange(0, nz, 16) + 15, nz - 1)
        live = sum(min(16, nz - i0) for i0, il in zip(range(0, nz, 16), last) if ramp[il] >= args.ramp_min) * nr
        tflops = MLP_FLOP_PER_POINT * live / (mlp_ms / 1e3) / 1e12
        learned = {"value": npts * args.learned_steps / (lms / 1e3), "unit": "grid-point updates/s",
                   "ms_per_step": lms / args.learned_steps, "steps": args.learned_steps,
                   "speedup_vs_lad": (lad_ms / (lms / args.learned_steps)) if lad_ms else None,
                   "refresh": args.learned_refresh, "sensor_threshold": args.sensor_threshold,
                   "ramp_min": args.ramp_min, "fallback_fraction": frac, "weights": os.path.basename(args.weights),
                   "network": "104-128-128-128-2, GELU, BF16 weights and activations, FP32 accumulation",
                   "inference_roofline": {"bound": "mfma", "achieved": tflops, "peak": BF16_PEAK_TFLOPS,
                                          "unit": "TFLOP/s", "frac": tflops / BF16_PEAK_TFLOPS,
                                          "kernel_ms": mlp_ms, "points": live,
                                          "flop_per_point": MLP_FLOP_PER_POINT}}
        log(f"learned: {learned['ms_per_step']:.4f} ms/step, x{learned['speedup_vs_lad'] or 0:.3f} vs LAD, "
            f"inference {mlp_ms:.4f} ms ({tflops:.1f} TFLOP/s), fallback {frac:.4f}")
        if args.shocked_timing_steps > 0:
            # the speed-up where the sensor has work: the under-expanded jet from a developed state
            sj = jc.make_jet(nz, nr, seed=1234, npr=1.5)
            lad_kw = dict(shock_capturing=1, C_mu=jc.LAD_C[0], C_beta=jc.LAD_C[1], C_kappa=jc.LAD_C[2])
            w = capi.Context(sj, **lad_kw)
            w.run(args.shocked_warm)
            start = w.state()
            w.close()
            sh = {"steps": args.shocked_timing_steps, "warm_steps": args.shocked_warm, "sensor_threshold": 0.05}
            for mode in ("lad", "learned"):
                c = capi.Context(sj, **lad_kw)
                c.set_state(start, sj.target)
                if mode == "learned":
                    c.set_learned_lad(model, 0.05, args.ramp_min, args.learned_refresh)
                sms, _ = timed_run(c, args.shocked_timing_steps, 0)
                sh[f"{mode}_ms_per_step"] = sms / args.shocked_timing_steps
                if mode == "learned":
                    sh["fallback_fraction"] = c.fallback_fraction()
                c.close()
            sh["speedup_vs_lad"] = sh["lad_ms_per_step"] / sh["learned_ms_per_step"]
            learned["shocked_npr1.5"] = sh
            log(f"learned on the npr 1.5 jet: {sh['learned_ms_per_step']:.4f} ms/step against "
                f"{sh['lad_ms_per_step']:.4f} under the reference LAD (x{sh['speedup_vs_lad']:.3f}), "
                f"fallback {sh['fallback_fraction']:.4f}")
        if args.score_steps > 0 or args.shocked_score_steps > 0:
            uj = 1.5 * float(np.sqrt(jc.TJ_TINF))
            learned["score"] = {}
            for name, sc, steps, warm, thr in (
                    ("developed", case, args.score_steps, args.score_warm, args.sensor_threshold),
                    ("shocked_npr1.5", jc.make_jet(nz, nr, seed=1234, npr=1.5) if args.shocked_score_steps > 0 else None,
                     args.shocked_score_steps, args.shocked_warm, 0.05)):
                if steps <= 0:
                    continue
                t0 = time.time()
                learned["score"][name] = ll.score(sc, model, thr, steps=steps, every=4, ramp_min=args.ramp_min,
                                                  refresh=args.learned_refresh, warm=warm, uj=uj)
                r = learned["score"][name]
                log(f"learned: {name} jet scored over {steps} steps after {warm} in {time.time() - t0:.1f} s: "
                    f"record {r.get('record_time_D_per_c', 0):.3f} D/c, OASPL error learned "
                    f"{r.get('learned', {}).get('oaspl_err_db_mean', float('nan')):.4f} dB (mean), without the network "
                    f"{r.get('zero', {}).get('oaspl_err_db_mean', float('nan')):.4f} dB")

    ensemble = None
    if args.ensemble_runs > 0 and args.ensemble_steps > 0:
        ctx.close()
        log(f"ensemble: {args.ensemble_runs} runs")
        members = jc.ensemble(args.ensemble_runs)
        ectx = [capi.Context(c, forcing=f) for c, f in members]
        pts = sum(c.nz * c.nr for c, _ in members)
        batch = capi.Batch(ectx)
        estream = torch.cuda.ExternalStream(ectx[0].stream())
        batch.run(args.warmup)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(estream)
        batch.run(args.ensemble_steps)
        e1.record(estream)
        torch.cuda.synchronize()
        ems = e0.elapsed_time(e1)
        for c in ectx:
            d, _ = c.dt_log()
            if not (d[-1] > 0 and d[-1] < 1.0):
                raise RuntimeError(f"ensemble: implausible time step {d[-1]}")
        ensemble = {"runs": len(members), "points": pts, "steps": args.ensemble_steps,
                    "value": pts * args.ensemble_steps / (ems / 1e3), "ms_per_step": ems / args.ensemble_steps,
                    "workload": "plain FP64 jets, grids uniform over {440..2200}x{64..320} step 8, Mach 1.2-1.8, "
                                "inflow excitation St 0.1-1.0, amplitude 0.1-5% Uj, noise-free initial profiles; "
                                "one launch per kernel per RK stage, HIP-graph replay; 400 steps, not 2000: the "
                                "runs diverge after ~600 steps under the reference's numerics (no filter)"}
        log(f"ensemble: {ems / args.ensemble_steps:.4f} ms/step, {ensemble['value']:.4e} updates/s")
        batch.close()
        for c in ectx:
            c.close()

    combined = None
    if args.combined_runs > 0 and args.combined_steps > 0 and os.path.exists(args.weights):
        import numpy as np
        from axisymmetricjet_amd import acoustics as ac
        from axisymmetricjet_amd import learned_lad as ll
        ctx.close()  # the main context's buffers are not needed any more
        model = ll.load(args.weights)
        specs = jc.combined_specs(args.combined_runs)
        pts = ac.arc_points(case.z[:, 0], case.r[0, :])
        c = jc.LAD_C

        def combined_batch(learned):
            runs = []
            for mach, st, amp, seed in specs:
                cs = jc.make_jet(nz, nr, seed=seed, mach=mach)
                r = capi.Context(cs, forcing=jc.excitation(mach, strouhal=st, amp=amp), shock_capturing=1,
                                 C_mu=c[0], C_beta=c[1], C_kappa=c[2])
                if learned:
                    r.set_learned_lad(model, args.sensor_threshold, args.ramp_min, args.learned_refresh)
                r.set_probes(pts)
                runs.append(r)
                del cs
            return runs, capi.Batch(runs)

        def timed_batch(runs, b, steps):
            bstream = torch.cuda.ExternalStream(runs[0].stream())
            b.run(args.warmup)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(bstream)
            b.run(steps)
            e1.record(bstream)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1)

        def probe_history(runs, b, steps, every=8):
            ps = []
            for _ in range(steps // every):
                b.run(every)
                vs = [r.probe() for r in runs]
                ps.append([v[:, 0] * v[:, 4] for v in vs])
            return np.array(ps)  # (samples, runs, probes)

        cpts = nz * nr * len(specs)
        combined = {"runs": len(specs), "points": cpts, "steps": args.combined_steps,
                    "workload": f"{len(specs)} jets of {nz}x{nr}, Mach 1.4-1.6, inflow excitation St 0.2-0.8, "
                                "amplitude 0.5-3% Uj, seeded noise; each run's coefficients its own launches, "
                                "stage kernels batched over all runs"}
        hist = {}
        for name in ("learned", "lad"):
            t0 = time.time()
            runs, b = combined_batch(name == "learned")
            log(f"combined ({name}): {len(runs)} runs built in {time.time() - t0:.1f} s")
            bms = timed_batch(runs, b, args.combined_steps)
            combined[name] = {"value": cpts * args.combined_steps / (bms / 1e3), "ms_per_step": bms / args.combined_steps}
            for r in runs:
                d, _ = r.dt_log()
                if not (d[-1] > 0 and d[-1] < 1.0):
                    raise RuntimeError(f"combined: implausible time step {d[-1]}")
            if args.combined_score_steps > 0:
                hist[name] = probe_history(runs, b, args.combined_score_steps)
            log(f"combined ({name}): {combined[name]['ms_per_step']:.3f} ms/step, {combined[name]['value']:.4e} updates/s")
            b.close()
            for r in runs:
                r.close()
        if args.combined_long_steps > 0:
            # the configured run length, once: both batches with probes every 8 steps
            N, every = args.combined_long_steps, 8
            long = {"steps": N, "probe_every": every}
            hist_l, bad = {}, {}
            for name in ("learned", "lad"):
                t0 = time.time()
                runs, b = combined_batch(name == "learned")
                bstream = torch.cuda.ExternalStream(runs[0].stream())
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(bstream)
                ps, first_bad = [], [None] * len(runs)
                for k in range(N // every):
                    b.run(every)
                    vs = [r.probe() for r in runs]
                    row = [v[:, 0] * v[:, 4] for v in vs]
                    for ri, p in enumerate(row):
                        if first_bad[ri] is None and not np.all(np.isfinite(p)):
                            first_bad[ri] = (k + 1) * every
                    ps.append(row)
                    if (k + 1) * every % 1000 == 0:
                        log(f"combined long ({name}): step {(k + 1) * every}/{N}, "
                            f"{sum(x is not None for x in first_bad)} runs non-finite, {time.time() - t0:.0f} s")
                e1.record(bstream)
                torch.cuda.synchronize()
                lms = e0.elapsed_time(e1)
                long[name] = {"ms_per_step": lms / N, "value": cpts * N / (lms / 1e3),
                              "runs_nonfinite": sum(x is not None for x in first_bad),
                              "first_nonfinite_step": min((x for x in first_bad if x is not None), default=None)}
                hist_l[name], bad[name] = np.array(ps), first_bad
                log(f"combined long ({name}): {lms / N:.3f} ms/step with probes, "
                    f"{long[name]['runs_nonfinite']} of {len(runs)} runs left the finite range "
                    f"(first at step {long[name]['first_nonfinite_step']})")
                b.close()
                for r in runs:
                    r.close()
            # OASPL per run over the samples both of its runs keep finite
            errs = []
            for ri in range(len(specs)):
                m = min(x // every - 1 if x is not None else N // every for x in (bad["learned"][ri], bad["lad"][ri]))
                if m >= 16:
                    errs.append(np.abs(ac.oaspl(hist_l["learned"][:m, ri, :]) - ac.oaspl(hist_l["lad"][:m, ri, :])))
            long["oaspl_err_db_max"] = float(np.max(errs)) if errs else None
            long["oaspl_err_db_mean"] = float(np.mean(errs)) if errs else None
            long["runs_scored"] = len(errs)
            combined["long"] = long
        combined["speedup_vs_lad"] = combined["lad"]["ms_per_step"] / combined["learned"]["ms_per_step"]
        if hist:
            oa = {k: np.stack([ac.oaspl(v[:, r, :]) for r in range(v.shape[1])]) for k, v in hist.items()}
            d = np.abs(oa["learned"] - oa["lad"])
            combined["score"] = {"steps": args.combined_score_steps, "probe_every": 8, "probes": len(pts),
                                 "oaspl_err_db_max": float(d.max()), "oaspl_err_db_mean": float(d.mean()),
                                 "finite": bool(np.isfinite(hist["learned"]).all() and np.isfinite(hist["lad"]).all())}

    cpu = None
    if not args.no_cpu:
        # every core this job may use: a machine shared by several GPU jobs gives each a share of its cores
        # (OMP_NUM_THREADS), whatever os.cpu_count() shows there
        host_cpus = os.cpu_count() or 1
        threads = min(host_cpus, int(os.environ.get("OMP_NUM_THREADS") or host_cpus))
        log(f"cpu baseline: oracle, {threads} threads of {host_cpus} ({cpu_model()}), ~{args.cpu_seconds:.0f} s")
        v, n, el = cpu_baseline(case, args.cpu_seconds, threads, jc.LAD_C if args.method != "plain" else None)
        cpu = {"value": v, "unit": "grid-point updates/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(), "host_cpus": host_cpus,
               "sample": f"{n} RK4 steps ({'lad' if args.method == 'learned' else args.method}) of the same {nz}x{nr} jet (after 1 warm-up step), "
                         f"{el:.1f} s, C restatement at -O2 without FMA, OpenMP over grid lines, {threads} threads "
                         f"(the job's share of the {host_cpus} host CPUs)"}
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
        "config": {"workload": f"FP64 axisymmetric jet RK4 step, {nz}x{nr} (z x r), Mach 1.5, Re 1e4, Pr 0.7, "
                               f"CFL 0.5, SAT/NSCBC/axis boundaries, "
                               + ("shock capturing and filtering on (LAD C = %g, %g, %g)" % jc.LAD_C
                                  if args.method != "plain" else "shock capturing off")
                               + (", learned LAD coefficient" if args.method == "learned" else ""),
                   "nz": nz, "nr": nr, "points": npts, "method": args.method},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": (measured_traffic(args.method) * npts
                                 if measured_traffic(args.method) and (nz, nr) == (4400, 640) else None),
                     "kernel": "the four RK stages of a step, each the artificial-diffusivity launches (lad) + the stage kernel",
                     "stage_ms": stage_ms, "bytes_per_update": BYTES_PER_UPDATE},
        "other_methods": others,
        "learned": learned,
        "ensemble": ensemble,
        "combined": combined,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
