"""Tile-shape autotuner of the stage and LAD kernels.

The fused stage kernel and the LAD strain and source kernels work on TZ x TR tiles (with a
2-point and a 7-point halo) of STAGE_THREADS threads.  The shape trades halo re-reads against
LDS per block, waves per SIMD and launch granularity, and it is a compile-time constant. So the
tuner does four things:
* builds one library per candidate (on the CPU, where hipcc runs);
* times each on the target workload on the GPU, in its own process (AJET_LIB);
* checks that the candidates' 10-step states of the 440 x 64 LAD jet agree within 1e-12: the
  shape decides which points fall in interior tiles (the fused arithmetic) and which in
  boundary tiles (the reference's operation order), so the last bits may differ;
* records the fastest on the headline workload (the LAD step at 4400 x 640) in tune.json,
  which axisymmetricjet_amd.build then compiles in, with every candidate's timings on the plain
  and learned steps and on the batched ensemble (jetcase.ensemble(64): 17.3 M points on grids
  of 440..2200 x 64..320) and the best shape of each workload.

    python -m axisymmetricjet_amd.autotune build     # CPU: candidate libraries in _tune/
    python -m axisymmetricjet_amd.autotune time      # GPU: timings, tune.json
    python -m axisymmetricjet_amd.build              # the tuned library
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TUNE = os.path.join(HERE, "tune.json")
# (z points, r points, threads): LDS per stage-kernel block with LAD is 9 fields x window + the
# two staged z-derivatives, e.g. 62 KB at 32 x 16 and 111 KB at 32 x 32
CANDIDATES = [(32, 16, 256), (64, 8, 256), (16, 32, 256), (64, 16, 512), (32, 32, 512)]


def lib_path(tile):
    return os.path.join(HERE, "_tune", "x".join(str(x) for x in tile), "libajet.so")


def build_all(cands=CANDIDATES, log=print):
    from . import build as hb
    for t in cands:
        log(f"autotune: building {t}")
        hb.build(tile=t, out=lib_path(t))


WORKER = r"""
import hashlib, json, sys
import numpy as np, torch
from axisymmetricjet_amd import capi, jetcase as jc
method, nz, nr, steps = sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4])
c = jc.LAD_C
lad = dict(shock_capturing=1, C_mu=c[0], C_beta=c[1], C_kappa=c[2])
small = capi.Context(jc.make_jet(440, 64, seed=1234), **lad)
small.run(10)
small_state = small.state()
digest = hashlib.sha256(np.ascontiguousarray(small_state).tobytes()).hexdigest()
small.close()
np.save(sys.argv[5], small_state)
small.close()
if method == "ensemble":
    members = jc.ensemble(64)
    runs = [capi.Context(c, forcing=f) for c, f in members]
    b = capi.Batch(runs)
    b.run(10)
    st = torch.cuda.ExternalStream(runs[0].stream())
    go = lambda: b.run(steps)
    ctx = runs[0]
else:
    ctx = capi.Context(jc.make_jet(nz, nr, seed=1234), **(lad if method != "plain" else {}))
    if method == "learned":
        from axisymmetricjet_amd import learned_lad as ll
        ctx.set_learned_lad(ll.load(), 0.1, 1e-8, 8)
    ctx.run(10)
    st = torch.cuda.ExternalStream(ctx.stream())
    go = lambda: ctx.run(steps)
torch.cuda.synchronize()
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record(st)
go()
e1.record(st)
torch.cuda.synchronize()
dts, _ = ctx.dt_log()
print(json.dumps({"ms_per_step": e0.elapsed_time(e1) / steps, "digest": digest, "dt_last": float(dts[-1])}))
"""


def measure(tile, method, nz, nr, steps):
    import numpy as np
    env = dict(os.environ, AJET_LIB=lib_path(tile), PYTHONPATH=ROOT)
    state = os.path.join(os.path.dirname(lib_path(tile)), f"state_{method}.npy")
    out = subprocess.run([sys.executable, "-c", WORKER, method, str(nz), str(nr), str(steps), state], env=env,
                         capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"autotune: {tile} failed:\n{out.stderr[-2000:]}")
    r = json.loads(out.stdout.strip().splitlines()[-1])
    r["state"] = np.load(state)
    return r


STATE_TOL = 1e-12  # agreement of the candidates' 10-step 440 x 64 LAD states (max norm)


def time_all(cands=CANDIDATES, nz=4400, nr=640, steps=200, methods=("lad", "plain", "learned", "ensemble"),
             ensemble_steps=100, log=print):
    import numpy as np
    results, states = {}, {}
    for t in cands:
        if not os.path.exists(lib_path(t)):
            log(f"autotune: {t} not built, skipped")
            continue
        key = "x".join(str(x) for x in t)
        results[key] = {}
        for m in methods:
            r = measure(t, m, nz, nr, ensemble_steps if m == "ensemble" else steps)
            results[key][m] = r["ms_per_step"]
            states[key] = r["state"]
            log(f"autotune: {t} {m}: {r['ms_per_step']:.4f} ms/step")
    ref = next(iter(states.values()))
    diff = max(float(np.abs(s - ref).max()) for s in states.values())
    if not diff <= STATE_TOL:
        raise RuntimeError(f"autotune: the candidates' 440x64 LAD states differ by {diff:.3e}")
    best = min(results, key=lambda k: results[k][methods[0]])
    tile = [int(x) for x in best.split("x")]
    doc = {"tile": tile, "chosen_by": f"{methods[0]} ms/step at {nz}x{nr}, {steps} steps",
           "timings_ms_per_step": results,
           "best_per_workload": {m: min(results, key=lambda k: results[k][m]) for m in methods},
           "ensemble_steps": ensemble_steps,
           "states_agree_within": diff, "identical_results": diff == 0.0}
    with open(TUNE, "w") as f:
        json.dump(doc, f, indent=1)
    log(f"autotune: best {tile}, written to {TUNE}")
    return doc


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("what", choices=["build", "time"])
    ap.add_argument("--nz", type=int, default=4400)
    ap.add_argument("--nr", type=int, default=640)
    ap.add_argument("--steps", type=int, default=200)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.what == "build":
        build_all(log=log)
    else:
        time_all(nz=a.nz, nr=a.nr, steps=a.steps, log=log)


if __name__ == "__main__":
    main()
