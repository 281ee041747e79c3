"""Tile-shape autotuner of the stage and LAD kernels.

The fused stage kernel and the LAD strain and source kernels work on TZ x TR tiles (with a
2-point and a 7-point halo) of STAGE_THREADS threads.  The shape trades halo re-reads against
LDS per block, waves per SIMD and launch granularity, and it is a compile-time constant. So the
tuner does four things:
* builds one library per candidate (on the CPU, where hipcc runs);
* times each on the target workload on the GPU, in its own process (AJET_LIB);
* checks that each candidate's 10-step state of the 440 x 64 LAD jet is bit-identical to the
  others' (the shape changes no arithmetic);
* records the fastest in tune.json, which axisymmetricjet_amd.build then compiles in.

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
digest = hashlib.sha256(np.ascontiguousarray(small.state()).tobytes()).hexdigest()
small.close()
ctx = capi.Context(jc.make_jet(nz, nr, seed=1234), **(lad if method != "plain" else {}))
if method == "learned":
    from axisymmetricjet_amd import learned_lad as ll
    ctx.set_learned_lad(ll.load(), 0.1, 1e-8, 4)
ctx.run(10)
st = torch.cuda.ExternalStream(ctx.stream())
torch.cuda.synchronize()
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record(st)
ctx.run(steps)
e1.record(st)
torch.cuda.synchronize()
dts, _ = ctx.dt_log()
print(json.dumps({"ms_per_step": e0.elapsed_time(e1) / steps, "digest": digest, "dt_last": float(dts[-1])}))
"""


def measure(tile, method, nz, nr, steps):
    env = dict(os.environ, AJET_LIB=lib_path(tile), PYTHONPATH=ROOT)
    out = subprocess.run([sys.executable, "-c", WORKER, method, str(nz), str(nr), str(steps)], env=env,
                         capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"autotune: {tile} failed:\n{out.stderr[-2000:]}")
    return json.loads(out.stdout.strip().splitlines()[-1])


def time_all(cands=CANDIDATES, nz=4400, nr=640, steps=200, methods=("lad", "learned"), log=print):
    results = {}
    for t in cands:
        if not os.path.exists(lib_path(t)):
            log(f"autotune: {t} not built, skipped")
            continue
        key = "x".join(str(x) for x in t)
        results[key] = {}
        for m in methods:
            r = measure(t, m, nz, nr, steps)
            results[key][m] = r["ms_per_step"]
            results[key]["digest"] = r["digest"]
            log(f"autotune: {t} {m}: {r['ms_per_step']:.4f} ms/step")
    digests = {v["digest"] for v in results.values()}
    if len(digests) != 1:
        raise RuntimeError(f"autotune: the candidates disagree on the 440x64 LAD state: {results}")
    best = min(results, key=lambda k: results[k][methods[0]])
    tile = [int(x) for x in best.split("x")]
    doc = {"tile": tile, "chosen_by": f"{methods[0]} ms/step at {nz}x{nr}, {steps} steps",
           "timings_ms_per_step": {k: {m: v[m] for m in methods} for k, v in results.items()},
           "identical_results": True}
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
