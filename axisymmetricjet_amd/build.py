"""Build the gfx950 library libajet.so in-tree with hipcc (no JIT cache, no pip install).

    python -m axisymmetricjet_amd.build
"""
from __future__ import annotations

import json
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libajet.so")
SOURCES = ["ajet_kernels.hip", "ajet_lad.hip", "ajet_mlp.hip", "ajet_capi.cpp", "ajet_setup.cpp", "ajet_trainer.cpp",
           "ajet_noalloc.cpp"]
HEADERS = ["ajet_internal.h", "ajet_kernels.h", "ajet_physics.h", "ajet_stencil.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: every multiply and add rounds on its own, as in the reference's
# unoptimised build; packed FP32 is disabled on the device side.
# -fmacf64-inst: FP64 multiply-adds take the three-operand form, whose addend may sit in scalar
# registers, rather than the accumulate form, which first moves a constant addend into a
# vector register pair (two extra vector instructions per constant; same arithmetic).
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Xclang", "-target-feature", "-Xclang", "-fmacf64-inst",
         "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
# Per-source additions.  The stage kernels are latency-bound at 2-3 waves per SIMD; the
# scheduler that clusters their memory instructions measured 1.375 -> 1.341 ms per LAD step
# (plain step within 0.5 %, results bit-identical; the ILP-first strategies were slower).
SOURCE_FLAGS = {"ajet_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=max-memory-clause"]}


TUNE = os.path.join(HERE, "tune.json")


def tile_defines(tile=None):
    """-D flags of the stage/LAD tile (z points, r points, threads of the stage kernel): the
    given tile, else the one the autotuner recorded in tune.json, else the sources' 32 x 16 x 256."""
    if tile is None and os.path.exists(TUNE):
        with open(TUNE) as f:
            tile = json.load(f).get("tile")
    if not tile:
        return []
    tz, tr, th = (int(x) for x in tile)
    return [f"-DAJET_TZ={tz}", f"-DAJET_TR={tr}", f"-DAJET_THREADS={th}"]


def build(verbose: bool = False, tile=None, out: str | None = None) -> str:
    """Compile every source (in parallel) and link libajet.so (or `out`); raises on any compiler
    error.  tile = (z points, r points, threads) overrides the tuned tile."""
    out = out or LIB
    objdir = os.path.join(os.path.dirname(out), "_build") if out != LIB else os.path.join(HERE, "_build")
    objs, procs = [], []
    os.makedirs(objdir, exist_ok=True)
    defs = tile_defines(tile)
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC] + FLAGS + defs + SOURCE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(shlex.quote(x) for x in cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
