"""CPU tests of the boundary: the C ABI exports, the patch builder behind it, and the Fortran
host's PLOT3D reader/writer (byte-identical to the Write_Soln layout)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from axisymmetricjet_amd import capi
from axisymmetricjet_amd import jetcase as jc
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_BUILD = os.path.join(ROOT, "axisymmetricjet_amd", "_build", "host")


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ajet_capi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*|void\*)\s+\**(ajet_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 15
    L = C.CDLL(capi.LIB)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(capi.EXPORTS)


def test_params_default_are_reference_constants():
    p = capi.default_params(100, 50)
    assert (p.nz, p.nr, p.viscous, p.shock_capturing, p.forcing) == (100, 50, 1, 0, 0)
    assert (p.gam, p.Re, p.Pr, p.CFL) == (1.4, 1000.0, 0.72, 0.5)
    assert (p.C_mu, p.C_beta, p.C_kappa) == (0.0, 0.0, 0.0)
    assert (p.sponge_start_top, p.sponge_A_top, p.sponge_end_left, p.sponge_A_left) == (12.5, 0.5, 6.0, 5.0)


@pytest.mark.parametrize("case", [
    jc.make_jet(88, 32, seed=1),
    jc.make_nozzle(96, 48, i_lip=24, j_in=16, j_out=22),
    jc.make_jet(440, 64, seed=1, bc_lines=jc.NOZZLE_BC_4400x640[:0] + jc.DEFAULT_BC),
])
def test_patches_match_oracle(case):
    got = capi.build_patches(case.nz, case.nr, case.bc_lines)
    assert got == po.Oracle(case).patches()


def test_nozzle_bc_file_patches_at_full_size():
    """The one bc file the reference ships, on its 4400 x 640 grid."""
    got = capi.build_patches(4400, 640, jc.NOZZLE_BC_4400x640)
    assert got[0] == (1, 4, 1, 150, 91, 1)
    assert got[1] == (1, 597, 147, 150, 94, -2)
    assert got[2] == (597, 600, 150, 169, 94, 1)
    assert got[3] == (1, 597, 169, 172, 94, 2)
    assert got[5] == (1, 4400, 637, 640, 92, -2)
    assert got[6] == (4397, 4400, 1, 640, 92, -1)
    assert got[7] == (1, 4400, 1, 4, 90, 2)


def ensure_host_built():
    exe = os.path.join(HOST_BUILD, "p3d_roundtrip")
    if not os.path.exists(exe):
        subprocess.check_call(["bash", os.path.join(ROOT, "axisymmetricjet_amd", "host", "build_host.sh")])
    return exe


def test_fortran_plot3d_roundtrip_is_byte_identical(tmp_path):
    exe = ensure_host_built()
    case = jc.make_nozzle(96, 48, i_lip=24, j_in=16, j_out=22)
    grid = str(tmp_path / "g.xyz")
    sin = str(tmp_path / "in.q")
    sout = str(tmp_path / "out.q")
    jc.write_grid(grid, case.z, case.r, case.iblank)
    jc.write_soln(sin, case.q0, tau=(12.0, 0.1, 0.2, 3.5))
    out = subprocess.check_output([exe, grid, sin, sout], text=True).split()
    nz, nr, ibsum = int(out[0]), int(out[1]), int(out[2])
    assert (nz, nr, ibsum) == (96, 48, int(case.iblank.sum()))
    np.testing.assert_allclose(float(out[5]), case.q0.sum(), rtol=1e-14)
    assert open(sin, "rb").read() == open(sout, "rb").read()
    q, tau = jc.read_soln(sout)
    assert tau == (12.0, 0.1, 0.2, 3.5)
    assert np.array_equal(q, case.q0)


def test_tuned_tile_is_a_candidate_and_reaches_the_build():
    import json
    from axisymmetricjet_amd import autotune, build
    with open(autotune.TUNE) as f:
        doc = json.load(f)
    assert tuple(doc["tile"]) in autotune.CANDIDATES
    assert doc["states_agree_within"] <= autotune.STATE_TOL
    tz, tr, th = doc["tile"]
    assert build.tile_defines() == [f"-DAJET_TZ={tz}", f"-DAJET_TR={tr}", f"-DAJET_THREADS={th}"]
    assert build.tile_defines((64, 8, 256)) == ["-DAJET_TZ=64", "-DAJET_TR=8", "-DAJET_THREADS=256"]
    # every timed candidate is listed, and the choice is the fastest on the stated metric
    t = doc["timings_ms_per_step"]
    assert min(t, key=lambda k: t[k]["lad"]) == "x".join(str(x) for x in doc["tile"])


def test_mixed_tile_list(tmp_path):
    """The launch list that puts interior and boundary tiles in one launch (host C++, built here
    with g++): each tile once, boundary tiles marked, a boundary tile first (it writes the step
    scalars), each XCD's range starting with its share of the boundary tiles."""
    exe = str(tmp_path / "tile_lists_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-ffunction-sections", "-fdata-sections", "-Wl,--gc-sections",
                           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "tile_lists_check.cpp"),
                           os.path.join(ROOT, "axisymmetricjet_amd", "csrc", "ajet_setup.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


def _case_file(path, case):
    lines = np.asarray(case.bc_lines, dtype=np.int64).reshape(-1, 7)
    with open(path, "wb") as f:
        f.write(np.array([case.nz, case.nr, len(lines)], dtype=np.int32).tobytes())
        f.write(np.asarray(case.z, dtype=np.float64).tobytes(order="F"))
        f.write(np.asarray(case.r, dtype=np.float64).tobytes(order="F"))
        f.write(np.asarray(case.iblank, dtype=np.int32).tobytes(order="F"))
        f.write(lines.tobytes())


@pytest.mark.parametrize("name", ["jet440", "nozzle96", "jet4400", "nozzle4400"])
def test_lad_source_tiles(tmp_path, name):
    """Which LAD source tiles compute the strain rate and dilatation themselves (host C++, built
    here with g++), checked point by point against the metadata: fused tiles read only interior
    points off the axis, inside the grid; every point a non-fused tile reads is in a strain
    tile.  On the plain jet every interior tile is fused; the nozzle's lip leaves some that are
    not."""
    case = {"jet440": lambda: jc.make_jet(440, 64, seed=1),
            "nozzle96": lambda: jc.make_nozzle(96, 48, i_lip=24, j_in=16, j_out=22),
            "jet4400": lambda: jc.make_jet(4400, 640, seed=1),
            "nozzle4400": lambda: jc.make_nozzle(4400, 640, i_lip=597, j_in=150, j_out=169)}[name]()
    exe, data = str(tmp_path / "lad_tiles_check"), str(tmp_path / "case.bin")
    _case_file(data, case)
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "lad_tiles_check.cpp"),
                           os.path.join(ROOT, "axisymmetricjet_amd", "csrc", "ajet_setup.cpp"), "-o", exe])
    out = subprocess.run([exe, data], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout
    fused, other, strain, total = (int(x) for x in re.findall(r"\d+", out.stdout))
    assert fused + other == total and fused > 0
    if name == "jet4400":
        # the strain kernel keeps the tiles within reach of the edges' boundary tiles
        assert strain < 0.3 * total and other < 0.2 * total


def test_interior_fast_math(tmp_path):
    """The interior tiles' FP64 helpers (csrc/ajet_fastmath.h), built for the host with hipcc:
    T**0.666 within 1e-15 of the C library's pow over its whole fast range (2^-8, 2^8) and the
    library's pow itself outside it; the reciprocal within an ulp; the factored D1 and D2
    interior stencils against the reference's unfused sums and exact on polynomials."""
    exe = str(tmp_path / "fastmath_check")
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                           "-ffp-contract=off", "-I", os.path.join(ROOT, "axisymmetricjet_amd", "csrc"),
                           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "fastmath_check.cpp"),
                           "-o", exe], stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
