"""The PLOT3D boundary of the path on the host (no GPU): the solution files the reference's own
WriteOutput writes (ModInputOutput.f90:201-263 -> Write_Soln, ModPLOT3D_IO.f90:539-599), kept as
raw bytes in tests/golden/plot3d88.npz by tests/golden/make_golden.py, against this project's
Fortran writer (axisymmetricjet_amd/host/ajet_plot3d.f90, through the p3d_roundtrip tool) and
the Python one (jetcase.write_soln); and the variants the reference's reader also accepts
(p3d_detect_, plot3d_format.c:16-163: single grid, single precision, planes, 2-D grids), which
this host rejects with a message rather than misreading.  With ND(1,3) = 1 the 'planes' layout
is the same bytes as 'whole', so it is accepted as such.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from axisymmetricjet_amd import jetcase as jc
import cases as gc
from test_host import ensure_host_built

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture():
    g = dict(np.load(os.path.join(HERE, "golden", "plot3d88.npz")))
    case = jc.make_jet(88, 32, seed=7)
    assert str(g["input_sha256"]) == gc.input_digest(case)
    return case, g


def test_reference_writes_the_initial_condition_as_this_host_does(tmp_path):
    case, g = fixture()
    exe = ensure_host_built()
    grid, sin, sout = (str(tmp_path / n) for n in ("g.xyz", "ic.q", "out.q"))
    jc.write_grid(grid, case.z, case.r, case.iblank)
    jc.write_soln(sin, case.q0, tau=tuple(g["tau_in"]))
    ref = g["ic_file"].tobytes()
    # the reference's WriteOutput of the initial condition is the input file itself (tau(1) =
    # iteration = 0, tau(2:4) passed through): the Python writer's bytes ...
    assert open(sin, "rb").read() == ref
    # ... and the Fortran host writer's, after reading it back
    subprocess.check_call([exe, grid, sin, sout], stdout=subprocess.DEVNULL)
    assert open(sout, "rb").read() == ref


def test_reference_restart_file_reads_back_byte_for_byte(tmp_path):
    """The reference's output after 20 steps: the host reads it (tau(1) = iteration 20, tau(4) =
    time, tau(2:3) from the input) and writes the same bytes."""
    case, g = fixture()
    exe = ensure_host_built()
    grid, sin, sout = (str(tmp_path / n) for n in ("g.xyz", "in.q", "out.q"))
    jc.write_grid(grid, case.z, case.r, case.iblank)
    open(sin, "wb").write(g["out1_file"].tobytes())
    q, tau = jc.read_soln(sin)
    assert tau[0] == 20.0 and tau[1:3] == tuple(g["tau_in"][1:3])
    assert abs(tau[3] - float(g["time1"])) == 0.0
    subprocess.check_call([exe, grid, sin, sout], stdout=subprocess.DEVNULL)
    assert open(sout, "rb").read() == g["out1_file"].tobytes()


def rec(payload: bytes) -> bytes:
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def soln_variant(kind, nz=8, nr=6, nk=1):
    """A solution file in one of the variants p3d_detect_ tells apart."""
    q = np.arange(nz * nr * nk * 5, dtype=np.float64)
    tau = struct.pack("<4d", 0.0, 0.0, 0.0, 0.0)
    if kind == "single-grid":   # no ngrid record
        return rec(struct.pack("<3i", nz, nr, nk)) + rec(tau) + rec(q.tobytes())
    if kind == "single-precision":
        return (rec(struct.pack("<i", 1)) + rec(struct.pack("<3i", nz, nr, nk)) +
                rec(struct.pack("<4f", 0, 0, 0, 0)) + rec(q.astype(np.float32).tobytes()))
    if kind == "planes":        # one record per k-plane
        body = b"".join(rec(q.reshape(5, nk, nr * nz)[:, k, :].tobytes()) for k in range(nk))
        return rec(struct.pack("<i", 1)) + rec(struct.pack("<3i", nz, nr, nk)) + rec(tau) + body
    if kind == "2d":            # NDIM = 2: two sizes, four variables
        return (rec(struct.pack("<i", 1)) + rec(struct.pack("<2i", nz, nr)) + rec(tau) +
                rec(np.zeros(nz * nr * 4).tobytes()))
    if kind == "truncated":
        good = rec(struct.pack("<i", 1)) + rec(struct.pack("<3i", nz, nr, 1)) + rec(tau) + rec(q.tobytes())
        return good[:-100]
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["single-grid", "single-precision", "planes", "2d", "truncated"])
def test_other_plot3d_variants_are_rejected(tmp_path, kind):
    exe = ensure_host_built()
    case = jc.make_jet(8, 6, seed=1)
    grid, sin, sout = (str(tmp_path / n) for n in ("g.xyz", "in.q", "out.q"))
    jc.write_grid(grid, case.z, case.r, case.iblank)
    nk = 2 if kind == "planes" else 1   # planes with one plane is the whole layout
    open(sin, "wb").write(soln_variant(kind, nk=nk))
    out = subprocess.run([exe, grid, sin, sout], capture_output=True, text=True)
    assert out.returncode != 0
    assert "PLOT3D:" in out.stdout or "PLOT3D:" in out.stderr
    assert not os.path.exists(sout)


def test_planes_layout_with_one_plane_is_the_whole_layout(tmp_path):
    exe = ensure_host_built()
    case = jc.make_jet(8, 6, seed=1)
    grid, sin, sout = (str(tmp_path / n) for n in ("g.xyz", "in.q", "out.q"))
    jc.write_grid(grid, case.z, case.r, case.iblank)
    data = soln_variant("planes", nk=1)
    open(sin, "wb").write(data)
    subprocess.check_call([exe, grid, sin, sout], stdout=subprocess.DEVNULL)
    assert open(sout, "rb").read() == data


def test_grid_variants_are_rejected(tmp_path):
    """A 2-D grid (NDIM = 2, two coordinates) and a single-precision grid."""
    exe = ensure_host_built()
    nz, nr = 8, 6
    sin, sout = str(tmp_path / "in.q"), str(tmp_path / "out.q")
    jc.write_soln(sin, np.zeros((nz, nr, 5)))
    for name, payload in (("2d", rec(struct.pack("<i", 1)) + rec(struct.pack("<2i", nz, nr)) +
                           rec(np.zeros(nz * nr * 2).tobytes() + np.ones(nz * nr, np.int32).tobytes())),
                          ("sp", rec(struct.pack("<i", 1)) + rec(struct.pack("<3i", nz, nr, 1)) +
                           rec(np.zeros(nz * nr * 3, np.float32).tobytes() + np.ones(nz * nr, np.int32).tobytes()))):
        grid = str(tmp_path / f"{name}.xyz")
        open(grid, "wb").write(payload)
        out = subprocess.run([exe, grid, sin, sout], capture_output=True, text=True)
        assert out.returncode != 0 and "PLOT3D:" in (out.stdout + out.stderr), name
