"""GPU parity of the localized artificial diffusivity (ComputeArtTransCoeffs and the RHS and
trajectories that use it), through the C ABI, against the reference's fixtures and the oracle.

The coefficients mu*, beta*, kappa* are compared normwise per field (max|gpu - ref| /
max|ref|); the GPU keeps the reference's operation order, and its only departure -- the in-place
z filter run as overlapping segments, each started 48 points early -- is damped below 1.3e-16
before any point it writes, so the bar is the RHS one, 1e-12.
"""
import numpy as np
import pytest

from axisymmetricjet_amd import capi
from axisymmetricjet_amd import jetcase as jc
from oracle import pyoracle as po
import cases as gc

pytestmark = pytest.mark.gpu

COEF_TOL = 1e-12
RHS_TOL = 1e-12
STATE_TOL = 1e-9


def lad_context(case, c=gc.LAD_C):
    return capi.Context(case, shock_capturing=1, C_mu=c[0], C_beta=c[1], C_kappa=c[2])


def lad_oracle(case, c=gc.LAD_C):
    return po.Oracle(case, shock=1, c_mu=c[0], c_beta=c[1], c_kappa=c[2])


def normwise(a, b):
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    return num / den if den > 0 else num


def normwise5(a, b):
    return max(normwise(a[:, :, e], b[:, :, e]) for e in range(5))


def compressed(case, amp=0.08):
    """The case's initial state with a Gaussian dip in Vz (a compression ahead of it) so that the
    dilatation switch and beta* are active."""
    q = np.array(case.q0, order="F")
    z = case.z
    r = case.r
    q[:, :, 3] -= amp * np.exp(-((z - 15.0) ** 2) / 2.0) * np.exp(-(r ** 2) / 0.5)
    return q


CASES = {
    "jet440": lambda: jc.make_jet(440, 64, seed=1234),
    "swirl64": lambda: jc.make_jet(64, 40, seed=11, swirl=1e-3, r_max=16.0),
    # lip closure rows at i = 150..153, inside the 48-point warm-up of the z-filter segment that
    # starts at i = 192 (segments of 64 columns on a grid this small)
    "nozzle320": lambda: jc.make_nozzle(320, 96, i_lip=150, j_in=30, j_out=40, seed=4),
    "walls200": lambda: jc.make_jet(200, 64, seed=5, bc_lines=gc.WALL_BC),
    # the full-size grid, whose z filter runs in 128-column segments on MI355X (1008 blocks, one
    # per resident slot); the grids above take 64
    "jet4400": lambda: jc.make_jet(4400, 640, seed=1234),
}


@pytest.mark.parametrize("name", list(CASES))
def test_lad_coeffs_vs_oracle(name):
    case = CASES[name]()
    q = compressed(case)
    print(f"lad coeffs {name} {case.nz}x{case.nr}", flush=True)
    ref = lad_oracle(case).lad(q)
    got = lad_context(case).lad_coeffs(q)
    errs = [normwise(g, r) for g, r in zip(got, ref)]
    print(f"{name}: mu*, beta*, kappa* normwise err {errs}", flush=True)
    assert all(np.abs(r).max() > 0 for r in ref)
    assert max(errs) <= COEF_TOL


@pytest.mark.parametrize("name", ["jet440_lad", "swirl64"])
def test_lad_coeffs_vs_reference_filtered_quantities(name):
    """The reference's own r-filtered quantities (its C is 0, so they are dumped before the
    scaling), scaled here by C Re (C_kappa Re Pr) and the ramp as ComputeArtTransCoeffs does."""
    case, g, shock, nsteps, extra = gc.load(name)
    ctx = lad_context(case)
    got = ctx.lad_coeffs(case.q0)
    re, pr = case.re, case.pr
    scale = [gc.LAD_C[0] * re, gc.LAD_C[1] * re, gc.LAD_C[2] * re * pr]
    ramp = (1.0 + np.tanh(5.0 * (case.z - 10.0))) / 2.0
    for f in range(3):
        want = scale[f] * g["lad_rfiltered"][:, :, f] * ramp
        err = normwise(got[f], want)
        print(f"{name} field {f}: normwise err {err:.3e}", flush=True)
        assert err <= COEF_TOL


def test_lad_rhs_vs_reference():
    """The reference's ComputeRHS with the artificial coefficients injected (swirl64)."""
    case, g, shock, nsteps, extra = gc.load("swirl64")
    ctx = lad_context(case)
    got = ctx.rhs(case.q0)
    err = normwise5(got, g["ladrhs"])
    print(f"swirl64 LAD rhs normwise err {err:.3e}", flush=True)
    assert err <= RHS_TOL


@pytest.mark.parametrize("name", ["nozzle320", "walls200"])
def test_lad_rhs_vs_oracle(name):
    case = CASES[name]()
    q = compressed(case)
    ref = lad_oracle(case).rhs(q)
    got = lad_context(case).rhs(q)
    err = normwise5(got, ref)
    print(f"{name}: LAD rhs normwise err {err:.3e}", flush=True)
    assert err <= RHS_TOL


@pytest.mark.parametrize("name,nsteps", [("jet440", 100), ("nozzle320", 50), ("swirl64", 50)])
def test_lad_trajectory_vs_oracle(name, nsteps):
    case = CASES[name]()
    case.q0[:] = compressed(case)
    print(f"lad trajectory {name} {nsteps} steps", flush=True)
    o = lad_oracle(case)
    dts_ref = o.run(nsteps)
    ctx = lad_context(case)
    ctx.run(nsteps)
    dts, t = ctx.dt_log()
    np.testing.assert_allclose(dts, dts_ref, rtol=1e-12, atol=0)
    assert abs(t - o.time()) <= 1e-12 * abs(o.time())
    err = np.abs(ctx.state() - o.state()).max()
    print(f"{name}: state max-norm err after {nsteps} LAD steps {err:.3e}", flush=True)
    assert err <= STATE_TOL


def test_lad_with_zero_constants_is_the_plain_shock_capturing_step():
    """C = 0 makes mu* = beta* = kappa* = 0 exactly (LAD runs, contributes nothing)."""
    case = jc.make_jet(200, 48, seed=8)
    ctx = lad_context(case, c=(0.0, 0.0, 0.0))
    for f in ctx.lad_coeffs(compressed(case)):
        assert not np.any(f)


def test_lad_coeffs_need_shock_capturing():
    case = jc.make_jet(64, 32, seed=2)
    ctx = capi.Context(case)
    with pytest.raises(capi.AjetError):
        ctx.lad_coeffs(case.q0)
