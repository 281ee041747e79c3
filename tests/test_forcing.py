"""Forcing terms in the oracle: ControlForcingTerms (ModControlForcing.f90:6-65) against the
reference's own routine, and the inflow excitation (this project's extension) against its closed
form.  CPU only."""
import numpy as np
import pytest

from axisymmetricjet_amd import jetcase as jc
from oracle import pyoracle as po
import cases as gc


def normwise(a, b):
    out = []
    for e in range(5):
        den = np.abs(b[:, :, e]).max()
        num = np.abs(a[:, :, e] - b[:, :, e]).max()
        out.append(num / den if den > 0 else num)
    return max(out)


def gaussian(case, f):
    return np.exp(-1.0 * ((case.z - f.z0) ** 2 / f.lz ** 2) - (case.r - f.r0) ** 2 / f.lr ** 2)


def test_feedback_forcing_matches_reference_routine():
    """The reference's ControlForcingTerms, called by the harness after ComputeRHS, against the
    oracle's forcing (applied before the BCs as ComputeRHS does): the forcing increments agree."""
    case, g, shock, nsteps, extra = gc.load("forcing88")
    f = gc.feedback_forcing()
    assert np.array_equal(g["forcing_F"], f.F)
    ref_delta = g["forcingrhs"] - g["rhs"]
    plain = po.Oracle(case).rhs(case.q0)
    forced = po.Oracle(case, forcing=f).rhs(case.q0)
    assert np.array_equal(plain, g["rhs"])
    got_delta = forced - plain
    # the increments differ only by where they are rounded in (before / after the BCs)
    assert normwise(got_delta, ref_delta) <= 1e-12
    assert np.abs(ref_delta).max() > 1e-6


def test_feedback_forcing_closed_form():
    case = jc.make_jet(88, 32, seed=3)
    f = gc.feedback_forcing()
    q = np.array(case.q0, order="F")
    q[:, :, 1] += 1e-2
    q[:, :, 4] += 2e-2
    delta = po.Oracle(case, forcing=f).rhs(q) - po.Oracle(case).rhs(q)
    G = gaussian(case, f)
    want = np.zeros_like(delta)
    for e in range(5):
        for k in range(5):
            want[:, :, e] = want[:, :, e] + 6.0 * f.F[e, k] * G * (q[:, :, k] - case.target[:, :, k])
    assert normwise(delta, want) <= 1e-12
    # the axis row takes the forcing too (the forcing loops run over every point)
    assert np.abs(delta[:, 0, 1]).max() > 0


def test_excitation_closed_form_and_zero_amplitude():
    case = jc.make_jet(96, 40, seed=4)
    f = jc.excitation(1.5, strouhal=0.45, amp=0.02, z0=2.0, r0=0.5, lz=1.0, lr=0.2)
    f.t0 = 1.25
    plain = po.Oracle(case).rhs(case.q0)
    delta = po.Oracle(case, forcing=f).rhs(case.q0) - plain
    G = gaussian(case, f)
    want = f.amp * f.uj * np.sin(2.0 * np.pi * f.strouhal * f.uj * (f.t0 + 0.0)) * G
    assert np.abs(delta[:, :, [0, 1, 2, 4]]).max() == 0.0
    np.testing.assert_allclose(delta[:, :, 3], want, rtol=1e-9, atol=1e-14)
    f0 = jc.excitation(1.5, strouhal=0.45, amp=0.0)
    assert np.array_equal(po.Oracle(case, forcing=f0).rhs(case.q0), plain)


def test_excitation_uses_the_stage_times():
    """One RK4 step of a state at rest under pure excitation: the Vz increment is the RK4
    quadrature of amp*Uj*sin(w t)*G over the step (stage times t, t+dt/2, t+dt/2, t+dt)."""
    case = jc.make_jet(64, 32, seed=1)
    case.q0[:, :, 1:4] = 0.0
    case.q0[:, :, 0] = 1.0
    case.q0[:, :, 4] = 1.0
    case.target[:] = case.q0
    f = jc.excitation(1.5, strouhal=0.3, amp=0.05, z0=10.0, r0=4.0, lz=2.0, lr=1.0)
    o = po.Oracle(case, forcing=f, viscous=0)
    dt = 1e-3
    o.compute_dt()
    lib = po.lib()
    lib.orc_set_dt(o.h, dt)
    lib.orc_rk4(o.h)
    q = o.state()
    G = gaussian(case, f)
    w = 2.0 * np.pi * f.strouhal * f.uj
    s = lambda t: f.amp * f.uj * np.sin(w * t)  # noqa: E731
    want = dt * (s(0.0) + 2 * s(dt / 2) + 2 * s(dt / 2) + s(dt)) / 6.0 * G
    interior = (slice(8, -8), slice(8, -8))
    # within one step the induced velocity is advected only to second order in dt
    np.testing.assert_allclose(q[:, :, 3][interior], want[interior], rtol=1e-6, atol=1e-6 * np.abs(want).max())
    # a wrong stage time (all four stages at the step start) is off by far more
    wrong = dt * s(0.0) * G
    assert np.abs(wrong - want).max() > 100 * 1e-6 * np.abs(want).max()
