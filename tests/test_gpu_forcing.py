"""GPU parity of the forcing terms (ControlForcingTerms feedback and the inflow excitation)
through the C ABI, against the oracle; bars as for the plain path (RHS 1e-12 normwise, state
1e-9 after many steps)."""
import numpy as np
import pytest

from axisymmetricjet_amd import capi
from axisymmetricjet_amd import jetcase as jc
from oracle import pyoracle as po
import cases as gc

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
STATE_TOL = 1e-9


def normwise(a, b):
    out = []
    for e in range(5):
        den = np.abs(b[:, :, e]).max()
        num = np.abs(a[:, :, e] - b[:, :, e]).max()
        out.append(num / den if den > 0 else num)
    return max(out)


def both():
    f = jc.excitation(1.5, strouhal=0.45, amp=0.02, z0=2.0, r0=0.5, lz=1.0, lr=0.2,
                      feedback=gc.feedback_forcing().F)
    f.t0 = 0.75
    return f


def test_forcing_rhs_vs_reference_increment():
    """The GPU forcing increment against the reference's own ControlForcingTerms."""
    case, g, shock, nsteps, extra = gc.load("forcing88")
    f = gc.feedback_forcing()
    got = capi.Context(case, forcing=f).rhs(case.q0) - capi.Context(case).rhs(case.q0)
    err = normwise(got, g["forcingrhs"] - g["rhs"])
    print(f"forcing88: increment normwise err {err:.3e}", flush=True)
    assert err <= RHS_TOL


def test_forcing_rhs_vs_oracle():
    case = jc.make_nozzle(160, 96, i_lip=40, j_in=24, j_out=30, seed=6)
    f = both()
    q = np.array(case.q0, order="F")
    q[:, :, 1] += 1e-3
    ref = po.Oracle(case, forcing=f).rhs(q)
    got = capi.Context(case, forcing=f).rhs(q)
    err = normwise(got, ref)
    print(f"nozzle160 forced rhs normwise err {err:.3e}", flush=True)
    assert err <= RHS_TOL


@pytest.mark.parametrize("shock", [0, 1])
def test_forced_trajectory_vs_oracle(shock):
    case = jc.make_jet(440, 64, seed=1234)
    f = both()
    c = gc.LAD_C if shock else (0.0, 0.0, 0.0)
    nsteps = 100 if not shock else 40
    print(f"forced jet440 shock={shock} {nsteps} steps", flush=True)
    o = po.Oracle(case, shock=shock, c_mu=c[0], c_beta=c[1], c_kappa=c[2], forcing=f)
    dts_ref = o.run(nsteps)
    ctx = capi.Context(case, forcing=f, shock_capturing=shock, C_mu=c[0], C_beta=c[1], C_kappa=c[2])
    ctx.run(nsteps)
    dts, t = ctx.dt_log()
    np.testing.assert_allclose(dts, dts_ref, rtol=1e-12, atol=0)
    assert abs(t - o.time()) <= 1e-12 * o.time()
    err = np.abs(ctx.state() - o.state()).max()
    print(f"forced jet440 shock={shock}: state max-norm err {err:.3e}", flush=True)
    assert err <= STATE_TOL


def test_zero_forcing_is_the_plain_step_bit_for_bit():
    case = jc.make_jet(96, 48, seed=9)
    f = jc.excitation(1.5, strouhal=0.5, amp=0.0)
    a = capi.Context(case, forcing=f)
    b = capi.Context(case)
    a.run(20)
    b.run(20)
    assert np.array_equal(a.state(), b.state())
    assert np.array_equal(a.dt_log()[0], b.dt_log()[0])


def test_forcing_parameters_required():
    case = jc.make_jet(64, 32, seed=2)
    p = capi.default_params(case.nz, case.nr, forcing=1, Re=case.re, Pr=case.pr)
    ctx = capi.Context(case, params=p)
    with pytest.raises(capi.AjetError):
        ctx.run(1)
