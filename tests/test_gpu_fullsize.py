"""Full-size parity: the 4400 x 640 Mach 1.5 jet, 1000 RK4 steps on the GPU, against the
oracle's trajectory of the same seeded case (tests/golden/make_full_size.py), for the plain step
and for the step with shock capturing and filtering on.

The fixture keeps a 16 x 16 subsample of the state, every boundary row and column, the per
variable max |q| after 10 and after 1000 steps, and the 1000 time steps.  Bars: the state within
1e-9 (max norm) after 1000 steps; the time steps within 1e-10 relative (each is CFL over a max
of the state, so it inherits the state's agreement).
"""
import os

import numpy as np
import pytest

from axisymmetricjet_amd import capi
from axisymmetricjet_amd import jetcase as jc
import cases as gc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
STATE_TOL = 1e-9


def edges(q):
    return np.concatenate([q[:, 0, :], q[:, -1, :], q[0, :, :], q[-1, :, :]], axis=0)


@pytest.mark.parametrize("method", ["plain", "lad"])
def test_full_size_1000_steps_vs_oracle(method):
    path = os.path.join(HERE, "golden", f"jet4400_{method}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} not written (tests/golden/make_full_size.py)")
    g = dict(np.load(path))
    case = jc.make_jet(4400, 640, seed=1234)
    assert str(g["input_sha256"]) == gc.input_digest(case)
    c = tuple(float(x) for x in g["lad_c"])
    if method == "lad":
        assert c == jc.LAD_C
        ctx = capi.Context(case, shock_capturing=1, C_mu=c[0], C_beta=c[1], C_kappa=c[2])
    else:
        ctx = capi.Context(case)
    done = 0
    for snap in (10, 1000):
        print(f"{method}: running to step {snap}", flush=True)
        ctx.run(snap - done)
        done = snap
        q = ctx.state()
        e_sub = np.abs(q[::16, ::16, :] - g[f"sub{snap}"]).max()
        e_edge = np.abs(edges(q) - g[f"edges{snap}"]).max()
        amax = np.abs(q).max(axis=(0, 1))
        print(f"{method} step {snap}: subsample err {e_sub:.3e}, boundary err {e_edge:.3e}", flush=True)
        assert np.isfinite(q).all()
        assert e_sub <= STATE_TOL and e_edge <= STATE_TOL
        np.testing.assert_allclose(amax, g[f"absmax{snap}"], rtol=STATE_TOL, atol=STATE_TOL)
    dts, t = ctx.dt_log()
    assert len(dts) == 1000
    np.testing.assert_allclose(dts, g["dts"], rtol=1e-10, atol=0)
    assert abs(t - float(g["time"])) <= 1e-10 * float(g["time"])
