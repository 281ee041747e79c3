"""Full-size parity at 4400 x 640 against the oracle's trajectory of the same seeded case
(tests/golden/make_full_size.py):

* the Mach 1.5 jet, 1000 RK4 steps, with the plain step and with shock capturing and filtering on;
* the same grid with the reference's own bc file, bc_4400x640_nozzle.dat (jetcase.NOZZLE_BC_4400x640:
  SAT inflow, NSCBC no-slip nozzle walls, NSCBC top and outflow, axis; the points inside the lip
  blanked), 100 steps with and without shock capturing.

Each fixture keeps, after 10 steps and at the end, a 16 x 16 subsample of the state, every boundary
row and column, the per-variable max |q|, and the sums of the state over i per row j and over j per
column i -- every point of the grid enters two of those sums -- and all the time steps.  Bars: the
state within 1e-9 (max norm) on the subsample and the boundaries, the row and column sums within
1e-9 per term summed (n x 1e-9 for a sum of n points), the time steps within 1e-10 relative (each is
CFL over a max of the state, so it inherits the state's agreement).
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
CASES = {
    "jet4400_plain": (lambda: jc.make_jet(4400, 640, seed=1234), 0),
    "jet4400_lad": (lambda: jc.make_jet(4400, 640, seed=1234), 1),
    "nozzle4400_plain": (lambda: jc.make_nozzle(4400, 640, i_lip=597, j_in=150, j_out=169, seed=1234), 0),
    "nozzle4400_lad": (lambda: jc.make_nozzle(4400, 640, i_lip=597, j_in=150, j_out=169, seed=1234), 1),
}


def edges(q):
    return np.concatenate([q[:, 0, :], q[:, -1, :], q[0, :, :], q[-1, :, :]], axis=0)


@pytest.mark.parametrize("name", list(CASES))
def test_full_size_vs_oracle(name):
    path = os.path.join(HERE, "golden", f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} not written (tests/golden/make_full_size.py)")
    g = dict(np.load(path))
    factory, shock = CASES[name]
    case = factory()
    assert str(g["input_sha256"]) == gc.input_digest(case)
    assert int(g["shock"]) == shock
    if name.startswith("nozzle"):
        assert [tuple(l) for l in case.bc_lines] == [tuple(l) for l in jc.NOZZLE_BC_4400x640]
    c = tuple(float(x) for x in g["lad_c"])
    if shock:
        assert c == jc.LAD_C
        ctx = capi.Context(case, shock_capturing=1, C_mu=c[0], C_beta=c[1], C_kappa=c[2])
    else:
        ctx = capi.Context(case)
    snaps = [int(x) for x in g["snaps"]]
    done = 0
    for snap in snaps:
        print(f"{name}: running to step {snap}", flush=True)
        ctx.run(snap - done)
        done = snap
        q = ctx.state()
        e_sub = np.abs(q[::16, ::16, :] - g[f"sub{snap}"]).max()
        e_edge = np.abs(edges(q) - g[f"edges{snap}"]).max()
        e_row = np.abs(q.sum(axis=0) - g[f"rowsum{snap}"]).max()
        e_col = np.abs(q.sum(axis=1) - g[f"colsum{snap}"]).max()
        amax = np.abs(q).max(axis=(0, 1))
        print(f"{name} step {snap}: subsample err {e_sub:.3e}, boundary err {e_edge:.3e}, "
              f"row-sum err {e_row:.3e}, column-sum err {e_col:.3e}", flush=True)
        assert np.isfinite(q).all()
        assert e_sub <= STATE_TOL and e_edge <= STATE_TOL
        assert e_row <= case.nz * STATE_TOL and e_col <= case.nr * STATE_TOL
        np.testing.assert_allclose(amax, g[f"absmax{snap}"], rtol=STATE_TOL, atol=STATE_TOL)
    dts, t = ctx.dt_log()
    assert len(dts) == snaps[-1]
    np.testing.assert_allclose(dts, g["dts"], rtol=1e-10, atol=0)
    assert abs(t - float(g["time"])) <= 1e-10 * float(g["time"])
