"""ajet_probe_at, the probe entry point that allocates nothing (its scratch area is a fixed
device array inside the library), against ajet_set_probes + ajet_probe, through the C ABI."""
import numpy as np
import pytest

from axisymmetricjet_amd import capi
from axisymmetricjet_amd import jetcase as jc

pytestmark = pytest.mark.gpu


def probe_alloc(ctx, pts):
    ctx.set_probes(pts)
    return ctx.probe()


def test_probe_at_matches_probe_on_two_calls():
    case = jc.make_jet(440, 64, seed=3)
    ctx = capi.Context(case)
    ctx.run(3)
    rng = np.random.default_rng(7)
    big = np.stack([rng.integers(0, case.nz, 3000), rng.integers(0, case.nr, 3000)], 1)
    print("probe_at: 3000 points", flush=True)
    got = ctx.probe_at(big)
    assert np.array_equal(got, probe_alloc(ctx, big))
    # a second, smaller call after the state has moved: only its own points come back
    ctx.run(2)
    small = big[::7][:150]
    print("probe_at: 150 points", flush=True)
    got2 = ctx.probe_at(small)
    assert got2.shape == (150, 5)
    assert np.array_equal(got2, probe_alloc(ctx, small))
    q = ctx.state()
    assert np.array_equal(got2, q[small[:, 0], small[:, 1], :])


def test_probe_at_second_context_and_limits():
    case = jc.make_jet(96, 48, seed=4)
    a, b = capi.Context(case), capi.Context(case)
    b.run(1)
    pts = np.array([[0, 0], [95, 47], [10, 20]])
    assert np.array_equal(a.probe_at(pts), probe_alloc(a, pts))
    assert np.array_equal(b.probe_at(pts), probe_alloc(b, pts))
    assert not np.array_equal(a.probe_at(pts), b.probe_at(pts))
    assert a.probe_at(np.zeros((0, 2))).shape == (0, 5)
    full = np.zeros((capi.PROBE_AT_MAX, 2), np.int32)
    assert a.probe_at(full).shape == (capi.PROBE_AT_MAX, 5)
    with pytest.raises(capi.AjetError, match="at most"):
        a.probe_at(np.zeros((capi.PROBE_AT_MAX + 1, 2), np.int32))
    with pytest.raises(capi.AjetError, match="outside"):
        a.probe_at(np.array([[96, 0]]))
