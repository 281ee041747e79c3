"""The learned-LAD oracle pinned against mathematics and an independent library (CPU only):
finite differences for the backward pass, torch's bfloat16 rounding, tanh GELU and Adam, and a
closed form for the stencil features."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def small_model(seed):
    m = mo.init_model(seed)
    rng = np.random.default_rng(seed + 100)
    for k in ("b1", "b2", "b3", "b4"):
        m[k] = (0.1 * rng.standard_normal(m[k].shape)).astype(np.float32)
    m["w4"] = (rng.standard_normal(m["w4"].shape) * 0.2).astype(np.float32)
    return m


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(100000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 100000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3.0e38], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mo.bf16(x), want)


def test_gelu_and_gradient_match_torch():
    x = np.linspace(-8, 8, 4001)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(xt, approximate="tanh")
    y.sum().backward()
    np.testing.assert_allclose(mo.gelu(x), y.detach().numpy(), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(mo.gelu_grad(x), xt.grad.numpy(), rtol=1e-13, atol=1e-14)


def test_backward_matches_finite_differences():
    """FP64 mode (no rounding): every parameter block's gradient against central differences."""
    m = small_model(1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((16, mo.NIN))
    t = rng.standard_normal((16, mo.NOUT))
    loss, g = mo.loss_and_grads(m, x, t, exact=True)
    h = 1e-6
    for k in mo.PARAMS:
        w = np.asarray(m[k], np.float64)
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            mp = {kk: np.asarray(v, np.float64).copy() for kk, v in m.items()}
            mm = {kk: np.asarray(v, np.float64).copy() for kk, v in m.items()}
            mp[k][idx] += h
            mm[k][idx] -= h
            fd = (mo.loss_and_grads(mp, x, t, exact=True)[0] - mo.loss_and_grads(mm, x, t, exact=True)[0]) / (2 * h)
            assert abs(g[k][idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, idx, g[k][idx], fd)


def test_wrong_gelu_derivative_fails_the_fd_check(monkeypatch):
    """The check above catches a plausible mistake: GELU' without its second term."""
    m = small_model(3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((8, mo.NIN))
    t = rng.standard_normal((8, mo.NOUT))
    monkeypatch.setattr(mo, "gelu_grad", lambda z: 0.5 * (1 + np.tanh(mo.K_GELU * (z + mo.A_GELU * z ** 3))))
    _, g = mo.loss_and_grads(m, x, t, exact=True)
    monkeypatch.undo()
    _, g_ok = mo.loss_and_grads(m, x, t, exact=True)
    assert np.abs(g["w1"] - g_ok["w1"]).max() > 1e-3 * np.abs(g_ok["w1"]).max()


def test_bf16_forward_follows_torch_bfloat16_pipeline():
    """The rounding points of the BF16 forward against the same layers in torch (bfloat16
    operands converted to float32, float32 products)."""
    m = small_model(5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((64, mo.NIN)).astype(np.float32)
    y, _ = mo.forward(m, x)
    b = lambda a: torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float64)  # noqa: E731
    a = b(x)
    for l in (1, 2, 3):
        z = (a @ b(m[f"w{l}"]).T).to(torch.float32) + torch.from_numpy(m[f"b{l}"])
        a = torch.nn.functional.gelu(z, approximate="tanh").to(torch.bfloat16).to(torch.float64)
    yt = ((a @ b(m["w4"]).T).to(torch.float32) + torch.from_numpy(m["b4"])).numpy()
    # float32 tanh of numpy and torch may differ by an ulp, which can move a BF16 rounding
    np.testing.assert_allclose(y, yt, rtol=2e-2, atol=2e-2 * np.abs(yt).max())
    assert np.mean(np.abs(y - yt)) <= 1e-3 * np.mean(np.abs(yt))


def test_bf16_gradients_close_to_exact():
    m = small_model(7)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((256, mo.NIN)).astype(np.float32)
    t = rng.standard_normal((256, mo.NOUT)).astype(np.float32)
    l16, g16 = mo.loss_and_grads(m, x, t)
    l64, g64 = mo.loss_and_grads(m, x, t, exact=True)
    assert abs(l16 - l64) <= 2e-2 * l64
    for k in mo.PARAMS:
        assert np.abs(g16[k] - g64[k]).max() <= 5e-2 * np.abs(g64[k]).max(), k


def test_adam_matches_torch():
    m = small_model(9)
    rng = np.random.default_rng(10)
    params = {k: torch.tensor(np.asarray(m[k], np.float32)) for k in mo.PARAMS}
    opt = torch.optim.Adam(list(params.values()), lr=3e-3)
    state = {}
    for step in range(5):
        grads = {k: (rng.standard_normal(m[k].shape) * 0.01).astype(np.float32) for k in mo.PARAMS}
        for k in mo.PARAMS:
            params[k].grad = torch.from_numpy(grads[k].copy())
        opt.step()
        mo.adam_step(m, grads, state, lr=3e-3)
        for k in mo.PARAMS:
            # torch's vectorised CPU lerp fuses a multiply-add, so a rare element differs in
            # its last bits; a wrong formula differs everywhere
            want = params[k].numpy()
            np.testing.assert_allclose(m[k], want, rtol=1e-4, atol=1e-8)
            close = np.isclose(m[k], want, rtol=2e-6, atol=0)
            assert (~close).sum() <= max(2, close.size // 1000), (k, (~close).sum())


def test_features_closed_form_on_a_linear_field():
    nz, nr = 12, 9
    zc = np.linspace(0.0, 11.0, nz)
    rc = np.linspace(0.0, 4.0, nr) ** 1.2
    q = np.zeros((nz, nr, 5))
    ii, jj = np.meshgrid(np.arange(nz), np.arange(nr), indexing="ij")
    coef = {0: (0.25, -0.5), 1: (1.0, 2.0), 3: (-0.75, 0.125), 4: (0.5, 0.0)}
    for v, (a, b) in coef.items():
        q[:, :, v] = 3.0 + a * ii + b * jj
    pts = np.array([[5, 4], [0, 0], [nz - 1, nr - 1], [1, 7]])
    x = mo.features(q, zc, rc, pts)
    for p, (i, j) in enumerate(pts):
        for vi, v in enumerate(mo.STENCIL_VARS):
            a, b = coef[v]
            for db in range(-2, 3):
                for da in range(-2, 3):
                    n = vi * 25 + (db + 2) * 5 + (da + 2)
                    ci, cj = np.clip(i + da, 0, nz - 1), np.clip(j + db, 0, nr - 1)
                    want = q[i, j, v] if (da == 0 and db == 0) else a * (ci - i) + b * (cj - j)
                    assert x[p, n] == np.float32(want)
    dz, dr, ramp = mo.lad_spacings(zc, rc)
    assert np.array_equal(x[:, 100], dz[pts[:, 0]].astype(np.float32))
    assert np.array_equal(x[:, 101], dr[pts[:, 1]].astype(np.float32))
    assert np.array_equal(x[:, 103], rc[pts[:, 1]].astype(np.float32))
    np.testing.assert_allclose(ramp, (1 + np.tanh(5 * (zc - 10))) / 2)
    assert dz[0] == zc[1] - zc[0] and dz[5] == (zc[6] - zc[4]) / 2


@pytest.mark.parametrize("n", [1, 33])
def test_coefficients_are_ramp_scaled_and_non_negative(n):
    m = small_model(11)
    m["out_scale"] = np.array([2.0, 0.5], np.float32)
    rng = np.random.default_rng(12)
    x = rng.standard_normal((n, mo.NIN)).astype(np.float32)
    ramp = rng.uniform(0, 1, n)
    b, k = mo.coefficients(m, x, ramp)
    y, _ = mo.forward(m, x)
    assert np.all(b >= 0) and np.all(k >= 0)
    np.testing.assert_allclose(b, ramp * 2.0 * np.maximum(y[:, 0], 0))
    np.testing.assert_allclose(k, ramp * 0.5 * np.maximum(y[:, 1], 0))
