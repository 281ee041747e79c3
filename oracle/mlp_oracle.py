"""CPU oracle of the learned LAD coefficient (test infrastructure: only tests, smoke() and the
bench's CPU leg use it, as the checker, never as the product path).

The network has no counterpart in the reference: it stands in for ComputeArtTransCoeffs
(ModShockCapturing.f90:5-106), whose beta* and kappa* it is trained to reproduce, scaled by
the reference's ramp (1 + tanh(5 (z - 10))) / 2 (ModShockCapturing.f90:85).  This module states
the arithmetic the GPU kernels must follow, rounding point by rounding point:

  features   x (104, FP32): the 5 x 5 stencil of rho, Vr, Vz, T around (i, j) (indices clamped
             to the grid), the centre as its value and the 24 neighbours as differences from
             it, then dz(i), dr(j) (the LAD spacings), ramp(i) and r(j)
  normalise  x_n = (x - shift) * scale in FP32, rounded to BF16 (round to nearest even)
  layers     z_l = a_{l-1} W_l^T + b_l with BF16 a and W and FP32 accumulation;
             a_l = bf16(gelu(z_l)) for l = 1..3 (tanh form of GELU, FP32); y = z_4 (FP32)
  output     beta* = ramp * out_scale[0] * max(y_0, 0), kappa* likewise with y_1; mu* = 0
  loss       L = sum((y - t)^2) / (2 n) over the n samples and both outputs
  backward   every GEMM operand in BF16 (dy, dz_l rounded), FP32 accumulation; bias gradients
             are FP32 sums of the BF16-rounded dz
  Adam       torch.optim.Adam's update in FP32: m += (1 - b1)(g - m), v = v b2 + ((1 - b2) g) g,
             p += (-lr / bc1) m / (sqrt(v) / sqrt(bc2) + eps)

Pinning: the backward is checked against central finite differences of the forward in the
FP64 mode (exact=True: no BF16 rounding), the GELU against torch's tanh GELU, the rounding
against torch.bfloat16 and Adam against torch.optim.Adam (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

NIN, NHID, NOUT = 104, 128, 2
STENCIL_VARS = (0, 1, 3, 4)  # rho, Vr, Vz, T of the primitive state
K_GELU = 0.7978845608028654  # sqrt(2 / pi)
A_GELU = 0.044715


def bf16(x):
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def gelu(x):
    dt = x.dtype
    k, a = dt.type(K_GELU), dt.type(A_GELU)
    t = np.tanh(k * (x + a * x * x * x))
    return dt.type(0.5) * x * (dt.type(1) + t)


def gelu_grad(x):
    dt = x.dtype
    k, a = dt.type(K_GELU), dt.type(A_GELU)
    t = np.tanh(k * (x + a * x * x * x))
    return dt.type(0.5) * (dt.type(1) + t) + dt.type(0.5) * x * (dt.type(1) - t * t) * k * (
        dt.type(1) + dt.type(3) * a * x * x)


def lad_spacings(zc, rc):
    """dz(i), dr(j), ramp(i) of the LAD: half the centred difference of the coordinates,
    one-sided at the ends (ModShockCapturing.f90:171-189), and the ramp (:85)."""
    def sp(c):
        d = np.zeros_like(c)
        if len(c) > 2:
            d[1:-1] = (c[2:] - c[:-2]) / 2.0
        if len(c) > 1:
            d[0] = c[1] - c[0]
            d[-1] = c[-1] - c[-2]
        return d
    return sp(zc), sp(rc), (1.0 + np.tanh(5.0 * (zc - 10.0))) / 2.0


def features(q, zc, rc, points):
    """Raw FP32 features (n, 104) of the points (n, 2) = (i, j), 0-based, of the state q
    (nz, nr, 5)."""
    nz, nr = q.shape[:2]
    pts = np.asarray(points, dtype=np.int64)
    i, j = pts[:, 0], pts[:, 1]
    dz, dr, ramp = lad_spacings(np.asarray(zc, np.float64), np.asarray(rc, np.float64))
    out = np.zeros((len(pts), NIN), dtype=np.float32)
    for vi, v in enumerate(STENCIL_VARS):
        f = q[:, :, v]
        c = f[i, j]
        for b in range(-2, 3):
            for a in range(-2, 3):
                n = vi * 25 + (b + 2) * 5 + (a + 2)
                ii = np.clip(i + a, 0, nz - 1)
                jj = np.clip(j + b, 0, nr - 1)
                out[:, n] = (c if (a == 0 and b == 0) else f[ii, jj] - c).astype(np.float32)
    out[:, 100] = dz[i].astype(np.float32)
    out[:, 101] = dr[j].astype(np.float32)
    out[:, 102] = ramp[i].astype(np.float32)
    out[:, 103] = np.asarray(rc, np.float64)[j].astype(np.float32)
    return out


def normalise(x, model):
    return (x.astype(np.float32) - model["shift"].astype(np.float32)) * model["scale"].astype(np.float32)


def _w(model, name, exact):
    w = np.asarray(model[name], dtype=np.float64 if exact else np.float32)
    return w if exact else bf16(w)


def _mm(a, b_t, exact):
    """a (n, k) times b_t^T (m, k) with FP32 (or FP64) accumulation."""
    if exact:
        return a.astype(np.float64) @ b_t.astype(np.float64).T
    return (a.astype(np.float64) @ b_t.astype(np.float64).T).astype(np.float32)


def forward(model, xn, exact=False):
    """The layers on normalised features xn (n, 104).  exact=True: FP64, no rounding."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda v: v) if exact else bf16
    cache = {"a0": rnd(np.asarray(xn, dtype=dt))}
    a = cache["a0"]
    for l in (1, 2, 3):
        z = (_mm(a, _w(model, f"w{l}", exact), exact) + np.asarray(model[f"b{l}"], dt)).astype(dt)
        cache[f"z{l}"] = z
        a = rnd(gelu(z))
        cache[f"a{l}"] = a
    y = (_mm(a, _w(model, "w4", exact), exact) + np.asarray(model["b4"], dt)).astype(dt)
    cache["y"] = y
    return y, cache


def loss_and_grads(model, xn, t, exact=False):
    """L = sum((y - t)^2) / (2 n) and its gradients with respect to w1..w4, b1..b4."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda v: v) if exact else bf16
    y, c = forward(model, xn, exact)
    n = y.shape[0]
    t = np.asarray(t, dt)
    loss = float(np.sum((y.astype(np.float64) - t) ** 2) / (2 * n))
    dz = rnd(((y - t) / dt(n)).astype(dt))
    g = {}
    for l in (4, 3, 2, 1):
        a_prev = c[f"a{l - 1}"]
        g[f"w{l}"] = _mm(dz.T, a_prev.T, exact)
        g[f"b{l}"] = dz.astype(np.float64).sum(axis=0).astype(dt)
        if l > 1:
            da = _mm(dz, _w(model, f"w{l}", exact).T, exact)
            dz = rnd((da * gelu_grad(c[f"z{l - 1}"])).astype(dt))
    return loss, g


PARAMS = ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")


def adam_step(model, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One torch.optim.Adam step (default hyper-parameters, no weight decay) in FP32."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    bc1 = 1.0 - beta1 ** t
    bc2 = 1.0 - beta2 ** t
    f = np.float32
    for k in PARAMS:
        p = np.asarray(model[k], np.float32)
        g = np.asarray(grads[k], np.float32)
        m = state.setdefault("m_" + k, np.zeros_like(p))
        v = state.setdefault("v_" + k, np.zeros_like(p))
        m[...] = m + f(1.0 - beta1) * (g - m)
        v[...] = v * f(beta2) + (f(1.0 - beta2) * g) * g
        denom = np.sqrt(v) / f(np.sqrt(bc2)) + f(eps)
        model[k] = (p + (f(-(lr / bc1)) * m) / denom).astype(np.float32)
    return model


def init_model(seed=0):
    """He-style initialisation of the four layers, identity normalisation."""
    rng = np.random.default_rng(seed)
    m = {}
    dims = [(NHID, NIN), (NHID, NHID), (NHID, NHID), (NOUT, NHID)]
    for l, (o, i) in enumerate(dims, start=1):
        m[f"w{l}"] = (rng.standard_normal((o, i)) * np.sqrt(2.0 / i)).astype(np.float32)
        m[f"b{l}"] = np.zeros(o, np.float32)
    m["w4"] *= np.float32(0.1)
    m["shift"] = np.zeros(NIN, np.float32)
    m["scale"] = np.ones(NIN, np.float32)
    m["out_scale"] = np.ones(NOUT, np.float32)
    return m


def coefficients(model, x_raw, ramp):
    """beta*, kappa* of the network for raw features x_raw (n, 104) and ramp (n,)."""
    y, _ = forward(model, normalise(x_raw, model))
    r = np.asarray(ramp, np.float64)
    s = np.asarray(model["out_scale"], np.float64)
    return r * s[0] * np.maximum(y[:, 0].astype(np.float64), 0.0), r * s[1] * np.maximum(y[:, 1].astype(np.float64), 0.0)
