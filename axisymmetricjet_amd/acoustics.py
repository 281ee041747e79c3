"""Scoring of a jet run (an extension: the reference writes solution files only): pressure
histories on an arc of probes, their spectra and overall sound pressure level, and the positions
of the compression peaks (shocks) along a grid line.

Pressure is the solver's p = rho T (AssignValues, ModAssign.f90:20), in units of the ambient
pressure, so OASPL here is 10 log10(<p'^2>) in dB relative to the ambient pressure; scores use
differences between runs, which do not depend on the reference level.
"""
from __future__ import annotations

import numpy as np

ARC_ANGLES = (30.0, 45.0, 60.0, 75.0, 90.0, 105.0, 120.0, 135.0, 150.0)


def arc_points(zc, rc, centre=15.0, radius=6.5, angles=ARC_ANGLES):
    """Grid points (i, j) nearest to an arc of the given radius around (z, r) = (centre, 0),
    angles in degrees from the downstream jet axis."""
    pts = []
    for a in angles:
        z = centre + radius * np.cos(np.radians(a))
        r = radius * np.sin(np.radians(a))
        pts.append((int(np.argmin(np.abs(zc - z))), int(np.argmin(np.abs(rc - r)))))
    return np.array(pts, dtype=np.int32)


def record(ctx, nsteps, every=2):
    """Advance ctx by nsteps (a multiple of `every`), sampling the probes' pressure after each
    `every` steps: returns times (m,) and pressures (m, nprobes) since the first call's start.
    Stops at the first sample that is not finite (or a step whose time-step reduction meets a
    non-finite value), returning the finite samples before it."""
    times, ps = [], []
    for _ in range(nsteps // every):
        try:
            ctx.run(every)
            v = ctx.probe()
            t = ctx.dt_log(0)[1]
        except RuntimeError:
            break
        p = v[:, 0] * v[:, 4]
        if not (np.all(np.isfinite(p)) and np.isfinite(t)):
            break
        ps.append(p)
        times.append(t)
    return np.array(times), np.array(ps).reshape(len(times), -1)


def oaspl(p):
    """OASPL per probe of pressure histories p (m, nprobes), in dB re the ambient pressure."""
    pf = p - p.mean(axis=0, keepdims=True)
    return 10.0 * np.log10(np.mean(pf * pf, axis=0) + 1e-300)


def lowest_strouhal(t, uj, nseg=4, d=1.0):
    """The lowest Strouhal number f D / Uj that spectrum() resolves over a record of times t:
    its segments span a quarter (1/nseg) of the record, so it is nseg D / (Uj T)."""
    return nseg * d / (uj * (t[-1] - t[0]))


def spectrum(t, p, nseg=4):
    """Welch power spectral density of each probe's history, resampled to a uniform time
    grid: (f, psd (nf, nprobes))."""
    from scipy.signal import welch

    tu = np.linspace(t[0], t[-1], len(t))
    pu = np.stack([np.interp(tu, t, p[:, k]) for k in range(p.shape[1])], axis=1)
    fs = 1.0 / (tu[1] - tu[0])
    f, psd = welch(pu - pu.mean(axis=0), fs=fs, nperseg=max(8, len(tu) // nseg), axis=0)
    return f, psd


def spectrum_error_db(psd_a, psd_b, floor=1e-30):
    """Mean absolute level difference in dB between two spectra, over the bins above 1e-3 of
    the peak of the second."""
    keep = psd_b > 1e-3 * psd_b.max(axis=0, keepdims=True)
    d = np.abs(10.0 * np.log10((psd_a + floor) / (psd_b + floor)))
    return float(d[keep].mean()) if keep.any() else 0.0


def compression(q, zc, j):
    """max(0, -dVz/dz) along grid line j of a state q (nz, nr, 5)."""
    vz = q[:, j, 3]
    d = np.gradient(vz, zc)
    return np.maximum(0.0, -d)


def shock_positions(q, zc, j, rel=0.3, zmin=10.0):
    """z of the local maxima of the compression along line j above rel times its maximum
    (downstream of zmin, where the LAD acts)."""
    c = compression(q, zc, j)
    c = np.where(zc >= zmin, c, 0.0)
    top = c.max()
    if top <= 0.0:
        return np.zeros(0)
    k = np.where((c[1:-1] > c[:-2]) & (c[1:-1] >= c[2:]) & (c[1:-1] >= rel * top))[0] + 1
    return zc[k]


def shock_shift(za, zb):
    """Mean distance from each peak of za to the nearest peak of zb (and back), in diameters;
    0 when both are empty, inf when only one is."""
    if len(za) == 0 and len(zb) == 0:
        return 0.0
    if len(za) == 0 or len(zb) == 0:
        return float("inf")
    d1 = np.abs(za[:, None] - zb[None, :]).min(axis=1)
    d2 = np.abs(zb[:, None] - za[None, :]).min(axis=1)
    return float(0.5 * (d1.mean() + d2.mean()))
