"""Fixture case registry shared by the fixture writer and the tests (no reference access)."""
from __future__ import annotations

import hashlib
import os

import numpy as np

from axisymmetricjet_amd import jetcase as jc

HERE = os.path.dirname(os.path.abspath(__file__))

LAD_C = jc.LAD_C  # C_mu, C_beta, C_kappa wherever LAD runs with C != 0

WALL_BC = [(1, 91, 1, 1, 1, 1, -1), (1, 93, -2, 1, -1, -1, -1), (1, 93, -1, -1, -1, 1, -1),
           (1, 90, 2, 1, -1, 1, 1)]


def cases():
    """name -> (case factory, shock flag, steps, extra modes)"""
    return {
        "jet88": (lambda: jc.make_jet(88, 32, seed=7), 0, 20, ()),
        "jet440": (lambda: jc.make_jet(440, 64, seed=1234), 0, 100, ()),
        "jet440_lad": (lambda: jc.make_jet(440, 64, seed=1234), 1, 100, ()),
        "swirl64": (lambda: jc.make_jet(64, 40, seed=11, swirl=1e-3, r_max=16.0), 1, 10, ("ladrhs",)),
        "nozzle96": (lambda: jc.make_nozzle(96, 48, i_lip=24, j_in=16, j_out=22), 0, 10, ()),
        "walls64": (lambda: jc.make_jet(64, 32, seed=5, bc_lines=WALL_BC), 0, 10, ()),
        "forcing88": (lambda: jc.make_jet(88, 32, seed=8), 0, 2, ("forcing",)),
    }


def feedback_forcing():
    """The forcing_mat and Gaussian of the forcing fixture (Vz_Control_Vr_Feedback.dat layout:
    25 values, row e outer, then z0, r0, lz, lr): seeded, O(0.1), centred on the shear layer."""
    rng = np.random.default_rng(2024)
    F = np.round(0.1 * rng.standard_normal((5, 5)), 6)
    return jc.Forcing(1, F, z0=3.0, r0=0.5, lz=1.5, lr=0.25)


def input_digest(case) -> str:
    h = hashlib.sha256()
    for a in (case.z, case.r, case.iblank, case.q0, case.target):
        h.update(np.ascontiguousarray(np.asarray(a).ravel(order="F")).tobytes())
    h.update(repr(case.bc_lines).encode())
    return h.hexdigest()


def load(name):
    """(case, fixture dict, shock, nsteps, extra) for a committed fixture."""
    factory, shock, nsteps, extra = cases()[name]
    case = factory()
    g = dict(np.load(os.path.join(HERE, f"{name}.npz")))
    if str(g["input_sha256"]) != input_digest(case):
        raise AssertionError(f"{name}: regenerated input differs from the one the fixture was made from")
    return case, g, shock, nsteps, extra
