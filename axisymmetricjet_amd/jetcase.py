"""Synthetic jet cases and the PLOT3D / bc-file formats the solver reads and writes.

The reference ships no grid, initial condition or target file (only a nozzle bc file), so the
cases are generated here:

* grid: z uniform on [0, 30] D, r tanh-stretched on [0, 8] D with r_1 = 0 on the axis, stored as
  an NDIM=3 multi-grid whole double PLOT3D grid with IBLANK, the layout the reference's mesh
  writer uses (WriteMeshAndSoln/write_soln_file.f90:204-214);
* initial condition: hyperbolic-tangent top-hat jet U_z(r) = 0.5 (1 + tanh((0.5 - r) / (2 theta)))
  U_j with Crocco-Busemann density and T = 1/rho (uniform p), the profile construction of
  write_soln_file.f90:34-77, plus seeded Gaussian noise on Vr and Vz (Vr = 0 on the axis);
* target: the noise-free initial condition;
* bc lines in the format ModPatch.f90:49 reads.

The solution slots hold the primitives rho, Vr, Vtheta, Vz, T (ModInputOutput.f90:238-250).
"""
from __future__ import annotations

import dataclasses
import struct

import numpy as np

GAMMA = 1.4
TJ_TINF = 0.6  # jet-to-ambient temperature ratio of the profile (write_soln_file.f90:31-33)


@dataclasses.dataclass
class JetCase:
    nz: int
    nr: int
    z: np.ndarray          # (nz, nr) Fortran order
    r: np.ndarray          # (nz, nr)
    iblank: np.ndarray     # (nz, nr) int32
    q0: np.ndarray         # (nz, nr, 5) initial primitives
    target: np.ndarray     # (nz, nr, 5)
    bc_lines: list         # [(grid, bc, dir, is, ie, js, je), ...]
    mach: float
    re: float
    pr: float


@dataclasses.dataclass
class Forcing:
    """Forcing terms (ajet_capi.h AJET_FORCING_*): mode bits 1 = ControlForcingTerms feedback
    alpha*F(e,k)*G*(q_k - q_target_k) (ModControlForcing.f90:46-62), 2 = inflow excitation
    amp*Uj*sin(2 pi St Uj (t0 + t))*G on the Vz equation; G the Gaussian at (z0, r0) of widths
    lz, lr.  F is forcing_mat, row e = equation."""
    mode: int
    F: np.ndarray
    z0: float = 1.0
    r0: float = 0.5
    lz: float = 0.5
    lr: float = 0.1
    amp: float = 0.0
    strouhal: float = 0.0
    uj: float = 0.0
    t0: float = 0.0


def excitation(mach: float, strouhal: float, amp: float, z0: float = 1.0, r0: float = 0.5, lz: float = 0.5,
               lr: float = 0.1, feedback=None) -> Forcing:
    """Inflow excitation of the jet of Mach `mach` (Uj from jet_profile's construction) at the
    shear layer, amplitude `amp` as a fraction of Uj; feedback = optional 5x5 forcing_mat too."""
    uj = mach * np.sqrt(TJ_TINF)
    F = np.zeros((5, 5)) if feedback is None else np.asarray(feedback, dtype=np.float64)
    mode = 2 | (1 if feedback is not None else 0)
    return Forcing(mode, F, z0, r0, lz, lr, amp, strouhal, float(uj), 0.0)


def ensemble(n: int = 64, seed: int = 2026):
    """The seeded ensemble of jets: grid sizes uniform over {440..2200} x {64..320} in steps of 8,
    jet Mach uniform in [1.2, 1.8], inflow excitation with Strouhal number uniform in [0.1, 1.0]
    and amplitude uniform in [0.1%, 5%] of Uj; the initial state is the noise-free profile, the
    excitation being the perturbation.  Returns [(case, forcing), ...].

    Under the reference's numerics (no filter with shock capturing off) these runs stay bounded
    for a few hundred steps and then diverge: the oracle, which matches the reference bit for bit,
    takes member 38 (520 x 216) to a non-finite state near step 650."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nz = int(rng.integers(440 // 8, 2200 // 8 + 1)) * 8
        nr = int(rng.integers(64 // 8, 320 // 8 + 1)) * 8
        mach = float(rng.uniform(1.2, 1.8))
        st = float(rng.uniform(0.1, 1.0))
        amp = float(rng.uniform(0.001, 0.05))
        case = make_jet(nz, nr, seed=seed * 1000 + k, mach=mach, noise=0.0)
        out.append((case, excitation(mach, strouhal=st, amp=amp)))
    return out


def combined_specs(n: int = 32, seed: int = 4242):
    """The seeded forcing sweep of the combined learned-LAD form, as parameters (the 4400 x 640
    cases are large, so the caller builds them one at a time): jet Mach uniform in [1.4, 1.6],
    Strouhal number in [0.2, 0.8], excitation amplitude in [0.5%, 3%] of Uj, noise seed per
    member.  Returns [(mach, strouhal, amp, noise_seed), ...]."""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(1.4, 1.6)), float(rng.uniform(0.2, 0.8)), float(rng.uniform(0.005, 0.03)),
             seed * 1000 + k) for k in range(n)]


# LAD constants C_mu, C_beta, C_kappa of the shock-capturing runs (the reference ships 0, 0, 0).
# C_beta is a tenth of the usual 1.75: the reference's time step ignores the artificial
# viscosities, and with 1.75 the seeded grid-scale noise drives beta* to ~70 at the outflow
# corner of the 4400 x 640 jet and the run diverges within 5 steps; with 0.175 it stays bounded.
LAD_C = (0.002, 0.175, 0.01)


def stretched_r(nr: int, r_max: float = 8.0, delta: float = 2.0) -> np.ndarray:
    """tanh stretching on [0, r_max], fine at the axis and the shear layer, r[0] = 0."""
    eta = np.arange(nr, dtype=np.float64) / (nr - 1)
    r = r_max * (1.0 - np.tanh(delta * (1.0 - eta)) / np.tanh(delta))
    r[0] = 0.0
    return r


def uniform_z(nz: int, z_max: float = 30.0) -> np.ndarray:
    return z_max * np.arange(nz, dtype=np.float64) / (nz - 1)


def jet_profile(r: np.ndarray, mach: float = 1.5, theta: float = 0.05, tj_tinf: float = TJ_TINF,
                gam: float = GAMMA, npr: float = 1.0):
    """Top-hat jet with Crocco-Busemann density (write_soln_file.f90:34-77 construction).

    Returns rho, Vz, T on the given radii.  npr = 1 (the reference generator's jet): T = 1/rho,
    so p = rho T = 1.  npr > 1: an under-expanded jet, p = 1 + (npr - 1) u(r) with the same top
    hat u and the same T, rho = p / T (exactly the npr = 1 profile when npr = 1).
    """
    u = 0.5 * (1.0 + np.tanh((0.5 - r) / (2.0 * theta)))
    uj = mach * np.sqrt(tj_tinf)
    term1 = 0.5 * (gam - 1.0) * u * (1.0 - u) * mach ** 2
    term2 = u
    term3 = (1.0 / tj_tinf) * (1.0 - u)
    rho_rhoj = 1.0 / (term1 + term2 + term3)
    rho = rho_rhoj / tj_tinf
    t = 1.0 / rho
    p = 1.0 + (npr - 1.0) * u
    return rho * p, u * uj, t


DEFAULT_BC = [
    (1, 91, 1, 1, 1, 1, -1),     # SAT far-field inflow on i = 1
    (1, 92, -2, 1, -1, -1, -1),  # NSCBC non-reflecting top j = Nr
    (1, 92, -1, -1, -1, 1, -1),  # NSCBC non-reflecting outflow i = Nz
    (1, 90, 2, 1, -1, 1, 1),     # axis j = 1
]


def make_jet(nz: int, nr: int, seed: int = 1234, mach: float = 1.5, noise: float = 1e-3,
             theta: float = 0.05, re: float = 1.0e4, pr: float = 0.7,
             bc_lines=None, r_max: float = 8.0, swirl: float = 0.0, npr: float = 1.0) -> JetCase:
    """The jet case on an nz x nr grid.  swirl > 0 adds seeded noise to Vtheta as well (a
    test variant: the configured cases keep Vtheta = 0); npr > 1 makes the jet under-expanded
    (jet static pressure npr times the ambient), which forms shock cells."""
    zc = uniform_z(nz)
    rc = stretched_r(nr, r_max=r_max)
    z = np.asfortranarray(np.repeat(zc[:, None], nr, axis=1))
    r = np.asfortranarray(np.repeat(rc[None, :], nz, axis=0))
    rho, vz, t = jet_profile(rc, mach=mach, theta=theta, npr=npr)
    target = np.zeros((nz, nr, 5), dtype=np.float64, order="F")
    target[:, :, 0] = rho[None, :]
    target[:, :, 3] = vz[None, :]
    target[:, :, 4] = t[None, :]
    q0 = np.array(target, order="F", copy=True)
    rng = np.random.default_rng(seed)
    # draw in Fortran order so the noise field does not depend on the array layout
    nvr = rng.standard_normal(nz * nr).reshape((nz, nr), order="F")
    nvz = rng.standard_normal(nz * nr).reshape((nz, nr), order="F")
    q0[:, :, 1] += noise * nvr
    q0[:, :, 3] += noise * nvz
    if swirl:
        q0[:, :, 2] += swirl * rng.standard_normal(nz * nr).reshape((nz, nr), order="F")
        q0[:, 0, 2] = 0.0
    q0[:, 0, 1] = 0.0
    iblank = np.ones((nz, nr), dtype=np.int32, order="F")
    return JetCase(nz, nr, z, r, iblank, q0, target,
                   list(bc_lines if bc_lines is not None else DEFAULT_BC), mach, re, pr)


# --------------------------------------------------------------------------- file formats

def _rec(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def write_grid(path: str, z: np.ndarray, r: np.ndarray, iblank: np.ndarray) -> None:
    """NDIM=3, multi-grid ('m'), whole ('w'), double ('d'), with IBLANK ('y')."""
    nz, nr = z.shape
    xyz = np.concatenate([np.asarray(z, np.float64).ravel(order="F"),
                          np.asarray(r, np.float64).ravel(order="F"),
                          np.zeros(nz * nr)]).astype("<f8")
    with open(path, "wb") as f:
        f.write(_rec(struct.pack("<i", 1)))
        f.write(_rec(struct.pack("<3i", nz, nr, 1)))
        f.write(_rec(xyz.tobytes() + np.asarray(iblank, "<i4").ravel(order="F").tobytes()))


def write_soln(path: str, q: np.ndarray, tau=(0.0, 0.0, 0.0, 0.0)) -> None:
    """Solution file as Write_Soln writes it (ModPLOT3D_IO.f90:539-599), NDIM=3, 'm','w','d'."""
    nz, nr, _ = q.shape
    with open(path, "wb") as f:
        f.write(_rec(struct.pack("<i", 1)))
        f.write(_rec(struct.pack("<3i", nz, nr, 1)))
        f.write(_rec(struct.pack("<4d", *tau)))
        f.write(_rec(np.asarray(q, "<f8").ravel(order="F").tobytes()))


def _records(path: str):
    return _records_of(open(path, "rb").read(), path)


def _records_of(data: bytes, path: str = "data"):
    pos, out = 0, []
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        out.append(data[pos + 4: pos + 4 + n])
        (m,) = struct.unpack_from("<i", data, pos + 4 + n)
        if m != n:
            raise ValueError(f"{path}: record marker mismatch at byte {pos}")
        pos += 8 + n
    return out


def read_grid(path: str):
    recs = _records(path)
    if len(recs[0]) != 4 or struct.unpack("<i", recs[0])[0] != 1:
        raise ValueError("only single-block multi-grid files are supported")
    nz, nr, nk = struct.unpack("<3i", recs[1])
    npts = nz * nr * nk
    body = recs[2]
    xyz = np.frombuffer(body[: 3 * npts * 8], "<f8")
    if len(body) == 3 * npts * 8 + npts * 4:
        ib = np.frombuffer(body[3 * npts * 8:], "<i4").reshape((nz, nr), order="F")
    else:
        ib = np.ones((nz, nr), np.int32)
    z = xyz[:npts].reshape((nz, nr), order="F")
    r = xyz[npts: 2 * npts].reshape((nz, nr), order="F")
    return np.asfortranarray(z), np.asfortranarray(r), np.asfortranarray(ib.astype(np.int32))


def read_soln(path: str):
    return _soln_of(_records(path))


def read_soln_bytes(data: bytes):
    """read_soln of a file's contents."""
    return _soln_of(_records_of(data))


def _soln_of(recs):
    nz, nr, nk = struct.unpack("<3i", recs[1])
    tau = struct.unpack("<4d", recs[2])
    q = np.frombuffer(recs[3], "<f8").reshape((nz, nr, 5), order="F")
    return np.asfortranarray(q), tau


def write_bc(path: str, lines) -> None:
    with open(path, "w") as f:
        for (g, bc, d, is_, ie, js, je) in lines:
            f.write(f"{g} {bc} {d} {is_} {ie} {js} {je} 1 1\n")


def read_bc(path: str):
    out = []
    for line in open(path):
        parts = line.split()
        if len(parts) < 7:
            break
        out.append(tuple(int(x) for x in parts[:7]))
    return out


NOZZLE_BC_4400x640 = [
    # bc_4400x640_nozzle.dat, the one bc file the reference ships
    (1, 91, 1, 1, 1, 1, 150), (1, 94, -2, 1, 597, 150, 150), (1, 94, 1, 597, 597, 150, 169),
    (1, 94, 2, 1, 597, 169, 169), (1, 91, 1, 1, 1, 169, -1), (1, 92, -2, 1, -1, -1, -1),
    (1, 92, -1, -1, -1, 1, -1), (1, 90, 2, 1, -1, 1, 1),
]


def make_nozzle(nz: int, nr: int, i_lip: int, j_in: int, j_out: int, seed: int = 99,
                mach: float = 1.5, re: float = 1.0e4, pr: float = 0.7, r_max: float = 8.0) -> JetCase:
    """A jet issuing from a nozzle lip: the bc layout of bc_4400x640_nozzle.dat scaled to any
    grid (inner wall j = j_in, outer wall j = j_out, lip face i = i_lip), with the points
    inside the lip (i < i_lip, j_in < j < j_out) blanked."""
    bc = [(1, 91, 1, 1, 1, 1, j_in), (1, 94, -2, 1, i_lip, j_in, j_in), (1, 94, 1, i_lip, i_lip, j_in, j_out),
          (1, 94, 2, 1, i_lip, j_out, j_out), (1, 91, 1, 1, 1, j_out, -1), (1, 92, -2, 1, -1, -1, -1),
          (1, 92, -1, -1, -1, 1, -1), (1, 90, 2, 1, -1, 1, 1)]
    case = make_jet(nz, nr, seed=seed, mach=mach, re=re, pr=pr, bc_lines=bc, r_max=r_max)
    case.iblank[: i_lip - 1, j_in: j_out - 1] = 0
    return case


def write_case(prefix: str, case: JetCase):
    """Write grid, IC, target and bc files; returns their paths."""
    paths = dict(grid=prefix + "mesh.xyz", ic=prefix + "ic.q", target=prefix + "target.q",
                 bc=prefix + "bc.dat")
    write_grid(paths["grid"], case.z, case.r, case.iblank)
    write_soln(paths["ic"], case.q0)
    write_soln(paths["target"], case.target)
    write_bc(paths["bc"], case.bc_lines)
    return paths
