"""ORACLE -- test infrastructure only.

ctypes binding of the C restatement in ajet_oracle.c (liboracle.so).  Only tests/, the smoke
check and the benchmark's CPU-baseline leg import this module, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class Params(C.Structure):
    _fields_ = [("nz", C.c_int32), ("nr", C.c_int32), ("viscous", C.c_int32),
                ("shock_capturing", C.c_int32), ("forcing", C.c_int32), ("reserved", C.c_int32),
                ("gam", C.c_double), ("Re", C.c_double), ("Pr", C.c_double), ("CFL", C.c_double),
                ("C_mu", C.c_double), ("C_beta", C.c_double), ("C_kappa", C.c_double),
                ("sponge_start_top", C.c_double), ("sponge_A_top", C.c_double),
                ("sponge_end_left", C.c_double), ("sponge_A_left", C.c_double)]


class BcLine(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("grid", "bc", "bc_dir", "i_start", "i_end", "j_start", "j_end")]


class Patch(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("i_start", "i_end", "j_start", "j_end", "bc", "bc_dir")]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "ajet_oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                               "-fopenmp", "-o", LIB, src, "-lm"])
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = np.ctypeslib.ndpointer(np.float64, flags="F_CONTIGUOUS")
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), dp, dp, C.c_void_p, C.POINTER(BcLine), C.c_int32]
        L.orc_destroy.argtypes = [C.c_void_p]
        for name, args in {
            "orc_set_state": [C.c_void_p, dp, dp],
            "orc_get_state": [C.c_void_p, dp],
            "orc_rhs": [C.c_void_p, dp, dp],
            "orc_rk4": [C.c_void_p],
            "orc_lad": [C.c_void_p, dp, dp, dp, dp],
            "orc_lad_stages": [C.c_void_p, dp, dp, dp, dp],
            "orc_metrics": [C.c_void_p, dp, dp, dp, dp],
            "orc_tables": [C.c_void_p, dp, dp, dp],
            "orc_d1": [C.c_void_p, dp, C.c_int, dp],
            "orc_d2": [C.c_void_p, dp, C.c_int, dp, dp],
            "orc_d2comp": [C.c_void_p, dp, C.c_int, dp],
            "orc_tgfilter": [C.c_void_p, dp, C.c_int, C.c_int, dp],
            "orc_set_dt": [C.c_void_p, C.c_double],
            "orc_set_stale_T": [C.c_void_p, C.c_double],
        }.items():
            getattr(L, name).argtypes = args
        L.orc_compute_dt.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_run.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_time.restype = C.c_double
        L.orc_time.argtypes = [C.c_void_p]
        L.orc_stale_T.restype = C.c_double
        L.orc_stale_T.argtypes = [C.c_void_p]
        L.orc_patches.argtypes = [C.c_void_p, C.POINTER(Patch), C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_forcing.argtypes = [C.c_void_p, C.c_int, C.c_void_p] + [C.c_double] * 8
        _lib = L
    return _lib


def make_params(nz, nr, re=1000.0, pr=0.72, cfl=0.5, shock=0, viscous=1, c_mu=0.0, c_beta=0.0,
                c_kappa=0.0, gam=1.4):
    return Params(nz, nr, viscous, shock, 0, 0, gam, re, pr, cfl, c_mu, c_beta, c_kappa,
                  12.5, 0.5, 6.0, 5.0)


def bc_array(lines):
    arr = (BcLine * max(1, len(lines)))()
    for k, l in enumerate(lines):
        arr[k] = BcLine(*[int(x) for x in l[:7]])
    return arr, len(lines)


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


class Oracle:
    """One single-domain restatement of the reference solver state."""

    def __init__(self, case, shock=0, c_mu=0.0, c_beta=0.0, c_kappa=0.0, viscous=1, cfl=0.5,
                 params=None, threads=0, forcing=None):
        L = lib()
        if threads:
            L.orc_set_threads(int(threads))
        self.nz, self.nr = case.nz, case.nr
        self.p = params or make_params(case.nz, case.nr, re=case.re, pr=case.pr, cfl=cfl, shock=shock,
                                       viscous=viscous, c_mu=c_mu, c_beta=c_beta, c_kappa=c_kappa)
        bcs, n = bc_array(case.bc_lines)
        ib = np.asfortranarray(case.iblank, dtype=np.int32)
        self._ib = ib
        self.h = L.orc_create(C.byref(self.p), F(case.z), F(case.r), ib.ctypes.data, bcs, n)
        if not self.h:
            raise RuntimeError("orc_create failed")
        L.orc_set_state(self.h, F(case.q0), F(case.target))
        if forcing is not None:
            fm = np.ascontiguousarray(forcing.F, dtype=np.float64)
            self._fm = fm
            L.orc_set_forcing(self.h, int(forcing.mode), fm.ctypes.data, forcing.z0, forcing.r0, forcing.lz,
                              forcing.lr, forcing.amp, forcing.strouhal, forcing.uj, forcing.t0)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, k=5):
        return (self.nz, self.nr, k) if k else (self.nz, self.nr)

    def set_state(self, q, target):
        lib().orc_set_state(self.h, F(q), F(target))

    def state(self):
        out = np.zeros(self.shape(), order="F")
        lib().orc_get_state(self.h, out)
        return out

    def rhs(self, q):
        out = np.zeros(self.shape(), order="F")
        lib().orc_rhs(self.h, F(q), out)
        return out

    def compute_dt(self):
        d = C.c_double()
        rc = lib().orc_compute_dt(self.h, C.byref(d))
        return d.value, rc

    def run(self, nsteps):
        dts = np.zeros(nsteps)
        rc = lib().orc_run(self.h, nsteps, dts.ctypes.data)
        if rc:
            raise RuntimeError(f"orc_run rc={rc}")
        return dts

    def time(self):
        return lib().orc_time(self.h)

    def stale_T(self):
        return lib().orc_stale_T(self.h)

    def lad(self, q):
        outs = [np.zeros(self.shape(0), order="F") for _ in range(3)]
        lib().orc_lad(self.h, F(q), *outs)
        return outs

    def lad_stages(self, q):
        outs = [np.zeros((self.nz, self.nr, 3), order="F") for _ in range(3)]
        lib().orc_lad_stages(self.h, F(q), *outs)
        return outs

    def metrics(self):
        outs = [np.zeros(self.shape(0), order="F") for _ in range(4)]
        lib().orc_metrics(self.h, *outs)
        return outs

    def tables(self):
        a = np.zeros((9, 6), order="F")
        b = np.zeros((9, 6), order="F")
        t = np.zeros((9, 11), order="F")
        lib().orc_tables(self.h, a, b, t)
        return a, b, t

    def patches(self):
        arr = (Patch * 64)()
        n = lib().orc_patches(self.h, arr, 64)
        return [(p.i_start, p.i_end, p.j_start, p.j_end, p.bc, p.bc_dir) for p in arr[:n]]

    def d1(self, f, dirn):
        out = np.zeros(self.shape(0), order="F")
        lib().orc_d1(self.h, F(f), dirn, out)
        return out

    def d2(self, f, dirn, first):
        out = np.zeros(self.shape(0), order="F")
        lib().orc_d2(self.h, F(f), dirn, F(first), out)
        return out

    def d2comp(self, f, dirn):
        out = np.zeros(self.shape(0), order="F")
        lib().orc_d2comp(self.h, F(f), dirn, out)
        return out

    def tgfilter(self, f, dirn, inplace):
        out = np.zeros(self.shape(0), order="F")
        lib().orc_tgfilter(self.h, F(f), dirn, int(inplace), out)
        return out
