"""ctypes binding of libajet.so (include/ajet_capi.h) and a small host-side mirror of the
reference's driver: ModSolver's setup, then {ComputeTimeStep; RK4} per iteration
(ModSolver.f90:19-125).  Loading fails loudly if the HIP library is not built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("AJET_LIB") or os.path.join(HERE, "libajet.so")

AJET_OK = 0
ERRORS = {-1: "EINVAL", -2: "EGRID", -3: "EPATCH", -4: "ENODEV", -5: "EBLOWUP", -6: "ENONFINITE",
          -7: "ETILE", -8: "EUNSUPPORTED"}

EXPORTS = ["ajet_params_default", "ajet_build_patches", "ajet_create", "ajet_destroy", "ajet_last_error",
           "ajet_set_state", "ajet_get_state", "ajet_compute_dt", "ajet_rk4_stage", "ajet_rk4_step", "ajet_run",
           "ajet_get_dt_log", "ajet_rhs", "ajet_lad_coeffs", "ajet_set_forcing", "ajet_batch_create",
           "ajet_batch_run", "ajet_batch_destroy", "ajet_batch_last_error", "ajet_stream", "ajet_version",
           "ajet_set_learned_lad", "ajet_lad_features", "ajet_lad_learned_coeffs", "ajet_lad_fallback_fraction",
           "ajet_trainer_create", "ajet_trainer_load", "ajet_trainer_run", "ajet_trainer_forward",
           "ajet_trainer_grads", "ajet_trainer_get", "ajet_trainer_destroy", "ajet_trainer_last_error",
           "ajet_set_probes", "ajet_probe", "ajet_probe_at"]
PROBE_AT_MAX = 4096  # AJET_PROBE_AT_MAX


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


class AjetError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run `python -m axisymmetricjet_amd.build`")
        L = C.CDLL(LIB)
        dp = np.ctypeslib.ndpointer(np.float64, flags="F_CONTIGUOUS")
        vp = C.c_void_p
        L.ajet_params_default.argtypes = [C.POINTER(Params), C.c_int32, C.c_int32]
        L.ajet_params_default.restype = None
        L.ajet_build_patches.argtypes = [C.c_int32, C.c_int32, C.POINTER(BcLine), C.c_int32, C.POINTER(Patch), C.c_int32]
        L.ajet_create.argtypes = [C.POINTER(Params), dp, dp, vp, C.POINTER(BcLine), C.c_int32, C.POINTER(vp)]
        L.ajet_destroy.argtypes = [vp]
        L.ajet_last_error.argtypes = [vp]
        L.ajet_last_error.restype = C.c_char_p
        L.ajet_set_state.argtypes = [vp, dp, vp]
        L.ajet_get_state.argtypes = [vp, dp]
        L.ajet_compute_dt.argtypes = [vp, vp]
        L.ajet_rk4_step.argtypes = [vp]
        L.ajet_rk4_stage.argtypes = [vp, C.c_int32]
        L.ajet_run.argtypes = [vp, C.c_int32]
        L.ajet_get_dt_log.argtypes = [vp, vp, C.c_int32, vp]
        L.ajet_rhs.argtypes = [vp, dp, dp]
        L.ajet_lad_coeffs.argtypes = [vp, dp, dp, dp, dp]
        L.ajet_set_forcing.argtypes = [vp, C.POINTER(ForcingC)]
        L.ajet_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
        L.ajet_batch_run.argtypes = [vp, C.c_int32]
        L.ajet_batch_destroy.argtypes = [vp]
        L.ajet_batch_last_error.argtypes = [vp]
        L.ajet_batch_last_error.restype = C.c_char_p
        L.ajet_set_learned_lad.argtypes = [vp, C.POINTER(MlpC), C.c_double, C.c_double, C.c_int32]
        L.ajet_lad_features.argtypes = [vp, dp, vp, C.c_int32, vp]
        L.ajet_lad_learned_coeffs.argtypes = [vp, dp, dp, dp, dp]
        L.ajet_lad_fallback_fraction.argtypes = [vp, C.POINTER(C.c_double)]
        L.ajet_trainer_create.argtypes = [C.c_int32, C.POINTER(MlpC), C.POINTER(vp)]
        L.ajet_trainer_load.argtypes = [vp, vp, vp, C.c_int64]
        L.ajet_trainer_run.argtypes = [vp, C.c_int32, C.c_float, C.c_uint64, vp]
        L.ajet_trainer_forward.argtypes = [vp, vp, C.c_int32, vp]
        L.ajet_trainer_grads.argtypes = [vp, vp, vp, C.c_int32, C.POINTER(MlpC), C.POINTER(C.c_double)]
        L.ajet_trainer_get.argtypes = [vp, C.POINTER(MlpC)]
        L.ajet_trainer_destroy.argtypes = [vp]
        L.ajet_trainer_last_error.argtypes = [vp]
        L.ajet_trainer_last_error.restype = C.c_char_p
        L.ajet_set_probes.argtypes = [vp, vp, C.c_int32]
        L.ajet_probe.argtypes = [vp, vp]
        L.ajet_probe_at.argtypes = [vp, vp, C.c_int32, vp]
        L.ajet_stream.argtypes = [vp]
        L.ajet_stream.restype = vp
        L.ajet_version.restype = C.c_char_p
        _lib = L
    return _lib


class ForcingC(C.Structure):
    _fields_ = [("F", C.c_double * 25)] + [(n, C.c_double) for n in
                                           ("z0", "r0", "lz", "lr", "amp", "strouhal", "uj", "t0")]


NIN, NHID, NOUT = 104, 128, 2
MLP_FIELDS = [("w1", (NHID, NIN)), ("b1", (NHID,)), ("w2", (NHID, NHID)), ("b2", (NHID,)),
              ("w3", (NHID, NHID)), ("b3", (NHID,)), ("w4", (NOUT, NHID)), ("b4", (NOUT,)),
              ("shift", (NIN,)), ("scale", (NIN,)), ("out_scale", (NOUT,))]


class MlpC(C.Structure):
    _fields_ = [(n, C.c_float * int(np.prod(shp))) for n, shp in MLP_FIELDS]


def mlp_to_c(model) -> MlpC:
    """A model dict (w1 (128, 104) .. out_scale (2,), float32) as the C struct ajet_mlp."""
    m = MlpC()
    for n, shp in MLP_FIELDS:
        a = np.ascontiguousarray(np.asarray(model[n], dtype=np.float32).reshape(shp)).ravel()
        C.memmove(getattr(m, n), a.ctypes.data, a.nbytes)
    return m


def mlp_from_c(m: MlpC) -> dict:
    return {n: np.frombuffer(getattr(m, n), dtype=np.float32).reshape(shp).copy() for n, shp in MLP_FIELDS}


def default_params(nz, nr, **kw) -> Params:
    p = Params()
    lib().ajet_params_default(C.byref(p), nz, nr)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def bc_array(lines):
    arr = (BcLine * max(1, len(lines)))()
    for k, l in enumerate(lines):
        arr[k] = BcLine(*[int(x) for x in l[:7]])
    return arr, len(lines)


def build_patches(nz, nr, lines):
    arr, n = bc_array(lines)
    out = (Patch * 64)()
    m = lib().ajet_build_patches(nz, nr, arr, n, out, 64)
    if m < 0:
        raise AjetError(m, "ajet_build_patches")
    return [(p.i_start, p.i_end, p.j_start, p.j_end, p.bc, p.bc_dir) for p in out[:m]]


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


class Context:
    """One device context: the time-step state of one jet run on the current GPU."""

    def __init__(self, case=None, *, params: Params | None = None, z=None, r=None, iblank=None, bc_lines=None,
                 forcing=None, **pkw):
        L = lib()
        if case is not None:
            z, r, iblank, bc_lines = case.z, case.r, case.iblank, case.bc_lines
            pkw.setdefault("Re", case.re)
            pkw.setdefault("Pr", case.pr)
        z, r = F(z), F(r)
        nz, nr = z.shape
        self.nz, self.nr = nz, nr
        if forcing is not None:
            pkw.setdefault("forcing", int(forcing.mode))
        self.params = params or default_params(nz, nr, **pkw)
        self._ib = None if iblank is None else np.asfortranarray(iblank, dtype=np.int32)
        bcs, n = bc_array(bc_lines or [])
        h = C.c_void_p()
        rc = L.ajet_create(C.byref(self.params), z, r, None if self._ib is None else self._ib.ctypes.data,
                           bcs, n, C.byref(h))
        self.h = h
        if rc != AJET_OK:
            msg = L.ajet_last_error(h).decode() if h.value else ""
            if h.value:
                L.ajet_destroy(h)
            self.h = None
            raise AjetError(rc, msg)
        if forcing is not None:
            self.set_forcing(forcing)
        if case is not None:
            self.set_state(case.q0, case.target)

    def _check(self, rc, what):
        if rc < 0 or rc > 0:
            raise AjetError(rc, f"{what}: {lib().ajet_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            lib().ajet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, q, target=None):
        # the converted target stays referenced until the call returns
        t = None if target is None else F(target)
        self._check(lib().ajet_set_state(self.h, F(q), None if t is None else t.ctypes.data), "ajet_set_state")

    def state(self):
        out = np.zeros((self.nz, self.nr, 5), order="F")
        self._check(lib().ajet_get_state(self.h, out), "ajet_get_state")
        return out

    def compute_dt(self, sync=True):
        d = C.c_double()
        self._check(lib().ajet_compute_dt(self.h, C.byref(d) if sync else None), "ajet_compute_dt")
        return d.value if sync else None

    def rk4_step(self):
        self._check(lib().ajet_rk4_step(self.h), "ajet_rk4_step")

    def rk4_stage(self, stage):
        self._check(lib().ajet_rk4_stage(self.h, int(stage)), "ajet_rk4_stage")

    def run(self, nsteps):
        self._check(lib().ajet_run(self.h, int(nsteps)), "ajet_run")

    def dt_log(self, n=None):
        t = C.c_double()
        cnt = lib().ajet_get_dt_log(self.h, None, 0, C.byref(t))
        if cnt < 0:
            self._check(cnt, "ajet_get_dt_log")
        n = cnt if n is None else min(n, cnt)
        out = np.zeros(max(n, 1))
        lib().ajet_get_dt_log(self.h, out.ctypes.data, n, None)
        return out[:n], t.value

    def rhs(self, q):
        out = np.zeros((self.nz, self.nr, 5), order="F")
        self._check(lib().ajet_rhs(self.h, F(q), out), "ajet_rhs")
        return out

    def set_forcing(self, f):
        fc = ForcingC()
        for k, v in enumerate(np.asarray(f.F, dtype=np.float64).ravel()):
            fc.F[k] = v
        for n in ("z0", "r0", "lz", "lr", "amp", "strouhal", "uj", "t0"):
            setattr(fc, n, float(getattr(f, n)))
        self._check(lib().ajet_set_forcing(self.h, C.byref(fc)), "ajet_set_forcing")

    def lad_coeffs(self, q):
        """mu*, beta*, kappa* of ComputeArtTransCoeffs for the state q, each (nz, nr)."""
        outs = [np.zeros((self.nz, self.nr), order="F") for _ in range(3)]
        self._check(lib().ajet_lad_coeffs(self.h, F(q), *outs), "ajet_lad_coeffs")
        return outs

    def set_learned_lad(self, model, sensor_threshold: float, ramp_min: float = 0.0, refresh: int = 1):
        """Learned LAD mode (model None: the reference's LAD again)."""
        if model is None:
            self._check(lib().ajet_set_learned_lad(self.h, None, 0.0, 0.0, 1), "ajet_set_learned_lad")
            return
        self._mlp = mlp_to_c(model)
        self._check(lib().ajet_set_learned_lad(self.h, C.byref(self._mlp), float(sensor_threshold), float(ramp_min),
                                               int(refresh)), "ajet_set_learned_lad")

    def lad_features(self, q, points):
        """Raw features (n, 104) float32 of the points (n, 2) = (i, j), 0-based."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 2))
        out = np.zeros((len(pts), NIN), dtype=np.float32)
        self._check(lib().ajet_lad_features(self.h, F(q), pts.ctypes.data, len(pts), out.ctypes.data),
                    "ajet_lad_features")
        return out

    def learned_coeffs(self, q):
        outs = [np.zeros((self.nz, self.nr), order="F") for _ in range(3)]
        self._check(lib().ajet_lad_learned_coeffs(self.h, F(q), *outs), "ajet_lad_learned_coeffs")
        return outs

    def fallback_fraction(self):
        d = C.c_double()
        self._check(lib().ajet_lad_fallback_fraction(self.h, C.byref(d)), "ajet_lad_fallback_fraction")
        return d.value

    def set_probes(self, points):
        self._probe_pts = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 2))
        self._check(lib().ajet_set_probes(self.h, self._probe_pts.ctypes.data, len(self._probe_pts)),
                    "ajet_set_probes")

    def probe(self):
        """(n, 5) primitives at the probe points of the current state."""
        out = np.zeros((len(self._probe_pts), 5))
        self._check(lib().ajet_probe(self.h, out.ctypes.data), "ajet_probe")
        return out

    def probe_at(self, points):
        """(n, 5) primitives at the points (n, 2) = (i, j), 0-based, of the current state, through
        the library's fixed scratch area (ajet_probe_at: no device allocation, n <= PROBE_AT_MAX)."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 2))
        out = np.zeros((len(pts), 5))
        self._check(lib().ajet_probe_at(self.h, pts.ctypes.data, len(pts), out.ctypes.data), "ajet_probe_at")
        return out

    def stream(self):
        return lib().ajet_stream(self.h)


class Batch:
    """Several contexts advanced together (ajet_batch_*): one launch per kernel per RK stage over
    all runs, later steps replayed from a captured HIP graph."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
        h = C.c_void_p()
        rc = lib().ajet_batch_create(arr, len(self.contexts), C.byref(h))
        self.h = h
        if rc != AJET_OK:
            msg = lib().ajet_batch_last_error(h).decode() if h.value else ""
            if h.value:
                lib().ajet_batch_destroy(h)
            self.h = None
            raise AjetError(rc, msg)

    def run(self, nsteps):
        rc = lib().ajet_batch_run(self.h, int(nsteps))
        if rc != AJET_OK:
            raise AjetError(rc, lib().ajet_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().ajet_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



class Trainer:
    """GPU training of the learned LAD network (ajet_trainer_*): the library's own BF16 MFMA
    forward and backward kernels and Adam."""

    def __init__(self, model, batch=4096):
        self.batch = int(batch)
        self._init = mlp_to_c(model)
        h = C.c_void_p()
        rc = lib().ajet_trainer_create(self.batch, C.byref(self._init), C.byref(h))
        self.h = h
        if rc != AJET_OK:
            msg = lib().ajet_trainer_last_error(h).decode() if h.value else ""
            if h.value:
                lib().ajet_trainer_destroy(h)
            self.h = None
            raise AjetError(rc, msg)

    def _check(self, rc, what):
        if rc != AJET_OK:
            raise AjetError(rc, f"{what}: {lib().ajet_trainer_last_error(self.h).decode()}")

    def load(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert x.shape[1] == NIN and t.shape == (x.shape[0], NOUT)
        self._check(lib().ajet_trainer_load(self.h, x.ctypes.data, t.ctypes.data, x.shape[0]), "ajet_trainer_load")

    def run(self, steps, lr, seed=0):
        losses = np.zeros(max(int(steps), 1))
        self._check(lib().ajet_trainer_run(self.h, int(steps), float(lr), int(seed), losses.ctypes.data),
                    "ajet_trainer_run")
        return losses[:steps]

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], NOUT), dtype=np.float32)
        self._check(lib().ajet_trainer_forward(self.h, x.ctypes.data, x.shape[0], y.ctypes.data),
                    "ajet_trainer_forward")
        return y

    def grads(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = MlpC()
        loss = C.c_double()
        self._check(lib().ajet_trainer_grads(self.h, x.ctypes.data, t.ctypes.data, x.shape[0], C.byref(g),
                                             C.byref(loss)), "ajet_trainer_grads")
        return loss.value, mlp_from_c(g)

    def model(self):
        m = MlpC()
        self._check(lib().ajet_trainer_get(self.h, C.byref(m)), "ajet_trainer_get")
        return mlp_from_c(m)

    def close(self):
        if getattr(self, "h", None):
            lib().ajet_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
