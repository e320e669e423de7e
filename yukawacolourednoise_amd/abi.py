"""ctypes binding of libycn.so (the C-ABI declared in include/ycn.h).

There is no CPU fallback: if ``libycn.so`` has not been built, importing this module
raises ``ImportError``; creating a context without a HIP device raises
``YcnError`` with code ``YCN_ERR_NO_DEVICE``.

Array conventions follow the reference (code/src/Lattice.cu:16-30,
code/src/Spinor.cuh:7-16): scalar fields are ``(nReplicas, Nt, Nx)`` float64 (or
``(Nt, Nx)`` for one replica); spinors are ``(V, 4)`` complex128, AoS
``[f1s1, f1s2, f2s1, f2s2]`` per site.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libycn.so")

YCN_OK = 0
YCN_ERR_INVALID = -1
YCN_ERR_HIP = -2
YCN_ERR_NO_DEVICE = -3
YCN_ERR_UNSUPPORTED = -4
YCN_ERR_NOT_CONVERGED = -5
YCN_NORMAL, YCN_DAGGER = 0, 1
YCN_TOL_ABS_RECURSIVE, YCN_TOL_REL_TRUE = 0, 1
YCN_CG_FP64, YCN_CG_MIXED = 0, 1
YCN_PRECOND_NONE, YCN_PRECOND_CNN = 0, 1

# every entry point declared in include/ycn.h
EXPORTED = [
    "ycn_default_params", "ycn_create", "ycn_destroy", "ycn_last_error", "ycn_abi_version",
    "ycn_tuning_report",
    "ycn_set_replicas", "ycn_set_eps_bar", "ycn_set_field", "ycn_get_field", "ycn_hot_start",
    "ycn_coloured_noise", "ycn_coloured_filter", "ycn_coloured_noise_noalloc", "ycn_dirac_apply", "ycn_cg_solve",
    "ycn_fermion_force", "ycn_langevin_step", "ycn_langevin_run", "ycn_host_syncs", "ycn_measure", "ycn_sync",
    "ycn_profile_begin", "ycn_profile_end",
    "ycn_precond_set_weights", "ycn_precond_get_weights", "ycn_precond_apply", "ycn_precond_loss_grad",
    "ycn_precond_train",
]
YCN_CNN_NWEIGHTS = 4896


class Params(C.Structure):
    _fields_ = [
        ("Nt", C.c_int), ("Nx", C.c_int), ("nReplicas", C.c_int), ("nNoise", C.c_int),
        ("m2", C.c_double), ("lam", C.c_double), ("g", C.c_double), ("M", C.c_double),
        ("r", C.c_double), ("cutFraction", C.c_double), ("driftMode", C.c_int),
        ("epsBar", C.c_double), ("adaptive", C.c_int), ("seed", C.c_uint64),
        ("cgTol", C.c_double), ("cgTolMode", C.c_int), ("cgMaxIter", C.c_int),
        ("cgPrecision", C.c_int), ("reliableEvery", C.c_int), ("precond", C.c_int),
        ("device", C.c_int), ("evenOdd", C.c_int),
    ]


class CgStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("restarts", C.c_int), ("converged", C.c_int),
                ("systems", C.c_int), ("rr", C.c_double), ("true_rel_resid", C.c_double),
                ("iters_mean", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Observables(C.Structure):
    _fields_ = [("magnetisation", C.c_double), ("phi2", C.c_double), ("trace", C.c_double),
                ("eps", C.c_double), ("langevin_time", C.c_double), ("clamp_events", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrainStats(C.Structure):
    _fields_ = [("pairs", C.c_int), ("loss_first", C.c_double), ("loss_last", C.c_double),
                ("eps_last", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class YcnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ycn error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(there is no CPU fallback for the HIP path)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.ycn_default_params.argtypes = [C.POINTER(Params)]
    L.ycn_default_params.restype = None
    L.ycn_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.ycn_destroy.argtypes = [vp]
    L.ycn_destroy.restype = None
    L.ycn_last_error.argtypes = [vp]
    L.ycn_last_error.restype = C.c_char_p
    L.ycn_tuning_report.argtypes = [vp]
    L.ycn_tuning_report.restype = C.c_char_p
    L.ycn_set_replicas.argtypes = [vp, dp, C.POINTER(C.c_uint64)]
    L.ycn_set_eps_bar.argtypes = [vp, C.c_double]
    L.ycn_set_field.argtypes = [vp, dp]
    L.ycn_get_field.argtypes = [vp, dp]
    L.ycn_hot_start.argtypes = [vp, C.c_uint64]
    L.ycn_coloured_noise.argtypes = [vp, C.c_uint64, dp]
    L.ycn_coloured_filter.argtypes = [vp, dp, dp]
    L.ycn_coloured_noise_noalloc.argtypes = [C.POINTER(Params), C.POINTER(C.c_uint64), C.c_uint64, vp, vp]
    L.ycn_dirac_apply.argtypes = [vp, dp, dp, dp, C.c_int]
    L.ycn_cg_solve.argtypes = [vp, dp, dp, dp, C.c_int, C.POINTER(CgStats)]
    L.ycn_fermion_force.argtypes = [vp, C.c_uint64, dp, C.POINTER(CgStats)]
    L.ycn_langevin_step.argtypes = [vp, C.c_uint64, dp, C.POINTER(CgStats)]
    L.ycn_langevin_run.argtypes = [vp, C.c_uint64, C.c_int, C.POINTER(CgStats)]
    L.ycn_host_syncs.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.ycn_measure.argtypes = [vp, C.c_int, C.c_uint64, C.POINTER(Observables), dp, dp]
    L.ycn_sync.argtypes = [vp]
    L.ycn_profile_begin.argtypes = [vp, C.c_char_p]
    L.ycn_profile_end.argtypes = [vp, dp, C.POINTER(C.c_int64), dp]
    fp = C.POINTER(C.c_float)
    L.ycn_precond_set_weights.argtypes = [vp, fp]
    L.ycn_precond_get_weights.argtypes = [vp, fp]
    L.ycn_precond_apply.argtypes = [vp, dp, dp]
    L.ycn_precond_loss_grad.argtypes = [vp, dp, dp, dp, dp, fp]
    L.ycn_precond_train.argtypes = [vp, C.c_uint64, C.c_int, C.c_int, C.c_double, C.POINTER(TrainStats)]
    return L


lib = _load()


def default_params(**kw) -> Params:
    p = Params()
    lib.ycn_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "lambda_":
            k = "lam"
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def coloured_noise_noalloc(params: Params, step: int, noise_dev_ptr: int, stream_ptr: int = 0, seeds=None):
    """ycn_coloured_noise_noalloc: the coloured noise of ``params`` written into a device
    buffer (``noise_dev_ptr``, nReplicas*Nt*Nx doubles) on a HIP stream, with no context
    and no allocation.  Returns the ABI code; raises nothing (callers check the code and
    ``lib.ycn_last_error(None)``)."""
    s = None
    if seeds is not None:
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        assert s.size == params.nReplicas
    return lib.ycn_coloured_noise_noalloc(C.byref(params),
                                          None if s is None else s.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          step, C.c_void_p(noise_dev_ptr or None), C.c_void_p(stream_ptr or None))


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _spin(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


@dataclass
class StepResult:
    eps: np.ndarray
    stats: dict


class Context:
    """One ycn_ctx: device buffers, stream and the hot-path entry points."""

    def __init__(self, params: Params):
        self.params = params
        h = C.c_void_p()
        rc = lib.ycn_create(C.byref(params), C.byref(h))
        if rc != YCN_OK:
            raise YcnError(rc, "ycn_create failed" + (" (no HIP device)" if rc == YCN_ERR_NO_DEVICE else ""))
        self._h = h
        self.Nt, self.Nx = params.Nt, params.Nx
        self.V = self.Nt * self.Nx
        self.nRep = params.nReplicas

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib.ycn_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, allow_nc=False):
        if rc == YCN_OK or (allow_nc and rc == YCN_ERR_NOT_CONVERGED):
            return rc
        raise YcnError(rc, lib.ycn_last_error(self._h).decode())

    def _field_shape(self):
        return (self.nRep, self.Nt, self.Nx) if self.nRep > 1 else (self.Nt, self.Nx)

    # -- state ----------------------------------------------------------------
    def set_replicas(self, m2=None, seeds=None):
        m = None if m2 is None else _f64(m2)
        s = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint64)
        self._check(lib.ycn_set_replicas(self._h, None if m is None else _ptr(m),
                                         None if s is None else s.ctypes.data_as(C.POINTER(C.c_uint64))))

    def set_eps_bar(self, e):
        self._check(lib.ycn_set_eps_bar(self._h, float(e)))

    def set_field(self, phi):
        phi = _f64(phi).reshape(-1)
        assert phi.size == self.nRep * self.V
        self._check(lib.ycn_set_field(self._h, _ptr(phi)))

    def get_field(self):
        out = np.empty(self.nRep * self.V, np.float64)
        self._check(lib.ycn_get_field(self._h, _ptr(out)))
        return out.reshape(self._field_shape())

    def hot_start(self, init_seed):
        self._check(lib.ycn_hot_start(self._h, init_seed))

    # -- hot path -------------------------------------------------------------
    def coloured_noise(self, step):
        out = np.empty(self.nRep * self.V, np.float64)
        self._check(lib.ycn_coloured_noise(self._h, step, _ptr(out)))
        return out.reshape(self._field_shape())

    def coloured_filter(self, xi):
        xi = _f64(xi).reshape(-1)
        assert xi.size == self.nRep * self.V
        out = np.empty_like(xi)
        self._check(lib.ycn_coloured_filter(self._h, _ptr(xi), _ptr(out)))
        return out.reshape(self._field_shape())

    def dirac_apply(self, psi, phi=None, dagger=False):
        psi = _spin(psi)
        out = np.empty(8 * self.V, np.float64)
        ph = None if phi is None else _f64(phi).reshape(-1)
        self._check(lib.ycn_dirac_apply(self._h, None if ph is None else _ptr(ph), _ptr(psi), _ptr(out),
                                        YCN_DAGGER if dagger else YCN_NORMAL))
        return out.view(np.complex128).reshape(self.V, 4)

    def cg_solve(self, b, phi=None, mtype=YCN_DAGGER):
        b = _spin(b)
        out = np.empty(8 * self.V, np.float64)
        ph = None if phi is None else _f64(phi).reshape(-1)
        st = CgStats()
        self._check(lib.ycn_cg_solve(self._h, None if ph is None else _ptr(ph), _ptr(b), _ptr(out), mtype,
                                     C.byref(st)), allow_nc=True)
        return out.view(np.complex128).reshape(self.V, 4), st.as_dict()

    def fermion_force(self, step):
        out = np.empty(self.nRep * self.V, np.float64)
        st = CgStats()
        self._check(lib.ycn_fermion_force(self._h, step, _ptr(out), C.byref(st)), allow_nc=True)
        return out.reshape(self._field_shape()), st.as_dict()

    def langevin_step(self, step):
        eps = np.empty(self.nRep, np.float64)
        st = CgStats()
        self._check(lib.ycn_langevin_step(self._h, step, _ptr(eps), C.byref(st)), allow_nc=True)
        return StepResult(eps, st.as_dict())

    def langevin_run(self, step0, nsteps):
        st = CgStats()
        self._check(lib.ycn_langevin_run(self._h, step0, nsteps, C.byref(st)), allow_nc=True)
        return st.as_dict()

    def host_syncs(self):
        """(host waits on the context's work since creation, captured steps the host
        finished)"""
        n, h = C.c_int64(), C.c_int64()
        self._check(lib.ycn_host_syncs(self._h, C.byref(n), C.byref(h)))
        return n.value, h.value

    def measure(self, replica=0, step=0, slices=True, corr=True):
        ob = Observables()
        sl = np.empty(self.Nt, np.float64) if slices else None
        cr = np.empty(self.Nt, np.float64) if corr else None
        self._check(lib.ycn_measure(self._h, replica, step, C.byref(ob), None if sl is None else _ptr(sl),
                                    None if cr is None else _ptr(cr)))
        return ob.as_dict(), sl, cr

    # -- learned preconditioner ------------------------------------------------
    def precond_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == YCN_CNN_NWEIGHTS
        self._check(lib.ycn_precond_set_weights(self._h, w.ctypes.data_as(C.POINTER(C.c_float))))

    def precond_get_weights(self):
        w = np.empty(YCN_CNN_NWEIGHTS, np.float32)
        self._check(lib.ycn_precond_get_weights(self._h, w.ctypes.data_as(C.POINTER(C.c_float))))
        return w

    def precond_apply(self, r):
        r = _spin(r)
        out = np.empty(8 * self.V, np.float64)
        self._check(lib.ycn_precond_apply(self._h, _ptr(r), _ptr(out)))
        return out.view(np.complex128).reshape(self.V, 4)

    def precond_loss_grad(self, r, e, phi=None):
        r, e = _spin(r), _spin(e)
        ph = None if phi is None else _f64(phi).reshape(-1)
        loss = C.c_double()
        g = np.empty(YCN_CNN_NWEIGHTS, np.float32)
        self._check(lib.ycn_precond_loss_grad(self._h, None if ph is None else _ptr(ph), _ptr(r), _ptr(e),
                                              C.byref(loss), g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def precond_train(self, step0, nsteps, pairs_per_solve=4, lr=1e-3):
        st = TrainStats()
        self._check(lib.ycn_precond_train(self._h, step0, nsteps, pairs_per_solve, lr, C.byref(st)))
        return st.as_dict()

    def tuning_report(self):
        return lib.ycn_tuning_report(self._h).decode()

    def sync(self):
        self._check(lib.ycn_sync(self._h))

    def profile_begin(self, kernel):
        self._check(lib.ycn_profile_begin(self._h, kernel.encode()))

    def profile_end(self):
        ms = C.c_double()
        n = C.c_int64()
        b = C.c_double()
        self._check(lib.ycn_profile_end(self._h, C.byref(ms), C.byref(n), C.byref(b)))
        return ms.value, n.value, b.value
