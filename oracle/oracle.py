"""ctypes bindings for the CPU FP64 oracle (``oracle/liborc.so``).

Test infrastructure: the reference the HIP library is tested against.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
module; the product package never does.

Arrays use the reference layouts: scalar fields are ``(Nt, Nx)`` float64 row-major,
spinors are ``(V, 4)`` complex128 (AoS ``[f1s1, f1s2, f2s1, f2s2]`` per site).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("Nt", C.c_int), ("Nx", C.c_int),
        ("m2", C.c_double), ("lam", C.c_double), ("g", C.c_double), ("M", C.c_double),
        ("r", C.c_double), ("cutFraction", C.c_double),
        ("driftMode", C.c_int), ("epsBar", C.c_double), ("adaptive", C.c_int),
        ("seed", C.c_uint64),
        ("cgTol", C.c_double), ("cgTolMode", C.c_int), ("cgMaxIter", C.c_int),
        ("nNoise", C.c_int), ("evenOdd", C.c_int),
    ]


class OrcCgStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("restarts", C.c_int), ("converged", C.c_int),
                ("rr", C.c_double), ("true_rel_resid", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        pp = C.POINTER(OrcParams)
        sp = C.POINTER(OrcCgStats)
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_int64, dp]
        L.orc_hot_start.argtypes = [C.c_uint64, C.c_int64, dp]
        L.orc_fft_c2c.argtypes = [C.c_int, dp, dp, C.c_int]
        L.orc_mask.argtypes = [C.c_int, C.c_int, C.c_double, dp]
        L.orc_coloured_filter.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_coloured_noise.argtypes = [pp, C.c_uint64, dp]
        L.orc_dirac.argtypes = [pp, dp, dp, dp, C.c_int]
        L.orc_normal_op.argtypes = [pp, dp, dp, dp, C.c_int]
        L.orc_cg.argtypes = [pp, dp, dp, dp, C.c_int, C.c_int, sp]
        L.orc_cg.restype = C.c_int
        L.orc_fermion_drift.argtypes = [pp, dp, C.c_uint64, dp, sp]
        L.orc_fermion_drift_eta.argtypes = [pp, dp, dp, dp, sp]
        L.orc_fermion_noise.argtypes = [C.c_uint64, C.c_int, C.c_uint64, C.c_int64, dp]
        L.orc_laplace_csr.argtypes = [C.c_int, C.c_int, np.ctypeslib.ndpointer(np.int32), np.ctypeslib.ndpointer(np.int32), dp]
        L.orc_bosonic_update.argtypes = [pp, dp, dp, dp, C.POINTER(C.c_double)]
        L.orc_bosonic_update.restype = C.c_double
        L.orc_langevin_step.argtypes = [pp, dp, C.c_uint64, C.POINTER(C.c_double), sp]
        L.orc_langevin_step.restype = C.c_int
        L.orc_time_slices.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_correlator.argtypes = [pp, dp, dp, sp]
        L.orc_correlator.restype = C.c_int
        L.orc_sum.argtypes = [C.c_int64, dp]
        L.orc_sum.restype = C.c_double
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.orc_cnn_forward.argtypes = [C.c_int, C.c_int, fp, C.c_int, C.c_double, C.c_double, dp, dp, dp]
        L.orc_cnn_loss_grad.argtypes = [pp, dp, fp, C.c_double, dp, dp, dp]
        L.orc_cnn_loss_grad.restype = C.c_double
        L.orc_bf16_round.argtypes = [C.c_double]
        L.orc_bf16_round.restype = C.c_double
        L.orc_eo_schur.argtypes = [pp, dp, dp, dp, C.c_int]
        L.orc_eo_rhs.argtypes = [pp, dp, dp, dp]
        L.orc_eo_reconstruct.argtypes = [pp, dp, dp, dp, dp]
        L.orc_eo_cg.argtypes = [pp, dp, dp, dp, C.c_int, sp]
        L.orc_eo_cg.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        _lib = L
    return _lib


def params(Nt=16, Nx=16, m2=-0.2, lam=0.5, g=0.3, M=0.3, r=1.0, cutFraction=0.5, driftMode=0,
           epsBar=0.01, adaptive=0, seed=1234, cgTol=1e-10, cgTolMode=1, cgMaxIter=10000,
           nNoise=1, evenOdd=0) -> OrcParams:
    return OrcParams(Nt, Nx, m2, lam, g, M, r, cutFraction, driftMode, epsBar, adaptive, seed,
                     cgTol, cgTolMode, cgMaxIter, nNoise, evenOdd)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _spin_to_f64(s):
    s = np.ascontiguousarray(s, dtype=np.complex128)
    return s.view(np.float64).reshape(-1)


def _f64_to_spin(a, V):
    return a.view(np.complex128).reshape(V, 4)


def set_threads(n: int) -> int:
    return lib().orc_set_threads(n)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def normals(seed, stream, step, n):
    out = np.empty(n, np.float64)
    lib().orc_normals(seed, stream, step, n, out)
    return out


def hot_start(init_seed, Nt, Nx):
    out = np.empty(Nt * Nx, np.float64)
    lib().orc_hot_start(init_seed, Nt * Nx, out)
    return out.reshape(Nt, Nx)


def fft(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_fft_c2c(len(x), x.view(np.float64), out.view(np.float64), sign)
    return out


def mask(Nt, Nx, s):
    out = np.empty(Nt * (Nx // 2 + 1), np.float64)
    lib().orc_mask(Nt, Nx, s, out)
    return out.reshape(Nt, Nx // 2 + 1)


def coloured_filter(xi, s):
    xi = _f64(xi)
    Nt, Nx = xi.shape
    out = np.empty_like(xi)
    lib().orc_coloured_filter(Nt, Nx, s, xi, out)
    return out


def coloured_noise(p: OrcParams, step):
    out = np.empty((p.Nt, p.Nx), np.float64)
    lib().orc_coloured_noise(C.byref(p), step, out)
    return out


def dirac(p: OrcParams, phi, psi, dagger=False):
    V = p.Nt * p.Nx
    out = np.empty(8 * V, np.float64)
    lib().orc_dirac(C.byref(p), _f64(phi).reshape(-1), _spin_to_f64(psi), out, int(dagger))
    return _f64_to_spin(out, V)


def normal_op(p: OrcParams, phi, psi, mtype=1):
    V = p.Nt * p.Nx
    out = np.empty(8 * V, np.float64)
    lib().orc_normal_op(C.byref(p), _f64(phi).reshape(-1), _spin_to_f64(psi), out, mtype)
    return _f64_to_spin(out, V)


def cg(p: OrcParams, phi, b, mtype=1, max_iter_override=0):
    V = p.Nt * p.Nx
    x = np.empty(8 * V, np.float64)
    st = OrcCgStats()
    lib().orc_cg(C.byref(p), _f64(phi).reshape(-1), _spin_to_f64(b), x, mtype, max_iter_override, C.byref(st))
    return _f64_to_spin(x, V), st.as_dict()


def fermion_noise(seed, j, step, V):
    out = np.empty(8 * V, np.float64)
    lib().orc_fermion_noise(seed, j, step, V, out)
    return _f64_to_spin(out, V)


def fermion_drift(p: OrcParams, phi, step):
    out = np.empty((p.Nt, p.Nx), np.float64)
    st = OrcCgStats()
    lib().orc_fermion_drift(C.byref(p), _f64(phi).reshape(-1), step, out, C.byref(st))
    return out, st.as_dict()


def fermion_drift_eta(p: OrcParams, phi, eta):
    out = np.empty((p.Nt, p.Nx), np.float64)
    st = OrcCgStats()
    e = np.ascontiguousarray(eta, dtype=np.complex128).view(np.float64).reshape(-1)
    lib().orc_fermion_drift_eta(C.byref(p), _f64(phi).reshape(-1), e, out, C.byref(st))
    return out, st.as_dict()


def laplace_csr(Nt, Nx):
    N = Nt * Nx
    I = np.empty(N + 1, np.int32)
    J = np.empty(5 * N, np.int32)
    val = np.empty(5 * N, np.float64)
    lib().orc_laplace_csr(Nt, Nx, I, J, val)
    return I, J, val


def bosonic_update(p: OrcParams, phi, fermion_drift, noise):
    phi = _f64(phi).copy()
    drift = _f64(fermion_drift).copy()
    eps = C.c_double()
    mx = lib().orc_bosonic_update(C.byref(p), phi.reshape(-1), drift.reshape(-1), _f64(noise).reshape(-1), C.byref(eps))
    return phi, drift, eps.value, mx


def langevin_step(p: OrcParams, phi, step):
    phi = _f64(phi).copy()
    eps = C.c_double()
    st = OrcCgStats()
    lib().orc_langevin_step(C.byref(p), phi.reshape(-1), step, C.byref(eps), C.byref(st))
    return phi, eps.value, st.as_dict()


def time_slices(phi):
    phi = _f64(phi)
    out = np.empty(phi.shape[0], np.float64)
    lib().orc_time_slices(phi.shape[0], phi.shape[1], phi, out)
    return out


def correlator(p: OrcParams, phi):
    out = np.empty(p.Nt, np.float64)
    st = OrcCgStats()
    lib().orc_correlator(C.byref(p), _f64(phi).reshape(-1), out, C.byref(st))
    return out, st.as_dict()


def eo_index(Nt, Nx, parity):
    """Full-lattice site indices of the compact half-lattice vector of `parity`, in
    compact order (t-major, x >> 1)."""
    t, x = np.meshgrid(np.arange(Nt), np.arange(Nx), indexing="ij")
    sites = (t * Nx + x).ravel()
    par = ((t + x) & 1).ravel()
    sel = sites[par == parity]
    comp = (t.ravel()[par == parity] * (Nx // 2) + (x.ravel()[par == parity] >> 1))
    out = np.empty_like(sel)
    out[comp] = sel
    return out


def eo_schur(p: OrcParams, phi, in_e, dagger=True):
    Vh = p.Nt * p.Nx // 2
    out = np.empty(8 * Vh, np.float64)
    lib().orc_eo_schur(C.byref(p), _f64(phi).reshape(-1), _spin_to_f64(in_e), out, int(dagger))
    return _f64_to_spin(out, Vh)


def eo_rhs(p: OrcParams, phi, eta):
    Vh = p.Nt * p.Nx // 2
    out = np.empty(8 * Vh, np.float64)
    lib().orc_eo_rhs(C.byref(p), _f64(phi).reshape(-1), _spin_to_f64(eta), out)
    return _f64_to_spin(out, Vh)


def eo_reconstruct(p: OrcParams, phi, eta, chi_e):
    V = p.Nt * p.Nx
    out = np.empty(8 * V, np.float64)
    lib().orc_eo_reconstruct(C.byref(p), _f64(phi).reshape(-1), _spin_to_f64(eta), _spin_to_f64(chi_e), out)
    return _f64_to_spin(out, V)


def eo_cg(p: OrcParams, phi, eta, max_iter_override=0):
    V = p.Nt * p.Nx
    chi = np.empty(8 * V, np.float64)
    st = OrcCgStats()
    lib().orc_eo_cg(C.byref(p), _f64(phi).reshape(-1), _spin_to_f64(eta), chi, max_iter_override, C.byref(st))
    return _f64_to_spin(chi, V), st.as_dict()


CNN_NW = 9 * 8 * 16 + 9 * 16 * 16 + 9 * 16 * 8 + 2 * 9 * 16   # W1, W2, W3, gates G1, G2


def cnn_c0(r=1.0, M=0.3):
    """Weight of the exact skip path z = c0 r + s net(r/s): c0 = 1 / (2r + M)^2."""
    return 1.0 / (2.0 * r + M) ** 2


def cnn_forward(Nt, Nx, W, r, bf16=True, c0=0.0, phi=None, g=0.0):
    """z = c0 r + s net(r/s) for an AoS (V, 4) complex spinor, the hidden layers gated by
    q = g phi (phi None: q = 0, gates 1 + tanh(0) = 1); bf16 emulates the GPU's rounding
    of weights and activations."""
    W = np.ascontiguousarray(W, dtype=np.float32)
    ph = np.zeros(Nt * Nx) if phi is None else _f64(phi).reshape(-1)
    out = np.empty(8 * Nt * Nx, np.float64)
    lib().orc_cnn_forward(Nt, Nx, W, int(bf16), c0, g, ph, _spin_to_f64(r), out)
    return _f64_to_spin(out, Nt * Nx)


def cnn_loss_grad(p: OrcParams, phi, W, r, e, c0=0.0):
    W = np.ascontiguousarray(W, dtype=np.float32)
    g = np.empty(CNN_NW, np.float64)
    L = lib().orc_cnn_loss_grad(C.byref(p), _f64(phi).reshape(-1), W, c0, _spin_to_f64(r), _spin_to_f64(e), g)
    return L, g
