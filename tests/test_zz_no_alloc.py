"""ycn_coloured_noise_noalloc, the allocation-free variant of ycn_coloured_noise, called
through the C-ABI on device buffers: two calls in a row, the second on a smaller lattice,
each equal bit for bit to the field the context entry point writes for the same
parameters, seeds and step (both run the same kernels on the same twiddles and mask)."""
import ctypes as C

import numpy as np
import pytest

import yukawacolourednoise_amd as Y
from yukawacolourednoise_amd import abi

pytestmark = pytest.mark.gpu

# first call at the scratch array's full capacity (one 1024^2 lattice), then a smaller
# batched lattice with odd FFT factors (48 = 16 * 3, 40 = 8 * 5) and per-replica seeds
CASES = [
    dict(Nt=1024, Nx=1024, nReplicas=1, cutFraction=0.5, seed=1234, step=3, seeds=None),
    dict(Nt=48, Nx=40, nReplicas=3, cutFraction=0.25, seed=7, step=11, seeds=[5, 17, 2 ** 40 + 3]),
]


def _context_noise(case):
    p = Y.default_params(Nt=case["Nt"], Nx=case["Nx"], nReplicas=case["nReplicas"],
                         cutFraction=case["cutFraction"], seed=case["seed"])
    with Y.Context(p) as c:
        if case["seeds"] is not None:
            c.set_replicas(seeds=case["seeds"])
        return c.coloured_noise(case["step"]).reshape(-1)


def _hip():
    """The HIP runtime libycn.so is linked against (device buffers and a stream for the
    caller's side of the test; torch's bundled runtime cannot share a process with it)."""
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def test_noalloc_noise_matches_context_entry_point(gpu):
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    for case in CASES:
        print(f"noalloc noise: {case['nReplicas']} x {case['Nt']}x{case['Nx']}, s={case['cutFraction']}, "
              f"step {case['step']}", flush=True)
        ref = _context_noise(case)
        p = Y.default_params(Nt=case["Nt"], Nx=case["Nx"], nReplicas=case["nReplicas"],
                             cutFraction=case["cutFraction"], seed=case["seed"])
        n = case["nReplicas"] * case["Nt"] * case["Nx"]
        dev = C.c_void_p()
        assert hip.hipMalloc(C.byref(dev), 8 * n) == 0
        try:
            assert hip.hipMemset(dev, 0xFF, 8 * n) == 0   # NaN everywhere until written
            rc = abi.coloured_noise_noalloc(p, case["step"], dev.value, stream.value, seeds=case["seeds"])
            assert rc == abi.YCN_OK, Y.lib.ycn_last_error(None).decode()
            assert hip.hipStreamSynchronize(stream) == 0
            got = np.empty(n, np.float64)
            assert hip.hipMemcpy(got.ctypes.data, dev, 8 * n, 2) == 0   # device to host
        finally:
            hip.hipFree(dev)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
    hip.hipStreamDestroy(stream)
