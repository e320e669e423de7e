"""C-ABI checks that need no GPU: the library loads, exports every entry point that
include/ycn.h declares, validates parameters and refuses to run without a device
(there is no CPU fallback)."""
import ctypes as C
import os
import re

import pytest

import yukawacolourednoise_amd as Y
from yukawacolourednoise_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "ycn.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ycn_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert header_symbols() == sorted(abi.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_abi_version_and_defaults():
    assert Y.lib.ycn_abi_version() == 2
    p = Y.default_params()
    # reference defaults: params.h:13-15, input.toml
    assert (p.Nt, p.Nx, p.cgTol, p.cgMaxIter, p.cgTolMode) == (16, 16, 1e-8, 1000, 0)
    assert (p.r, p.M, p.epsBar, p.adaptive, p.seed, p.nReplicas, p.nNoise) == (1.0, 1.0, 0.02, 1, 1234, 1, 1)


@pytest.mark.parametrize("field,value", [("Nt", 1), ("Nx", 0), ("nReplicas", 0), ("nNoise", 0),
                                         ("cgTolMode", 7), ("driftMode", 3), ("cgTol", -1.0)])
def test_invalid_params_rejected(field, value):
    p = Y.default_params()
    setattr(p, field, value)
    h = C.c_void_p()
    assert Y.lib.ycn_create(C.byref(p), C.byref(h)) == abi.YCN_ERR_INVALID
    assert not h.value


def test_null_arguments_rejected():
    assert Y.lib.ycn_create(None, None) == abi.YCN_ERR_INVALID
    assert Y.lib.ycn_sync(None) == abi.YCN_ERR_INVALID
    assert Y.lib.ycn_langevin_step(None, 0, None, None) == abi.YCN_ERR_INVALID


def test_no_cpu_fallback_without_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a HIP device node is present")
    with pytest.raises(Y.YcnError) as e:
        Y.Context(Y.default_params())
    assert e.value.code == abi.YCN_ERR_NO_DEVICE


@pytest.mark.parametrize("field,value", [("Nt", 2049), ("Nx", 4096), ("nReplicas", 65), ("nReplicas", 0), ("Nt", 1)])
def test_noalloc_noise_rejects_sizes_beyond_its_scratch(field, value):
    # validated on the host before any device call: no GPU needed
    p = Y.default_params(Nt=64, Nx=64)
    setattr(p, field, value)
    assert abi.coloured_noise_noalloc(p, 0, 0x1000, 0) == abi.YCN_ERR_INVALID
    assert b"ycn_coloured_noise_noalloc" in Y.lib.ycn_last_error(None)


def test_noalloc_noise_rejects_a_lattice_beyond_one_1024x1024_spectrum():
    # 2 replicas of 1024^2 (or one 2048x1024 lattice) need twice the half-spectrum capacity
    for kw in (dict(Nt=1024, Nx=1024, nReplicas=2), dict(Nt=2048, Nx=1024, nReplicas=1)):
        p = Y.default_params(**kw)
        assert abi.coloured_noise_noalloc(p, 0, 0x1000, 0) == abi.YCN_ERR_INVALID
        assert b"scratch capacity" in Y.lib.ycn_last_error(None)
    assert Y.lib.ycn_coloured_noise_noalloc(None, None, 0, C.c_void_p(0x1000), None) == abi.YCN_ERR_INVALID
