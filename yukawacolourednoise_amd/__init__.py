"""MI355X-native Langevin update of the 2D lattice Yukawa model with coloured noise.

The hot path lives in ``libycn.so`` (hand-written gfx950 HIP kernels behind the C-ABI of
``include/ycn.h``); :mod:`yukawacolourednoise_amd.abi` is its ctypes binding.
"""
from .abi import (Context, Params, CgStats, Observables, YcnError, default_params, lib,  # noqa: F401
                  YCN_NORMAL, YCN_DAGGER, YCN_TOL_ABS_RECURSIVE, YCN_TOL_REL_TRUE,
                  YCN_CG_FP64, YCN_CG_MIXED, YCN_PRECOND_NONE, YCN_PRECOND_CNN,
                  YCN_ERR_NO_DEVICE, YCN_ERR_NOT_CONVERGED, EXPORTED)

__all__ = ["Context", "Params", "CgStats", "Observables", "YcnError", "default_params", "lib", "EXPORTED"]
