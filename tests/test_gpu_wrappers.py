"""The reference-named C++ wrappers of yukawacolourednoise_amd/host/ycn_backend.hpp
(DiracOP, CGsolver, FermionicDrift, ColouredNoise: Dirac.cuh:21-22, CGsolver.cuh:36,
FermionicDrift.cuh:13-15, colouredNoise.h:28-33), driven by the compiled program
tests/host/wrappers_check (built by __graft_entry__.build()) on a 16x16 GPU context with
the 1024^2 workload's couplings, against the oracle: noise and both Dirac applications
to 1e-12, the CG solution's residual recomputed by the oracle to the solver tolerance,
and the fermion drift (same Philox noise) to the solver's accuracy."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "host", "wrappers_check")


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_reference_named_wrappers(gpu, tmp_path):
    assert os.path.exists(EXE), "tests/host/wrappers_check is built by __graft_entry__.build()"
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip())
    ld = lambda n: np.fromfile(str(tmp_path / n), dtype=np.float64)
    Nt = Nx = 16
    V = Nt * Nx
    p = O.params(Nt=Nt, Nx=Nx, m2=-0.4, lam=1.0, g=0.5, M=0.05, r=1.0, cutFraction=0.5, epsBar=0.005,
                 adaptive=0, seed=1234, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    phi = ld("phi.bin").reshape(Nt, Nx)
    assert np.array_equal(phi, O.hot_start(7, Nt, Nx))
    psi = ld("psi.bin").view(np.complex128).reshape(V, 4)
    assert rel(ld("noise.bin").reshape(Nt, Nx), O.coloured_noise(p, 3)) < 1e-12
    assert rel(ld("Dpsi.bin").view(np.complex128).reshape(V, 4), O.dirac(p, phi, psi, False)) < 1e-12
    assert rel(ld("Ddagpsi.bin").view(np.complex128).reshape(V, 4), O.dirac(p, phi, psi, True)) < 1e-12
    x = ld("x.bin").view(np.complex128).reshape(V, 4)
    res = np.linalg.norm(psi - O.normal_op(p, phi, x, 1)) / np.linalg.norm(psi)
    assert res <= 1.5e-10, res
    drift_ref, st = O.fermion_drift(p, phi, 0)
    assert rel(ld("drift.bin").reshape(Nt, Nx), drift_ref) < 1e-8
