"""rocFFT cross-check of the coloured-noise filter (colouredNoise.cpp:155-178: forward FFT
of the white field, the momentum cutoff |p| < Lambda with the 1 / V normalisation in the
mask, inverse FFT).

tests/host/rocfft_check (built by __graft_entry__.build()) runs the same filter through
rocFFT's double-precision real-to-complex and complex-to-real transforms, with the
oracle's half-spectrum mask.  The library's LDS Stockham transforms (noise.hip) must give
the same field to 1e-12 relative on power-of-two, mixed-radix and small lattices, and so
must the oracle's own transform, which pins the check itself."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import mk

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "host", "rocfft_check")


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("Nt,Nx", [(1024, 1024), (64, 48), (22, 32), (10, 12)])
@pytest.mark.parametrize("s", [1.0, 0.5, 0.25])
def test_filter_matches_rocfft(gpu, tmp_path, Nt, Nx, s):
    Y = gpu
    assert os.path.exists(EXE), "tests/host/rocfft_check is built by __graft_entry__.build()"
    xi = np.random.default_rng(Nt * 7 + Nx).standard_normal((Nt, Nx))
    xi.tofile(tmp_path / "xi.bin")
    O.mask(Nt, Nx, s).tofile(tmp_path / "mask.bin")
    r = subprocess.run([EXE, str(Nt), str(Nx), str(tmp_path / "xi.bin"), str(tmp_path / "mask.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(Nt, Nx)
    p, _ = mk(Y, Nt=Nt, Nx=Nx, cutFraction=s)
    with Y.Context(p) as c:
        got = c.coloured_filter(xi)
    e_lib, e_orc = rel(got, ref), rel(O.coloured_filter(xi, s), ref)
    print(f"# {Nt}x{Nx} s={s}: |ycn - rocFFT| / |rocFFT| = {e_lib:.2e}, |oracle - rocFFT| = {e_orc:.2e}")
    assert e_lib < 1e-12, e_lib
    assert e_orc < 1e-12, e_orc
