"""The drop-in executable `ycn_langevin input.toml` on the GPU: a short run of the
reference's control flow (burn-in, epsBar calibration, measurements) writes well-formed
traces.csv, slice.dat and data.csv that the analysis module reads; with
[cg] precond = "cnn" the first steps train the preconditioner online and the run still
completes with converged solves."""
import os
import subprocess

import numpy as np
import pytest

import yukawacolourednoise_amd.analysis as A

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "yukawacolourednoise_amd", "ycn_langevin")

INPUT = """[physics]
useMass = "true"
mass = -0.2
g = 0.5
kappa = 0.18
lambda = 0.02
cutFraction = 0.5

[langevin]
averageEpsilon = 0.01
MaxLangevinTime = 2.0
ExportTime = 0.25
burnCount = 2
MeasureDriftCount = 2

[io]
configFileName = "test.hdf"
export = "false"
timeSliceFileName = "slice.dat"

[random]
seed = 1234

[fermions]
yukawa_coupling = 0.3
fermion_mass = 0.3
driftMode = 0
WilsonParam = 1.0

[lattice]
Nt = 16
Nx = 16

[cg]
tol = 1e-10
tolMode = "rel"
maxIter = 10000
{extra}
"""


def _run(tmp_path, extra=""):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    (tmp_path / "input.toml").write_text(INPUT.format(extra=extra))
    print(f"# ycn_langevin in {tmp_path} with {extra!r}", flush=True)
    r = subprocess.run([EXE, "input.toml"], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _check_files(tmp_path, stdout):
    assert "Thermalization done!" in stdout
    D = A.Dataset(16, 16).add_data(str(tmp_path))
    n = len(D.fields_data["sigma"])
    assert n >= 4, stdout
    assert D.S_t.shape == (n, 16) and D.Sq_t.shape == (n, 16)
    for k in ("tr", "sigma", "phi"):
        assert np.all(np.isfinite(D.fields_data[k]))
    # time-slice sums / Nx average to sigma
    assert np.allclose(D.S_t.mean(axis=1), D.fields_data["sigma"], rtol=1e-4, atol=1e-6)
    mag, err = D.compute_mag()
    assert np.isfinite(mag) and err >= 0
    # stats.csv: <phi^2> positive, every solve converged to the 1e-10 bar
    s = D.stats
    assert s is not None and len(s["phi2"]) == n
    assert np.all(s["phi2"] > 0) and np.all(s["cg_solves"] >= 1) and np.all(s["cg_iters_mean"] > 0)
    assert np.all(s["cg_not_converged"] == 0) and np.all(s["cg_true_rel_max"] <= 1e-10)


def test_driver_fp64(gpu, tmp_path):
    out = _run(tmp_path)
    _check_files(tmp_path, out)


def test_driver_even_odd_mixed(gpu, tmp_path):
    out = _run(tmp_path, 'precision = "mixed"\nevenOdd = true\n')
    _check_files(tmp_path, out)


def test_driver_mixed_with_trained_cnn(gpu, tmp_path):
    out = _run(tmp_path, 'precision = "mixed"\nprecond = "cnn"\ncnnTrainPairs = 40\ncnnPairsPerSolve = 8\n')
    _check_files(tmp_path, out)
