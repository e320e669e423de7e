"""Pins the statistical estimators (yukawacolourednoise_amd/analysis.py) with closed forms,
brute force and processes of known autocorrelation, and the file readers against what the
C++ driver writes.  CPU only."""
import os
import subprocess

import numpy as np
import pytest

import yukawacolourednoise_amd.analysis as A


def _gamma_direct(x):
    """Gamma(t) by direct double sums (brute force for the FFT path)."""
    d = x - x.mean()
    N = len(x)
    return np.array([np.dot(d[:N - t], d[t:]) / (N - t) for t in range(N // 2 + 1)])


def test_gamma_fft_equals_direct_sums():
    x = np.random.default_rng(3).standard_normal(301).cumsum()
    N = len(x)
    tmax = N // 2
    f = np.fft.rfft(x - x.mean(), n=N + tmax)
    acf = np.fft.irfft(np.abs(f) ** 2, n=N + tmax)[:tmax + 1] / (N - np.arange(tmax + 1))
    assert np.allclose(acf, _gamma_direct(x), rtol=1e-12, atol=1e-12)


def test_statistics_iid():
    """Independent samples: tau_int = 1/2 within its error; the error is std/sqrt(N)."""
    x = np.random.default_rng(0).standard_normal(20000)
    r = A.ComputeStatistics(x)
    assert r.average == pytest.approx(x.mean(), abs=1e-14)
    assert abs(r.ACtime - 0.5) < 3 * r.ACtimeError + 0.02
    assert r.sigmaF == pytest.approx(x.std() / np.sqrt(len(x)), rel=0.1)
    assert r.sigmaF_error == pytest.approx(r.sigmaF * np.sqrt(0.5 / len(x) + 1.0 / len(x)), rel=2.0)


@pytest.mark.parametrize("rho", [0.5, 0.9])
def test_statistics_ar1(rho):
    """AR(1) x_t = rho x_{t-1} + e_t: tau_int = (1 + rho) / (2 (1 - rho)) and the error of
    the mean is sqrt(2 tau_int var / N)."""
    rng = np.random.default_rng(int(rho * 10))
    N = 200000
    e = rng.standard_normal(N)
    x = np.empty(N)
    x[0] = e[0] / np.sqrt(1 - rho ** 2)
    for i in range(1, N):
        x[i] = rho * x[i - 1] + e[i]
    r = A.ComputeStatistics(x)
    tau = (1 + rho) / (2 * (1 - rho))
    assert abs(r.ACtime - tau) < 4 * r.ACtimeError, (r, tau)
    var = 1.0 / (1 - rho ** 2)
    assert r.sigmaF == pytest.approx(np.sqrt(2 * tau * var / N), rel=0.1)


def test_statistics_constant_and_alternating():
    assert A.ComputeStatistics(np.full(10, 3.0)) == (3.0, 0.0, 0.0, 0.0, 0.0)
    # strictly alternating data: rho(1) = -1 -> the window closes at W = 1 (tau_W = 1e-5)
    x = np.tile([1.0, -1.0], 50)
    r = A.ComputeStatistics(x)
    g = _gamma_direct(x)
    CF = g[0] + 2 * g[1]
    g = g + CF / len(x)
    assert r.ACtime == pytest.approx((g[0] + g[1]) / g[0] - 0.5, rel=1e-12)


def test_jackknife_of_mean_is_standard_error():
    x = np.random.default_rng(1).standard_normal((50, 3))
    m, e = A.better_jackknife(x)
    assert np.allclose(m, x.mean(axis=0), rtol=0, atol=1e-14)
    assert np.allclose(e, x.std(axis=0, ddof=1) / np.sqrt(len(x)), rtol=1e-12)
    # leave-one-out by deletion (the reference's `jackknife` construction)
    sub = np.array([np.delete(x, i, axis=0).mean(axis=0) for i in range(len(x))])
    e2 = np.sqrt((len(x) - 1) * np.mean((sub - x.mean(axis=0)) ** 2, axis=0))
    assert np.allclose(A.jackknife(x)[1], e2, rtol=1e-12)


def test_connected_correlator_of_plane_waves():
    """S(t) = a cos(2 pi m t / Nt + th_k) with evenly spaced phases: <S> = 0 and every
    configuration gives G(t) = a^2/2 cos(2 pi m t / Nt) exactly, with zero error."""
    Nt, K, a, m = 12, 7, 1.7, 2
    t = np.arange(Nt)
    S = np.array([a * np.cos(2 * np.pi * m * t / Nt + 2 * np.pi * k / K) for k in range(K)])
    G, Ge = A.get_connected_2pt_fct(S)
    assert np.allclose(G, a * a / 2 * np.cos(2 * np.pi * m * t / Nt), atol=1e-13)
    assert np.all(Ge < 1e-7)


def test_second_moment_mass():
    """G = delta_0 + c (delta_1 + delta_{-1}) in 2D: chi2 = Vs (1 + 2c), mu2 = 2 Vs 2c,
    m^2 = 4 chi2 / mu2 = (1 + 2c) / c."""
    Nt, Nx, c = 8, 5, 0.3
    G = np.zeros(Nt)
    G[0], G[1], G[-1] = 1.0, c, c
    m2, _ = A.ren_mass2_from_connected(G, np.full(Nt, 1e-3), (Nt, Nx))
    assert m2 == pytest.approx((1 + 2 * c) / c, rel=1e-14)


def test_binder_and_susceptibility_closed_forms():
    s = np.tile([1.0, -1.0], 500)                 # two-valued magnetisation: U = 2/3
    U, dU = A.compute_binder_cumulant(s)
    assert U == pytest.approx(2.0 / 3.0, abs=1e-14)
    chi, _ = A.compute_susceptibility(s, 64)      # <(s - 0)^2> = 1 -> chi = V
    assert chi == pytest.approx(64.0, rel=1e-14)
    g = np.random.default_rng(5).standard_normal(400000)   # Gaussian: U -> 0
    U, dU = A.compute_binder_cumulant(g)
    assert abs(U) < 5 * dU + 1e-3
    # the reference's own error expression is available on request
    _, dq = A.compute_binder_cumulant(g, reference_quirk=True)
    assert np.isfinite(dq)


def test_quark_mass_fits():
    Nt = 16
    t = np.arange(Nt)
    m, amp = 0.37, 2.5
    corr = amp * np.sinh(m * (Nt / 2 - t))
    val, err = A.get_phys_quark_mass_via_sinh(corr, (Nt, 8))
    assert val == pytest.approx(m, rel=1e-6) and err < 1e-6
    val, _ = A.get_phys_quark_mass_via_exp(amp * np.exp(-m * t), (Nt, 8))
    assert val == pytest.approx(m, rel=1e-6)


HOST = os.path.join(os.path.dirname(__file__), "..", "yukawacolourednoise_amd", "host")


def test_dataset_reads_driver_output(tmp_path):
    """The readers parse what the C++ driver writes: run it with the deterministic test
    backend (tests/host/driver_check.cpp), whose measurement at step counter c reports
    trace 2c, magnetisation 0.5c - 40, slices t + c/4 and correlator exp(-t/10) + c/1000.
    The driver writes tr = trace / V, sigma = magnetisation / sqrt(2 kappa) / V (useMass
    false), slices / Nx, all with 6 significant digits."""
    from test_host_loop import INPUT, build_driver
    exe = build_driver(tmp_path)
    inp = tmp_path / "input.toml"
    inp.write_text(INPUT.format(useMass="false", tmax=4.0, adaptive="false"))
    subprocess.run([exe, str(inp), str(tmp_path)], check=True, capture_output=True, text=True)
    Nt, Nx = 6, 8
    V = Nt * Nx
    D = A.Dataset(Nt, Nx).add_data(str(tmp_path))
    tr = D.fields_data
    n = len(tr["sigma"])
    assert n >= 3
    assert D.S_t.shape == (n, Nt) and D.Sq_t.shape == (n, Nt)
    c = np.round(tr["tr"] * V / 2)                 # the step counters of the measurements
    assert np.all(np.diff(c) > 0)
    assert np.allclose(tr["sigma"], (0.5 * c - 40) / np.sqrt(2 * 0.18) / V, rtol=1e-5)
    assert np.allclose(D.S_t, (np.arange(Nt)[None, :] + 0.25 * c[:, None]) / Nx, rtol=1e-5)
    assert np.allclose(D.Sq_t, np.exp(-0.1 * np.arange(Nt))[None, :] + 1e-3 * c[:, None], rtol=1e-5)
    mag, _ = D.compute_mag()
    assert mag == pytest.approx(tr["sigma"].mean())
    assert D.stats is not None and len(D.stats["phi2"]) == n
    assert np.allclose(D.stats["phi2"], 3 * c / V, rtol=1e-5)
    assert np.allclose(D.stats["cg_iters_mean"], 100 + c, rtol=1e-5)
