"""Statistical analysis of Langevin runs: the reference's estimators, restated.

The reference analyses the files its Langevin program writes (``traces.csv``,
``slice.dat``, ``data.csv``) with Python (``analysis/functions.py``,
``analysis/correlations.py``, ``analysis/read_in_data.py``) and a C++ implementation of
Wolff's Gamma method behind ctypes (``analysis/AutoCorrelation/ac.cpp``).  This module
gives the same estimators, under the same names, for the files ``ycn_langevin`` writes
and for arrays of measurements, so a long GPU run can be judged against the oracle (or
against a reference run) within statistical errors.

Two error estimates of the reference are not standard error propagation; a
``reference_quirk=True`` argument reproduces them, the default propagates errors:

* ``compute_binder_cumulant``: the reference's error is (U/3) * sqrt((x/dx)^2 + (y/dy)^2)
  with x, y = <s^4>, <s^2> (functions.py:73-78,219; the comment there says (dx/x)^2 +
  (dy/y)^2).  Propagated: U = 1 - R/3, R = x/y^2, dU = (R/3) sqrt((dx/x)^2 + (2 dy/y)^2).
* ``get_phys_quark_mass_via_sinh`` / ``_exp``: ``fitToSinh`` / ``fitToExp``
  (functions.py:126-157) return the variance pcov[0][0] as the error of the mass; the
  standard error is its square root.
"""
from __future__ import annotations

import os
from collections import namedtuple

import numpy as np

ACreturn = namedtuple("ACreturn", ["average", "sigmaF", "sigmaF_error", "ACtime", "ACtimeError"])
ACreturn.__doc__ = ("Mean, its error, the error of the error, integrated autocorrelation time and its error "
                    "(fields of ACreturn, analysis/AutoCorrelation/ac.h:21-27).")


def ComputeStatistics(data, Stau: float = 1.5) -> ACreturn:
    """Wolff's Gamma method with automatic windowing (U. Wolff, Comput. Phys. Commun.
    156 (2004) 143), as ``AutoCorrelation`` does it (analysis/AutoCorrelation/ac.cpp:72-182).

    Gamma(t) = sum_i d_i d_{i+t} / (N - t) for t <= N/2 (d = data - mean; computed with a
    zero-padded FFT of length N + N/2, which gives the linear sums exactly); the window W is
    the first t where exp(-t/tau_W) - tau_W/sqrt(t N) < 0 with
    tau_W = Stau / log((G+1)/G), G = sum_{s=1..t} rho(s) (tau_W = 1e-5 if G <= 0), or N/2
    if there is none.  C_F = Gamma(0) + 2 sum_{t=1..W} Gamma(t); every Gamma(t) is then
    bias-corrected by + C_F/N, C_F recomputed, error = sqrt(C_F/N), its error
    error * sqrt((W + 1/2)/N), tau_int = sum_{t=0..W} Gamma(t) / Gamma(0) - 1/2 and
    its error 2 tau_int sqrt((W - tau_int + 1/2)/N).  Constant data give their value with
    zero errors and autocorrelation time (ac.cpp:83 returns all zeros)."""
    x = np.asarray(data, dtype=np.float64).ravel()
    N = x.size
    if N == 0:
        raise ValueError("ComputeStatistics needs at least one value")
    abb = x.sum() / N
    d = x - abb
    G0 = float(np.dot(d, d)) / N
    if G0 == 0.0:   # constant data (the reference returns a zero average here too)
        return ACreturn(float(abb), 0.0, 0.0, 0.0, 0.0)
    tmax = N // 2
    L = N + tmax
    f = np.fft.rfft(d, n=L)
    acf = np.fft.irfft(np.abs(f) ** 2, n=L)[:tmax + 1]
    gamma = acf / (N - np.arange(tmax + 1))
    rho = gamma / gamma[0]
    eps = 1e-5
    W, found, gint = 0, False, 0.0
    for t in range(1, tmax + 1):
        gint += rho[t]
        tauW = eps if gint <= 0.0 else Stau / np.log((gint + 1.0) / gint)
        gW = np.exp(-t / tauW) - tauW / np.sqrt(t * N)
        if gW < 0.0:
            W, found = t, True
            break
    if not found:
        W = tmax
    CF = gamma[0] + 2.0 * gamma[1:W + 1].sum()
    gamma = gamma + CF / N
    CF = gamma[0] + 2.0 * gamma[1:W + 1].sum()
    tau = gamma[:W + 1].sum() / gamma[0] - 0.5
    dtau = tau * 2.0 * np.sqrt((W - tau + 0.5) / N)
    with np.errstate(invalid="ignore"):
        dvalue = np.sqrt(CF / N)
    ddvalue = dvalue * np.sqrt((W + 0.5) / N)
    return ACreturn(float(abb), float(dvalue), float(ddvalue), float(tau), float(dtau))


def better_jackknife(samples):
    """Jackknife mean (bias-corrected) and error of the mean of `samples` along axis 0,
    from the closed-form leave-one-out means (analysis/correlations.py:59-71)."""
    s = np.asarray(samples, dtype=np.float64)
    N = s.shape[0]
    mean = s.mean(axis=0)
    sub = (N * mean - s) / (N - 1)
    unbiased = N * mean - (N - 1) * sub.mean(axis=0)
    err = np.sqrt((N - 1) * np.mean((sub - mean) ** 2, axis=0))
    return unbiased, err


def jackknife(samples):
    """Same estimator as ``better_jackknife`` (analysis/correlations.py:89-104, which
    builds the leave-one-out means by deletion)."""
    return better_jackknife(samples)


def get_mag(cfgs):
    """Mean and jackknife error of the magnetisation of configurations cfgs[n, ...]
    (analysis/correlations.py:17-20)."""
    c = np.asarray(cfgs, dtype=np.float64)
    return jackknife(c.reshape(c.shape[0], -1).mean(axis=1))


def get_abs_mag(cfgs):
    """Mean and jackknife error of |magnetisation| (analysis/correlations.py:22-25)."""
    c = np.asarray(cfgs, dtype=np.float64)
    return jackknife(np.abs(c.reshape(c.shape[0], -1).mean(axis=1)))


def get_chi2(cfgs):
    """Jackknife of (m - <m>)^2 over configurations (analysis/correlations.py:27-34)."""
    c = np.asarray(cfgs, dtype=np.float64)
    m = c.reshape(c.shape[0], -1).mean(axis=1)
    return jackknife((m - m.mean()) ** 2)


def get_connected_2pt_fct(S_t):
    """Connected time-slice correlator G(t) = < mean_{t0} dS(t0) dS(t0 - t) > over
    configurations, dS = S - <S>, with jackknife errors; S_t[n, t]
    (analysis/correlations.py:220-230)."""
    S = np.asarray(S_t, dtype=np.float64)
    d = S - S.mean(axis=0)
    per_cfg = np.stack([np.mean(d * np.roll(d, t, axis=1), axis=1) for t in range(S.shape[1])], axis=1)
    return better_jackknife(per_cfg)


def _distances(Nt):
    a = np.arange(Nt)
    a[a > Nt // 2] = Nt - a[a > Nt // 2]
    return a


def ren_mass2_from_connected(G, G_err, lattice_size):
    """Second-moment renormalised mass^2 from a connected time-slice correlator:
    chi2 = V_s sum G, mu2 = d V_s sum G a^2, m^2 = 2 d chi2 / mu2, with the reference's
    error propagation (analysis/correlations.py:193-207)."""
    G = np.asarray(G, dtype=np.float64)
    Ge = np.asarray(G_err, dtype=np.float64)
    Vs = float(np.prod(lattice_size[1:]))
    d = len(lattice_size)
    a = _distances(int(lattice_size[0]))
    chi2 = Vs * G.sum()
    mu2 = d * Vs * np.sum(G * a ** 2)
    m2 = 2 * d * chi2 / mu2
    chi_err = Vs * np.sqrt(np.sum(Ge ** 2))
    mu_err = 2 * Vs * np.sqrt(np.sum((a ** 2 * Ge) ** 2))
    m2_err = m2 * np.sqrt((chi_err / chi2) ** 2 + (mu_err / mu2) ** 2)
    return m2, m2_err


def get_ren_mass2_right_via_timeslices(S_t, lattice_size):
    """Renormalised mass^2 from the time slices (analysis/correlations.py:187-210)."""
    G, Ge = get_connected_2pt_fct(S_t)
    return ren_mass2_from_connected(G, Ge, lattice_size)


def get_ren_mass_right_via_timeslices(S_t, lattice_size):
    """Renormalised mass from the time slices (analysis/correlations.py:143-163)."""
    m2, m2_err = get_ren_mass2_right_via_timeslices(S_t, lattice_size)
    m = np.sqrt(m2)
    return m, m * 0.5 * m2_err / m2


def get_fermionic_correlator(Sq_t):
    """Mean and standard deviation over measurements of the fermion correlator
    (analysis/functions.py:160-163)."""
    s = np.asarray(Sq_t, dtype=np.float64)
    return s.mean(axis=0), s.std(axis=0)


def _fit_mass(model, corr, Nt, interval, step=1, reference_quirk=False):
    from scipy.optimize import curve_fit
    start, end = (1, Nt - 1) if interval is None else interval
    y = np.asarray(corr, dtype=np.float64)[start:end]
    x = np.arange(start, end, step, dtype=np.float64)[:y.size]
    popt, pcov = curve_fit(model, x, y, p0=[1.0, 1.0], maxfev=5000)
    err = pcov[0][0] if reference_quirk else np.sqrt(pcov[0][0])
    return float(popt[0]), float(err)


def get_phys_quark_mass_via_sinh(corr, volume, interval=None, filtering=False, reference_quirk=False):
    """Fit A sinh(m (Nt/2 - t)) to the fermion correlator on [start, end)
    (analysis/functions.py:126-157,165-183); `filtering` keeps the even entries."""
    Nt = int(volume[0])
    c = np.asarray(corr, dtype=np.float64)
    if filtering:
        c = c[::2]
    return _fit_mass(lambda t, m, A: A * np.sinh(m * (Nt / 2 - t)), c, Nt, interval, 2 if filtering else 1,
                     reference_quirk)


def get_phys_quark_mass_via_exp(corr, volume, interval=None, reference_quirk=False):
    """Fit A exp(-m t) to the fermion correlator on [start, end)
    (analysis/functions.py:126-139,185-194)."""
    Nt = int(volume[0])
    return _fit_mass(lambda t, m, A: A * np.exp(-t * m), corr, Nt, interval, 1, reference_quirk)


def read_traces(path):
    """``traces.csv`` (header ``tr,sigma,phi``) as a dict of arrays."""
    rows = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    return {"tr": rows[:, 0], "sigma": rows[:, 1], "phi": rows[:, 2]}


def read_stats(path):
    """``stats.csv`` written by the C++ driver beside the reference's files (Langevin
    time, eps, <phi^2>, and the CG statistics of the solves since the previous
    measurement) as a dict of arrays."""
    with open(path) as f:
        names = f.readline().strip().split(",")
    rows = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    return {n: rows[:, k] for k, n in enumerate(names)}


def read_time_slices(path, Nt=None):
    """``slice.dat``: one line of Nt time-slice sums per measurement; an incomplete last
    line (a run ended mid-write) is dropped (analysis/read_in_data.py:65-95)."""
    lines = [ln.split() for ln in open(path, encoding="ISO-8859-1") if ln.strip()]
    if not lines:
        return np.zeros((0, Nt or 0))
    n = Nt or len(lines[0])
    return np.array([[float(v) for v in ln] for ln in lines if len(ln) == n], dtype=np.float64)


def read_fermion_correlator(path, Nt):
    """``data.csv`` (header ``corr``, Nt lines per measurement) as an array [n, Nt]
    (analysis/functions.py:28-37)."""
    v = np.loadtxt(path, skiprows=1, ndmin=1)
    n = v.size // Nt
    return v[:n * Nt].reshape(n, Nt)


class Dataset:
    """The reference's ``Dataset`` (analysis/functions.py:12-87) over one run folder."""

    def __init__(self, Nt, Nx):
        self.Nt, self.Nx = int(Nt), int(Nx)
        self.volume = (self.Nt, self.Nx)

    def add_data(self, folder):
        self.fields_data = read_traces(os.path.join(folder, "traces.csv"))
        self.S_t = read_time_slices(os.path.join(folder, "slice.dat"), self.Nt)
        self.Sq_t = read_fermion_correlator(os.path.join(folder, "data.csv"), self.Nt)
        st = os.path.join(folder, "stats.csv")
        self.stats = read_stats(st) if os.path.exists(st) else None
        return self

    def compute_mag(self):
        b = ComputeStatistics(self.fields_data["sigma"])
        return b.average, b.sigmaF

    def compute_abs_mag(self):
        b = ComputeStatistics(self.fields_data["phi"])
        return b.average, b.sigmaF

    def compute_condensate(self):
        b = ComputeStatistics(self.fields_data["tr"])
        return b.average, b.sigmaF

    def compute_susceptibility(self):
        return compute_susceptibility(self.fields_data["sigma"], self.Nt * self.Nx)

    def compute_binder_cumulant(self, reference_quirk=False):
        return compute_binder_cumulant(self.fields_data["sigma"], reference_quirk)

    def compute_mphir(self):
        return get_ren_mass_right_via_timeslices(self.S_t, self.volume)

    def compute_mqphys(self, interval=None, func=None, filtering=False, reference_quirk=False):
        val, _ = get_fermionic_correlator(self.Sq_t)
        if func is None or func == "sinh":
            return get_phys_quark_mass_via_sinh(val, self.volume, interval, filtering, reference_quirk)
        if func == "exp":
            return get_phys_quark_mass_via_exp(val, self.volume, interval, reference_quirk)
        raise ValueError(f"unknown fit function {func!r}")


def compute_susceptibility(sigma, volume):
    """V <(sigma - <sigma>)^2> with Gamma-method error (analysis/functions.py:58-61)."""
    s = np.asarray(sigma, dtype=np.float64)
    b = ComputeStatistics(s)
    b = ComputeStatistics((s - b.average) ** 2)
    return volume * b.average, volume * b.sigmaF


def compute_binder_cumulant(sigma, reference_quirk=False):
    """U = 1 - <sigma^4> / (3 <sigma^2>^2) with the error propagated from the Gamma-method
    errors of both moments (analysis/functions.py:73-78)."""
    s = np.asarray(sigma, dtype=np.float64)
    b4 = ComputeStatistics(s ** 4)
    b2 = ComputeStatistics(s ** 2)
    val = 1.0 - b4.average / (3.0 * b2.average ** 2)
    if reference_quirk:
        err = val / 3.0 * np.sqrt((b4.average / b4.sigmaF) ** 2 + (b2.average / b2.sigmaF) ** 2)
    else:
        # U = 1 - R/3, R = <s^4>/<s^2>^2: dU = (R/3) sqrt((d4/<s^4>)^2 + (2 d2/<s^2>)^2)
        err = (1.0 - val) * np.sqrt((b4.sigmaF / b4.average) ** 2 + (2.0 * b2.sigmaF / b2.average) ** 2)
    return val, err
