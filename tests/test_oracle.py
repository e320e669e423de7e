"""Pins the CPU FP64 oracle (oracle/ycn_oracle.c) before the parity tests trust it.

Each check is independent of the oracle's own code: published known-answer vectors,
library routines (numpy.fft, numpy.linalg), closed forms, and an independent numpy
construction of the Wilson-Dirac matrix from its defining formula
D = (2r + M + g phi) 1 - 1/2 sum_mu [(r - gamma_mu) T_{+mu} + (r + gamma_mu) T_{-mu}],
gamma_t = sigma_3, gamma_x = sigma_1, antiperiodic in time
(code/src/Dirac.cu:49-70,144-190; Thesis.pdf p.19-21).
"""
import math

import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------------------
# Philox4x32-10 known-answer vectors (Random123 distribution, kat_vectors)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, expect):
    assert O.philox(ctr, key) == expect


def test_normals_distribution():
    n = O.normals(1234, 0, 3, 400_000)
    assert abs(n.mean()) < 5 * 1 / math.sqrt(n.size)
    assert abs(n.var() - 1) < 5 * math.sqrt(2 / n.size)
    # Kolmogorov-Smirnov against N(0,1)
    from scipy import stats
    assert stats.kstest(n[:100_000], "norm").pvalue > 1e-3
    # pairs from one Philox block are uncorrelated; streams and steps differ
    assert abs(np.corrcoef(n[0::2], n[1::2])[0, 1]) < 0.01
    assert not np.array_equal(O.normals(1234, 0, 4, 16), O.normals(1234, 0, 3, 16))
    assert not np.array_equal(O.normals(1234, 1, 3, 16), O.normals(1234, 0, 3, 16))


def test_hot_start_range():
    phi = O.hot_start(7, 64, 64)
    assert phi.min() >= -1 and phi.max() <= 1
    assert abs(phi.mean()) < 0.05


# ---------------------------------------------------------------------------
# FFT: numpy.fft and the O(n^2) DFT
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 8, 10, 11, 13, 14, 16, 22, 32, 60, 97, 128, 1024])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = max(1.0, np.abs(x).sum())
    assert np.abs(O.fft(x, -1) - np.fft.fft(x)).max() < 1e-14 * scale
    assert np.abs(O.fft(x, +1) - np.fft.ifft(x) * n).max() < 1e-14 * scale
    if n <= 32:
        k = np.arange(n)
        dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
        assert np.abs(O.fft(x, -1) - dft).max() < 1e-13 * scale


# ---------------------------------------------------------------------------
# coloured-noise mask and filter (colouredNoise.cpp:16-29,138-178)
# ---------------------------------------------------------------------------
# kept half-spectrum modes: count of (kt, kx<=N/2) with p_t^2 + p_x^2 <= 2 (s pi)^2
@pytest.mark.parametrize("N,s,kept", [
    (32, 0.5, 214), (256, 0.5, 12963), (512, 0.5, 51632), (1024, 0.5, 206239),
    (32, 0.25, 56), (256, 0.25, 3264), (512, 0.25, 12963), (1024, 0.25, 51632),
])
def test_mask_counts(N, s, kept):
    m = O.mask(N, N, s)
    assert m.shape == (N, N // 2 + 1)
    assert int((m > 0).sum()) == kept
    assert np.all((m == 0) | (m == 1.0 / (N * N)))


def test_mask_closed_form_small():
    Nt, Nx, s = 12, 10, 0.6
    m = O.mask(Nt, Nx, s)
    kt = np.arange(Nt)
    kx = np.arange(Nx // 2 + 1)
    pt = np.where(kt <= Nt // 2, kt, kt - Nt) * 2 * np.pi / Nt
    px = kx * 2 * np.pi / Nx
    keep = (pt[:, None] ** 2 + px[None, :] ** 2) <= 2 * (s * np.pi) ** 2
    assert np.array_equal(m > 0, keep)


@pytest.mark.parametrize("Nt,Nx", [(16, 16), (6, 32), (2, 32), (10, 12), (9, 7)])
@pytest.mark.parametrize("s", [1.0, 0.5, 0.25])
def test_filter_vs_numpy_rfft(Nt, Nx, s):
    rng = np.random.default_rng(Nt * Nx)
    xi = rng.standard_normal((Nt, Nx))
    m = O.mask(Nt, Nx, s) > 0
    ref = np.fft.irfft2(np.fft.rfft2(xi) * m, s=(Nt, Nx))
    out = O.coloured_filter(xi, s)
    assert np.abs(out - ref).max() < 1e-13 * max(1, np.abs(ref).max())
    if s == 1.0:
        assert np.abs(out - xi).max() < 1e-13  # every mode kept: white noise unchanged
    # a real symmetric projection: idempotent and self-adjoint
    assert np.abs(O.coloured_filter(out, s) - out).max() < 1e-13
    y = rng.standard_normal((Nt, Nx))
    assert abs(np.vdot(out, y) - np.vdot(xi, O.coloured_filter(y, s))) < 1e-12 * Nt * Nx


def test_coloured_noise_variance():
    # Var(noise) = kept fraction of all V modes (Parseval for a projection of white noise)
    N, s = 64, 0.5
    p = O.params(Nt=N, Nx=N, cutFraction=s)
    m = O.mask(N, N, s) > 0
    full = np.zeros((N, N), bool)
    full[:, : N // 2 + 1] = m
    full[:, N // 2 + 1:] = m[(-np.arange(N)) % N][:, 1:N // 2][:, ::-1]
    frac = full.sum() / (N * N)
    v = np.mean([O.coloured_noise(p, st).var() for st in range(20)])
    assert abs(v / frac - 1) < 0.05


# ---------------------------------------------------------------------------
# Wilson-Dirac operator
# ---------------------------------------------------------------------------
S1 = np.array([[0, 1], [1, 0]], complex)
S3 = np.array([[1, 0], [0, -1]], complex)


def wilson_dense(Nt, Nx, phi, M, g, r):
    """Independent construction from the defining formula (not from the oracle)."""
    V = Nt * Nx
    D = np.zeros((4 * V, 4 * V), complex)
    I2 = np.eye(2)

    def idx(t, x):
        return (t % Nt) * Nx + (x % Nx)

    for t in range(Nt):
        for x in range(Nx):
            i = idx(t, x)
            for f in range(2):
                sl = slice(4 * i + 2 * f, 4 * i + 2 * f + 2)
                D[sl, sl] += (2 * r + M + g * phi[t, x]) * I2
                for (dt, dx, gam) in ((1, 0, S3), (0, 1, S1)):
                    jp = idx(t + dt, x + dx)
                    jm = idx(t - dt, x - dx)
                    sp = -1.0 if (dt and t == Nt - 1) else 1.0
                    sm = -1.0 if (dt and t == 0) else 1.0
                    D[sl, 4 * jp + 2 * f:4 * jp + 2 * f + 2] += -0.5 * sp * (r * I2 - gam)
                    D[sl, 4 * jm + 2 * f:4 * jm + 2 * f + 2] += -0.5 * sm * (r * I2 + gam)
    return D


def oracle_dense(p, phi, dagger):
    V = p.Nt * p.Nx
    D = np.zeros((4 * V, 4 * V), complex)
    for j in range(4 * V):
        e = np.zeros(4 * V, complex)
        e[j] = 1
        D[:, j] = O.dirac(p, phi, e.reshape(V, 4), dagger).reshape(-1)
    return D


@pytest.mark.parametrize("Nt,Nx,r", [(4, 6, 1.0), (6, 4, 0.7), (2, 4, 1.0), (4, 3, 0.0)])
def test_dirac_vs_formula(Nt, Nx, r):
    p = O.params(Nt=Nt, Nx=Nx, M=0.3, g=0.5, r=r)
    phi = np.random.default_rng(1).uniform(-1, 1, (Nt, Nx))
    D = oracle_dense(p, phi, False)
    ref = wilson_dense(Nt, Nx, phi, 0.3, 0.5, r)
    assert np.abs(D - ref).max() < 1e-15
    Dd = oracle_dense(p, phi, True)
    assert np.abs(Dd - D.T).max() == 0.0          # D real, Dagger = transpose
    assert np.abs(D.imag).max() == 0.0
    g5 = np.kron(np.eye(2 * Nt * Nx), np.array([[0, -1j], [1j, 0]]))   # sigma_2 per flavour
    assert np.abs(g5 @ D @ g5 - D.T).max() < 1e-15


def test_dirac_free_spectrum_and_trace():
    Nt, Nx, M, g, r, phi0 = 6, 4, 0.3, 0.5, 1.0, 0.7
    p = O.params(Nt=Nt, Nx=Nx, M=M, g=g, r=r)
    D = oracle_dense(p, np.full((Nt, Nx), phi0), False)
    ev = list(np.linalg.eigvals(D))
    cf, tr = [], 0.0
    for kt in range(Nt):
        for kx in range(Nx):
            pt, px = 2 * np.pi * (kt + 0.5) / Nt, 2 * np.pi * kx / Nx
            A = M + g * phi0 + r * ((1 - np.cos(pt)) + (1 - np.cos(px)))
            s = np.sqrt(np.sin(pt) ** 2 + np.sin(px) ** 2)
            cf += [A + 1j * s, A - 1j * s] * 2
            tr += 4 * A / (A * A + s * s)
    err = 0.0
    for e in ev:
        j = int(np.argmin(np.abs(np.array(cf) - e)))
        err = max(err, abs(cf.pop(j) - e))
    assert err < 1e-13
    assert abs(np.trace(np.linalg.inv(D)).real - tr) < 1e-12 * abs(tr)


# ---------------------------------------------------------------------------
# CG (CGsolver.cu:72-170)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mtype", [0, 1])
def test_cg_vs_dense(mtype):
    Nt, Nx = 4, 6
    p = O.params(Nt=Nt, Nx=Nx, M=0.1, g=0.5, cgTol=1e-12, cgTolMode=1)
    rng = np.random.default_rng(2)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    b = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    D = wilson_dense(Nt, Nx, phi, 0.1, 0.5, 1.0)
    A = D.conj().T @ D if mtype == 1 else D @ D.conj().T
    x, st = O.cg(p, phi, b, mtype)
    xd = np.linalg.solve(A, b.reshape(-1))
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-12
    assert np.abs(x.reshape(-1) - xd).max() < 1e-9 * np.abs(xd).max()


def test_cg_absolute_reference_rule_and_cap():
    Nt, Nx = 8, 8
    phi = O.hot_start(1, Nt, Nx)
    b = np.random.default_rng(0).standard_normal((Nt * Nx, 4)) + 0j
    p = O.params(Nt=Nt, Nx=Nx, M=0.2, g=0.3, cgTol=1e-8, cgTolMode=0, cgMaxIter=1000)
    x, st = O.cg(p, phi, b, 1)
    assert st["converged"] and math.sqrt(st["rr"]) <= 1e-8
    # iteration cap: reported as not converged
    p.cgMaxIter = 3
    x, st = O.cg(p, phi, b, 1)
    assert st["iters"] == 3 and st["converged"] == 0
    # b = 0: zero iterations, x = 0
    x, st = O.cg(p, phi, np.zeros((Nt * Nx, 4), complex), 1)
    assert st["iters"] == 0 and np.all(x == 0)


def test_fermion_drift_definition():
    """drift_x = -g Re sum_k conj((D^dag)^-1 eta)_xk eta_xk, checked with a dense inverse."""
    Nt, Nx = 4, 4
    p = O.params(Nt=Nt, Nx=Nx, M=0.3, g=0.4, cgTol=1e-13, cgTolMode=1)
    phi = O.hot_start(2, Nt, Nx)
    eta = O.fermion_noise(p.seed, 0, 5, Nt * Nx)
    assert np.all(eta.imag == 0)
    D = wilson_dense(Nt, Nx, phi, 0.3, 0.4, 1.0)
    chi = np.linalg.solve(D.conj().T, eta.reshape(-1)).reshape(-1, 4)
    ref = -0.4 * np.real(np.sum(chi.conj() * eta, axis=1)).reshape(Nt, Nx)
    d, _ = O.fermion_drift(p, phi, 5)
    assert np.abs(d - ref).max() < 1e-10
    p.driftMode = 1
    d1, _ = O.fermion_drift(p, phi, 5)
    assert np.abs(d1 - p.cutFraction ** 2 * d).max() < 1e-15


def test_drift_estimates_trace():
    """E[drift_x] = -g [D^-1]_xx (sum over spin and flavour): sample mean over eta."""
    Nt, Nx, g = 4, 4, 0.5
    p = O.params(Nt=Nt, Nx=Nx, M=0.3, g=g, cgTol=1e-12, cgTolMode=1)
    phi = O.hot_start(3, Nt, Nx)
    D = wilson_dense(Nt, Nx, phi, 0.3, g, 1.0)
    Dinv = np.linalg.inv(D)
    diag = np.real(np.diag(Dinv)).reshape(-1, 4).sum(axis=1).reshape(Nt, Nx)
    n = 3000
    samples = np.stack([O.fermion_drift(p, phi, st)[0] for st in range(n)])
    mean, sem = samples.mean(0), samples.std(0) / math.sqrt(n)
    assert np.all(np.abs(mean - (-g * diag)) < 5 * sem + 1e-12)


# ---------------------------------------------------------------------------
# bosonic drift, Laplacian and update (langevin_gpu_v2.cu:9-117, Laplace.cu:61-100)
# ---------------------------------------------------------------------------
def test_laplace_csr_dense():
    Nt, Nx = 5, 3
    I, J, val = O.laplace_csr(Nt, Nx)
    L = np.zeros((Nt * Nx, Nt * Nx))
    for r in range(Nt * Nx):
        cols = J[I[r]:I[r + 1]]
        assert np.all(np.diff(cols) >= 0)
        for c, v in zip(cols, val[I[r]:I[r + 1]]):
            L[r, c] += v
    ref = np.zeros_like(L)
    for t in range(Nt):
        for x in range(Nx):
            i = t * Nx + x
            ref[i, i] -= 4
            for (a, b) in (((t + 1) % Nt, x), ((t - 1) % Nt, x), (t, (x + 1) % Nx), (t, (x - 1) % Nx)):
                ref[i, a * Nx + b] += 1
    assert np.array_equal(L, ref)


def test_bosonic_plane_wave_and_update():
    Nt, Nx = 16, 12
    kt, kx = 3, 2
    pt, px = 2 * np.pi * kt / Nt, 2 * np.pi * kx / Nx
    t, x = np.meshgrid(np.arange(Nt), np.arange(Nx), indexing="ij")
    phi = 0.1 * np.cos(pt * t + px * x)
    m2, lam, eps = -0.3, 0.7, 0.01
    p = O.params(Nt=Nt, Nx=Nx, m2=m2, lam=lam, epsBar=eps, adaptive=0)
    noise = np.random.default_rng(0).standard_normal((Nt, Nx))
    fd = np.random.default_rng(1).standard_normal((Nt, Nx))
    phin, drift, e, mx = O.bosonic_update(p, phi, fd, noise)
    lap = (4 * np.sin(pt / 2) ** 2 + 4 * np.sin(px / 2) ** 2) * phi   # -Laplacian of the wave
    ref = fd + phi * (m2 + lam / 6 * phi ** 2) + lap
    assert np.abs(drift - ref).max() < 1e-14
    assert e == eps
    assert np.abs(phin - (phi - (eps * drift + math.sqrt(2 * eps) * noise))).max() < 1e-15
    # adaptive step (langevin_gpu_v2.cu:65-117) and clamps (:248-256)
    p.adaptive = 1
    _, drift, e, mx = O.bosonic_update(p, phi, fd, noise)
    assert mx == np.abs(drift).max() and abs(e - eps * 30 / mx) < 1e-18
    _, _, e, mx = O.bosonic_update(p, np.full((Nt, Nx), 1e3), np.zeros((Nt, Nx)), noise)
    assert mx == 1e5 and e == eps * 30 / 1e5
    _, _, e, mx = O.bosonic_update(p, np.zeros((Nt, Nx)), np.zeros((Nt, Nx)), noise)
    assert mx == 1e-5


def test_langevin_step_composition():
    Nt, Nx = 8, 8
    p = O.params(Nt=Nt, Nx=Nx, adaptive=0, epsBar=0.01)
    phi = O.hot_start(7, Nt, Nx)
    new, eps, st = O.langevin_step(p, phi, 4)
    noise = O.coloured_noise(p, 4)
    d, _ = O.fermion_drift(p, phi, 4)
    ref, _, _, _ = O.bosonic_update(p, phi, d, noise)
    assert np.array_equal(new, ref)


# ---------------------------------------------------------------------------
# measurement: free Wilson fermion mass = log(1 + m_q) (Thesis.pdf p.33 eq. 4.3,
# p.35 Table 4.1)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mq", [1.0, 0.5])
def test_free_fermion_mass(mq):
    Nt, Nx = 96, 4   # finite-Nt contamination ~ exp(-m (Nt - 2t)) < 1e-13 at t = 8
    p = O.params(Nt=Nt, Nx=Nx, M=mq, g=0.0, cgTol=1e-13, cgTolMode=1)
    corr, st = O.correlator(p, np.zeros((Nt, Nx)))
    t = 8
    meff = math.log(corr[t] / corr[t + 1])
    assert abs(meff - math.log(1 + mq)) < 1e-8
    if mq == 1.0:
        assert abs(meff - 0.6931537171644739) < 1e-5   # Table 4.1, m_q = 1.0


# ---------------------------------------------------------------------------
# measurement reductions (reductions.cu:65-104,176-191; main.cu:441-469)
# ---------------------------------------------------------------------------
def test_time_slices_are_row_sums():
    """gpuTimeSlices: S(t) = sum_x phi(t, x), rows t of the row-major (t, x) field."""
    phi = np.random.default_rng(8).standard_normal((6, 10))
    phi[2, :] += 100.0     # a row offset lands in slice 2 only
    s = O.time_slices(phi)
    assert np.allclose(s, phi.sum(axis=1), rtol=0, atol=1e-12)
    assert np.argmax(s) == 2


@pytest.mark.parametrize("Nt,Nx", [(6, 8), (8, 6)])
def test_correlator_vs_dense_inverse(Nt, Nx):
    """main.cu:441-469: corr[t] = sum_x sum_k Re (D^-1 src)_(t,x),k for the point source
    src = delta_{x,0} (1, 1, 0, 0) (flavour 1, both spins, real), computed as CG on D D^dag
    then D^dag.  Checked against the dense inverse of the independently built Wilson
    matrix: the normalisation (no factor, no division by the volume) and the t-row
    summation.  D is flavour-diagonal and the source has no flavour-2 part, so the
    flavour-2 components of D^-1 src vanish and summing k < 2 or k < 4 is the same sum."""
    rng = np.random.default_rng(Nt * Nx)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    p = O.params(Nt=Nt, Nx=Nx, M=0.3, g=0.5, cgTol=1e-13, cgTolMode=1, cgMaxIter=100000)
    corr, st = O.correlator(p, phi)
    assert st["converged"] == 1
    D = wilson_dense(Nt, Nx, phi, 0.3, 0.5, 1.0)
    src = np.zeros(4 * Nt * Nx, complex)
    src[0] = src[1] = 1.0
    psi = np.linalg.solve(D, src).reshape(Nt, Nx, 4)
    assert np.abs(psi[..., 2:]).max() < 1e-14
    want = psi.real.sum(axis=(1, 2))
    assert np.abs(corr - want).max() <= 1e-10 * np.abs(want).max()
    # a source in the wrong spinor slot or a per-site normalisation would fail
    assert np.abs(corr - want / (Nt * Nx)).max() > 1e-3 * np.abs(want).max()
