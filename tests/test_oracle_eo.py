"""Pins the oracle's even-odd preconditioned fermion force against dense linear algebra:
the Schur complement of the dense Dirac matrix, its transpose, the dense solve of
D^dag chi = eta, and the unpreconditioned drift."""
import numpy as np
import pytest

import oracle as O

NT, NX = 6, 8
V = NT * NX


def _dense_dirac(p, phi, dagger):
    """Dense 4V x 4V matrix of D (or D^dag) in the AoS (site, component) order; D is real."""
    n = 4 * V
    M = np.empty((n, n))
    for j in range(n):
        e = np.zeros((V, 4), np.complex128)
        e.reshape(-1)[j] = 1.0
        M[:, j] = O.dirac(p, phi, e, dagger).reshape(-1).real
    return M


def _blocks(M):
    ie = O.eo_index(NT, NX, 0)
    io = O.eo_index(NT, NX, 1)
    ce = (4 * ie[:, None] + np.arange(4)[None, :]).ravel()
    co = (4 * io[:, None] + np.arange(4)[None, :]).ravel()
    return M[np.ix_(ce, ce)], M[np.ix_(ce, co)], M[np.ix_(co, ce)], M[np.ix_(co, co)], ce, co


@pytest.fixture(scope="module")
def setup():
    p = O.params(Nt=NT, Nx=NX, m2=-0.2, lam=0.5, g=0.4, M=0.2, r=1.0)
    phi = np.random.default_rng(0).uniform(-1, 1, (NT, NX))
    return p, phi


def test_eo_index_is_a_checkerboard():
    ie, io = O.eo_index(NT, NX, 0), O.eo_index(NT, NX, 1)
    assert len(ie) == len(io) == V // 2
    assert sorted(np.concatenate([ie, io])) == list(range(V))
    t, x = ie // NX, ie % NX
    assert np.all((t + x) % 2 == 0)
    assert np.array_equal(t * (NX // 2) + x // 2, np.arange(V // 2))


@pytest.mark.parametrize("dagger", [True, False])
def test_schur_matches_dense(setup, dagger):
    """S = M_ee - M_eo M_oo^-1 M_oe for M = D^dag (dagger) or D; the dense blocks use a
    general inverse, so the diagonal structure of M_oo is checked too."""
    p, phi = setup
    Mee, Meo, Moe, Moo, _, _ = _blocks(_dense_dirac(p, phi, dagger))
    assert np.allclose(Moo, np.diag(np.diag(Moo)))
    S = Mee - Meo @ np.linalg.solve(Moo, Moe)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((V // 2, 4)) + 1j * rng.standard_normal((V // 2, 4))
    out = O.eo_schur(p, phi, v, dagger)
    ref = (S @ v.reshape(-1).real + 1j * (S @ v.reshape(-1).imag)).reshape(V // 2, 4)
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


def test_schur_transpose(setup):
    p, phi = setup
    rng = np.random.default_rng(2)
    a = rng.standard_normal((V // 2, 4)) + 1j * rng.standard_normal((V // 2, 4))
    b = rng.standard_normal((V // 2, 4)) + 1j * rng.standard_normal((V // 2, 4))
    lhs = np.vdot(a, O.eo_schur(p, phi, b, True))
    rhs = np.vdot(O.eo_schur(p, phi, a, False), b)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_eo_solve_matches_dense(setup):
    p, phi = setup
    q = O.params(Nt=NT, Nx=NX, m2=-0.2, lam=0.5, g=0.4, M=0.2, r=1.0, cgTol=1e-13, cgTolMode=1, evenOdd=1)
    eta = O.fermion_noise(5, 0, 3, V)
    chi, st = O.eo_cg(q, phi, eta)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-13
    Md = _dense_dirac(p, phi, True)
    ref = np.linalg.solve(Md, eta.reshape(-1).real).reshape(V, 4)
    assert np.abs(chi - ref).max() < 1e-10 * np.abs(ref).max()
    # the reported true residual is that of the unpreconditioned system
    res = eta.reshape(-1) - Md @ chi.reshape(-1)
    assert np.linalg.norm(res) / np.linalg.norm(eta) == pytest.approx(st["true_rel_resid"], rel=1e-3, abs=1e-15)


def test_rhs_and_reconstruct(setup):
    """chi_e of the dense solution reproduces chi_o through the reconstruction, and solves
    S chi_e = b_e."""
    p, phi = setup
    eta = O.fermion_noise(9, 0, 1, V)
    ref = np.linalg.solve(_dense_dirac(p, phi, True), eta.reshape(-1).real).reshape(V, 4)
    chi_e = ref[O.eo_index(NT, NX, 0)]
    full = O.eo_reconstruct(p, phi, eta, chi_e)
    assert np.abs(full - ref).max() < 1e-12 * np.abs(ref).max()
    b = O.eo_rhs(p, phi, eta)
    assert np.abs(O.eo_schur(p, phi, chi_e, True) - b).max() < 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("nNoise", [1, 2])
def test_eo_drift_equals_plain_drift(setup, nNoise):
    _, phi = setup
    kw = dict(Nt=NT, Nx=NX, m2=-0.2, lam=0.5, g=0.4, M=0.2, r=1.0, cgTol=1e-12, cgTolMode=1, nNoise=nNoise)
    d0, s0 = O.fermion_drift(O.params(**kw, evenOdd=0), phi, 4)
    d1, s1 = O.fermion_drift(O.params(**kw, evenOdd=1), phi, 4)
    assert s1["converged"] == 1
    assert np.abs(d1 - d0).max() < 1e-9 * np.abs(d0).max()
    assert s1["iters"] < s0["iters"]
