"""Learned preconditioner on the GPU (BF16 MFMA CNN, flexible CG, online training),
checked against the CPU FP64 oracle of the same network (tests/test_oracle_cnn.py
pins that oracle with finite differences).

Tolerances: the GPU rounds weights and activations to bfloat16 (8-bit mantissa) and
accumulates in FP32; the oracle emulates the bf16 rounding and accumulates in FP64, so
an activation can land on the other side of a bf16 rounding boundary.  Forward outputs
agree to 1e-2 of the largest output (relative L2 error < 3e-3); the gradient, whose
oracle has no bf16 rounding at all, agrees in direction (cosine > 0.99) and norm (10%).
The gates run in FP32 on the GPU and in FP64 in the oracle.  Solves reach the same FP64
true residual as plain CG (<= 1e-10)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rand_spinor(rng, V, scale=1.0):
    return scale * (rng.standard_normal((V, 4)) + 1j * rng.standard_normal((V, 4)))


def ctx(Y, Nt, Nx, **kw):
    base = dict(Nt=Nt, Nx=Nx, m2=-0.4, lam=1.0, g=0.5, M=0.05, r=1.0, cutFraction=0.5, epsBar=0.005,
                adaptive=0, seed=1234, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    base.update(kw)
    return Y.Context(Y.default_params(**base))


def rand_weights(rng, scale=0.3):
    return (scale * rng.standard_normal(O.CNN_NW)).astype(np.float32)


@pytest.mark.parametrize("Nt,Nx", [(16, 16), (32, 48)])
def test_cnn_forward_vs_oracle(gpu, Nt, Nx):
    Y = gpu
    rng = np.random.default_rng(Nt + Nx)
    W = rand_weights(rng)
    r = rand_spinor(rng, Nt * Nx, 1e-3)
    phi = O.hot_start(3, Nt, Nx)
    with ctx(Y, Nt, Nx) as c:
        c.set_field(phi)
        c.precond_set_weights(W)
        assert np.array_equal(c.precond_get_weights(), W)
        z = c.precond_apply(r)
        z2 = c.precond_apply(7.5 * r)
    ref = O.cnn_forward(Nt, Nx, W, r, bf16=True, c0=O.cnn_c0(1.0, 0.05), phi=phi, g=0.5)
    err = np.abs(z - ref).max() / np.abs(ref).max()
    l2 = np.linalg.norm(z - ref) / np.linalg.norm(ref)
    assert err < 1e-2 and l2 < 3e-3, (err, l2)
    # linear in r (the gates depend on the field only): z(a r) = a z(r)
    assert np.linalg.norm(z2 - 7.5 * z) / np.linalg.norm(7.5 * z) < 1e-3
    # the gates are live: the same weights without the field give another output
    ref0 = O.cnn_forward(Nt, Nx, W, r, bf16=True, c0=O.cnn_c0(1.0, 0.05))
    assert np.linalg.norm(z - ref0) > 10 * np.linalg.norm(z - ref)


def test_cnn_gradient_vs_oracle(gpu):
    Y = gpu
    Nt, Nx = 16, 16
    rng = np.random.default_rng(2)
    W = rand_weights(rng, 0.2)
    p = O.params(Nt=Nt, Nx=Nx, M=0.05, g=0.5)
    phi = O.hot_start(5, Nt, Nx)
    e = rand_spinor(rng, Nt * Nx)
    r = O.normal_op(p, phi, e, 1)
    with ctx(Y, Nt, Nx) as c:
        c.precond_set_weights(W)
        L, g = c.precond_loss_grad(r, e, phi=phi)
    Lr, gr = O.cnn_loss_grad(p, phi, W, r, e, c0=O.cnn_c0(1.0, 0.05))
    assert abs(L - Lr) <= 2e-2 * abs(Lr), (L, Lr)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.99, cos
    assert abs(np.linalg.norm(g) / np.linalg.norm(gr) - 1) < 0.1


def test_fcg_converges_to_fp64_residual(gpu):
    Y = gpu
    Nt, Nx = 32, 32
    rng = np.random.default_rng(3)
    phi = O.hot_start(9, Nt, Nx)
    b = rand_spinor(rng, Nt * Nx)
    p = O.params(Nt=Nt, Nx=Nx, M=0.05, g=0.5, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    xr, sr = O.cg(p, phi, b, 1)
    for precond in (Y.YCN_PRECOND_CNN, Y.YCN_PRECOND_NONE):
        with ctx(Y, Nt, Nx, precond=precond, cgPrecision=Y.YCN_CG_MIXED) as c:
            x, st = c.cg_solve(b, phi=phi, mtype=Y.YCN_DAGGER)
        assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
        Ax = O.normal_op(p, phi, x, 1)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= 1.5e-10
        assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-7


def test_fcg_bad_network_falls_back_to_plain_cg(gpu):
    """A useless (large random) network must not break the solve: flexible CG stalls,
    switches the preconditioner off and plain CG still reaches the FP64 tolerance."""
    Y = gpu
    Nt, Nx = 32, 32
    rng = np.random.default_rng(4)
    phi = O.hot_start(1, Nt, Nx)
    b = rand_spinor(rng, Nt * Nx)
    with ctx(Y, Nt, Nx, precond=Y.YCN_PRECOND_CNN) as c:
        c.precond_set_weights(rand_weights(rng, 3.0))
        x, st = c.cg_solve(b, phi=phi, mtype=Y.YCN_DAGGER)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st


def test_online_training_reduces_loss(gpu):
    Y = gpu
    Nt, Nx = 32, 32
    rng = np.random.default_rng(6)
    p = O.params(Nt=Nt, Nx=Nx, M=0.05, g=0.5, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    with ctx(Y, Nt, Nx, precond=Y.YCN_PRECOND_CNN) as c:
        c.hot_start(7)
        # held-out pair from a solve at the start field
        phi = c.get_field()
        b = rand_spinor(rng, Nt * Nx)
        e, _ = O.cg(p, phi, b, 1)
        L0, _ = c.precond_loss_grad(b, e, phi=phi)
        st = c.precond_train(0, 40, pairs_per_solve=4, lr=2e-3)
        L1, _ = c.precond_loss_grad(b, e, phi=phi)
    assert st["pairs"] >= 100
    assert L1 < L0, (L0, L1, st)
