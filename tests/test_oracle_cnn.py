"""Pins the CPU FP64 oracle of the learned preconditioner: the analytic gradient of the
residual-space loss against central finite differences, scale equivariance, linearity
in r, the identity network, closed-form delta responses that fix the convolution and
gate tap orientation and channel order, and bfloat16 round-to-nearest-even."""
import numpy as np
import pytest

import oracle as O


def test_bf16_rounding():
    L = O.lib()
    assert L.orc_bf16_round(1.0) == 1.0
    assert L.orc_bf16_round(1.0 + 2 ** -8) == 1.0            # tie -> even
    assert L.orc_bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6 or L.orc_bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -7
    assert L.orc_bf16_round(1.0 + 2 ** -7) == 1.0 + 2 ** -7   # representable
    assert abs(L.orc_bf16_round(np.pi) - np.pi) / np.pi < 2 ** -8


N1, N2, N3, NG = 9 * 8 * 16, 9 * 16 * 16, 9 * 16 * 8, 9 * 16   # weight block sizes


def test_identity_network():
    """Centre taps as identity (W1 = a I, W2 = I, W3 = I/a) and zero gates: the linear
    network is the identity, net(u) = u."""
    Nt, Nx = 4, 6
    W = np.zeros(O.CNN_NW, np.float32)
    a = 2.0 ** -10   # a and 1/a exact in float32
    for ci in range(8):
        W[(4 * 8 + ci) * 16 + ci] = a
    for ci in range(16):
        W[N1 + (4 * 16 + ci) * 16 + ci] = 1.0
    for ci in range(8):
        W[N1 + N2 + (4 * 16 + ci) * 8 + ci] = 1.0 / a
    r = np.random.default_rng(0).standard_normal((Nt * Nx, 4)) + 0j
    phi = np.random.default_rng(1).uniform(-1, 1, (Nt, Nx))
    z = O.cnn_forward(Nt, Nx, W, r, bf16=False, phi=phi, g=0.5)
    assert np.abs(z - r).max() < 1e-12 * np.abs(r).max()


def test_linear_in_r():
    """The gates depend on the field only, so z(a r1 + b r2) = a z(r1) + b z(r2)."""
    Nt, Nx = 4, 6
    rng = np.random.default_rng(3)
    W = (0.3 * rng.standard_normal(O.CNN_NW)).astype(np.float32)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    r1 = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    r2 = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    f = lambda r: O.cnn_forward(Nt, Nx, W, r, bf16=False, c0=0.2, phi=phi, g=0.5)
    z = f(0.7 * r1 - 1.9 * r2)
    assert np.abs(z - (0.7 * f(r1) - 1.9 * f(r2))).max() < 1e-12 * np.abs(z).max()


def test_scale_equivariance():
    Nt, Nx = 4, 6
    rng = np.random.default_rng(1)
    W = (0.3 * rng.standard_normal(O.CNN_NW)).astype(np.float32)
    r = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    phi = rng.uniform(-1, 1, (Nt, Nx))
    z1 = O.cnn_forward(Nt, Nx, W, r, bf16=False, phi=phi, g=0.5)
    z2 = O.cnn_forward(Nt, Nx, W, 1e-6 * r, bf16=False, phi=phi, g=0.5)
    assert np.abs(z2 - 1e-6 * z1).max() < 1e-12 * np.abs(z1).max()


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_finite_differences(seed):
    Nt, Nx = 4, 6
    V = Nt * Nx
    rng = np.random.default_rng(seed)
    W = (0.2 * rng.standard_normal(O.CNN_NW)).astype(np.float32)
    p = O.params(Nt=Nt, Nx=Nx, M=0.3, g=0.5)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    e = rng.standard_normal((V, 4)) + 1j * rng.standard_normal((V, 4))
    r = O.normal_op(p, phi, e, 1)
    c0 = O.cnn_c0(1.0, 0.3)
    L, g = O.cnn_loss_grad(p, phi, W, r, e, c0=c0)
    # one weight in every block (W1, W2, W3, G1, G2), including first/last entries
    for i in [0, 5, 1151, 1152, 2000, 3455, 3456, 4000, 4607, 4608, 4700, 4751, 4752, 4895] + \
            list(rng.choice(O.CNN_NW, 6, replace=False)):
        h = np.float32(1e-3)
        Wp, Wm = W.copy(), W.copy()
        Wp[i] += h
        Wm[i] -= h
        fd = (O.cnn_loss_grad(p, phi, Wp, r, e, c0=c0)[0] - O.cnn_loss_grad(p, phi, Wm, r, e, c0=c0)[0]) / (float(Wp[i]) - float(Wm[i]))
        assert abs(g[i] - fd) <= 1e-4 * max(1.0, abs(fd)), (i, g[i], fd)
    # with z = 0 (all weights zero, no skip path) the loss is exactly 1 for a consistent
    # pair: |A e|^2 / |r|^2 with r = A e
    L0, _ = O.cnn_loss_grad(p, phi, np.zeros(O.CNN_NW, np.float32), r, e)
    assert abs(L0 - 1.0) < 1e-12
    # the skip path alone: L = |r - c0 A r|^2 / |r|^2
    Ar = O.normal_op(p, phi, r, 1).reshape(-1)
    rv = r.reshape(-1)
    L1, _ = O.cnn_loss_grad(p, phi, np.zeros(O.CNN_NW, np.float32), r, e, c0=c0)
    assert abs(L1 - np.vdot(rv - c0 * Ar, rv - c0 * Ar).real / np.vdot(rv, rv).real) < 1e-12


@pytest.mark.parametrize("t0,x0", [(0, 0), (5, 9)])
def test_tap_orientation_and_channel_order(t0, x0):
    """Pins the weight layout [tap][in][out], tap = (dt+1)*3 + (dx+1), out(site) =
    sum W * in(site + (dt, dx)), periodic wrap, and the channel map 2c / 2c+1 = Re / Im
    of component c: a delta in channel 0 (Re of component 0) at (t0, x0) through a
    (dt, dx) = (+1, 0) tap 0 -> 5, a centre tap 5 -> 7 and a (0, +1) tap 7 -> 2 lands in
    channel 2 (Re of component 1) at (t0 - 1, x0 - 1) only, with the value a b v (the
    gates are 1 without a field)."""
    Nt, Nx = 8, 12
    V = Nt * Nx
    a, b = 0.01, 0.7
    W = np.zeros(O.CNN_NW, np.float64)
    W[((1 + 1) * 3 + (0 + 1)) * 8 * 16 + 0 * 16 + 5] = a           # W1 tap (+1, 0), in 0 -> out 5
    W[N1 + (4 * 16 + 5) * 16 + 7] = 1.0                             # W2 centre, 5 -> 7
    W[N1 + N2 + ((0 + 1) * 3 + (1 + 1)) * 16 * 8 + 7 * 8 + 2] = b   # W3 tap (0, +1), 7 -> 2
    W = W.astype(np.float32)
    r = np.zeros((V, 4), np.complex128)
    v = 2.5
    r[t0 * Nx + x0, 0] = v
    z = O.cnn_forward(Nt, Nx, W, r, bf16=False, c0=0.0)
    want = np.zeros((V, 4), np.complex128)
    tp, xp = (t0 - 1) % Nt, (x0 - 1) % Nx
    af, bf = float(np.float32(a)), float(np.float32(b))                 # the weights as stored
    want[tp * Nx + xp, 1] = bf * af * v
    assert np.abs(z - want).max() < 1e-12 * abs(want).max()


@pytest.mark.parametrize("t0,x0", [(0, 0), (5, 11)])
def test_gate_orientation_and_channel(t0, x0):
    """Pins the gate layout G[tap][out] with q = g phi read at site + (dt, dx): a delta
    of r (channel 1, Im of component 0) at P = (t0, x0) runs through centre taps
    1 -> 5 -> 7 -> 1.  A field delta phi0 at P + (-1, +1) seen through G1's (-1, +1) tap
    on channel 5, and a second delta phi1 at P + (0, -1) seen through G2's (0, -1) tap on
    channel 7, scale the output at P by (1 + tanh(w1 g phi0)) (1 + tanh(w2 g phi1)); a
    gate tap on any other channel or offset leaves it unchanged."""
    Nt, Nx = 8, 12
    V = Nt * Nx
    g, w1, w2, a = 0.5, 0.8, -1.3, 0.25
    W = np.zeros(O.CNN_NW, np.float64)
    W[(4 * 8 + 1) * 16 + 5] = a                                     # W1 centre, 1 -> 5
    W[N1 + (4 * 16 + 5) * 16 + 7] = 1.0                             # W2 centre, 5 -> 7
    W[N1 + N2 + (4 * 16 + 7) * 8 + 1] = 1.0                         # W3 centre, 7 -> 1
    W[N1 + N2 + N3 + ((-1 + 1) * 3 + (1 + 1)) * 16 + 5] = w1        # G1 tap (-1, +1), channel 5
    W[N1 + N2 + N3 + NG + ((0 + 1) * 3 + (-1 + 1)) * 16 + 7] = w2   # G2 tap (0, -1), channel 7
    W[N1 + N2 + N3 + (4 * 16) + 6] = 3.0                            # G1 centre on an unused channel
    W = W.astype(np.float32)
    phi = np.zeros((Nt, Nx))
    phi0, phi1 = 0.9, -0.4
    phi[(t0 - 1) % Nt, (x0 + 1) % Nx] = phi0
    phi[t0, (x0 - 1) % Nx] = phi1
    r = np.zeros((V, 4), np.complex128)
    v = -1.5
    r[t0 * Nx + x0, 0] = 1j * v
    z = O.cnn_forward(Nt, Nx, W, r, bf16=False, c0=0.0, phi=phi, g=g)
    want = np.zeros((V, 4), np.complex128)
    f1, f2 = float(np.float32(w1)), float(np.float32(w2))
    want[t0 * Nx + x0, 0] = 1j * float(np.float32(a)) * v * (1 + np.tanh(f1 * g * phi0)) * (1 + np.tanh(f2 * g * phi1))
    assert np.abs(z - want).max() < 1e-12 * abs(want).max()
