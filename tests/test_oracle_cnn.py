"""Pins the CPU FP64 oracle of the learned preconditioner: the analytic gradient of the
energy-norm loss against central finite differences, scale equivariance, the identity
network, and bfloat16 round-to-nearest-even."""
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


def test_identity_network():
    """Centre taps as identity and tiny weights elsewhere: net(u) ~ u in the linear
    range of tanh (W1 = a I, W2 = I, W3 = I/a)."""
    Nt, Nx = 4, 6
    W = np.zeros(O.CNN_NW, np.float32)
    a = 1e-3
    for ci in range(8):
        W[(4 * 8 + ci) * 16 + ci] = a
    for ci in range(16):
        W[9 * 8 * 16 + (4 * 16 + ci) * 16 + ci] = 1.0
    for ci in range(8):
        W[9 * 8 * 16 + 9 * 16 * 16 + (4 * 16 + ci) * 8 + ci] = 1.0 / a
    r = np.random.default_rng(0).standard_normal((Nt * Nx, 4)) + 0j
    z = O.cnn_forward(Nt, Nx, W, r, bf16=False)
    assert np.abs(z - r).max() < 1e-5 * np.abs(r).max()


def test_scale_equivariance():
    Nt, Nx = 4, 6
    rng = np.random.default_rng(1)
    W = (0.3 * rng.standard_normal(O.CNN_NW)).astype(np.float32)
    r = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    z1 = O.cnn_forward(Nt, Nx, W, r, bf16=False)
    z2 = O.cnn_forward(Nt, Nx, W, 1e-6 * r, bf16=False)
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
    # one weight in every layer, including the first tap and channel edges
    for i in [0, 5, 1151, 1152, 2000, 3455, 3456, 4000, 4607] + list(rng.choice(O.CNN_NW, 6, replace=False)):
        h = np.float32(1e-3)
        Wp, Wm = W.copy(), W.copy()
        Wp[i] += h
        Wm[i] -= h
        fd = (O.cnn_loss_grad(p, phi, Wp, r, e, c0=c0)[0] - O.cnn_loss_grad(p, phi, Wm, r, e, c0=c0)[0]) / (float(Wp[i]) - float(Wm[i]))
        assert abs(g[i] - fd) <= 1e-4 * max(1.0, abs(fd)), (i, g[i], fd)
    # with z = 0 (all weights zero, no skip path) the loss is exactly 1 for a consistent pair
    L0, _ = O.cnn_loss_grad(p, phi, np.zeros(O.CNN_NW, np.float32), r, e)
    assert abs(L0 - 1.0) < 1e-12
    # the skip path alone: L = <e - c0 r, A (e - c0 r)> / <e, r>
    d = (e - c0 * r).reshape(-1)
    Ad = O.normal_op(p, phi, d.reshape(-1, 4), 1).reshape(-1)
    L1, _ = O.cnn_loss_grad(p, phi, np.zeros(O.CNN_NW, np.float32), r, e, c0=c0)
    assert abs(L1 - np.vdot(d, Ad).real / np.vdot(e.reshape(-1), r.reshape(-1)).real) < 1e-12


@pytest.mark.parametrize("t0,x0", [(0, 0), (5, 9)])
def test_tap_orientation_and_channel_order(t0, x0):
    """Pins the weight layout [tap][in][out], tap = (dt+1)*3 + (dx+1), out(site) =
    sum W * in(site + (dt, dx)), periodic wrap, and the channel map 2c / 2c+1 = Re / Im
    of component c: a delta in channel 0 (Re of component 0) at (t0, x0) through a
    (dt, dx) = (+1, 0) tap 0 -> 5, a centre tap 5 -> 7 and a (0, +1) tap 7 -> 2 lands in
    channel 2 (Re of component 1) at (t0 - 1, x0 - 1) only, with the value fixed by the
    two tanh layers."""
    Nt, Nx = 8, 12
    V = Nt * Nx
    a, b = 0.01, 0.7
    W = np.zeros(O.CNN_NW, np.float64)
    n1, n2 = 9 * 8 * 16, 9 * 16 * 16
    W[((1 + 1) * 3 + (0 + 1)) * 8 * 16 + 0 * 16 + 5] = a           # W1 tap (+1, 0), in 0 -> out 5
    W[n1 + (4 * 16 + 5) * 16 + 7] = 1.0                             # W2 centre, 5 -> 7
    W[n1 + n2 + ((0 + 1) * 3 + (1 + 1)) * 16 * 8 + 7 * 8 + 2] = b   # W3 tap (0, +1), 7 -> 2
    W = W.astype(np.float32)
    r = np.zeros((V, 4), np.complex128)
    v = 2.5
    r[t0 * Nx + x0, 0] = v
    z = O.cnn_forward(Nt, Nx, W, r, bf16=False, c0=0.0)
    s = v / np.sqrt(8 * V)
    want = np.zeros((V, 4), np.complex128)
    tp, xp = (t0 - 1) % Nt, (x0 - 1) % Nx
    af, bf = float(np.float32(a)), float(np.float32(b))                 # the weights as stored
    want[tp * Nx + xp, 1] = s * bf * np.tanh(np.tanh(af * v / s))
    assert np.abs(z - want).max() < 1e-12 * abs(want).max()
