"""Long-run statistical agreement of the GPU Langevin chain with the CPU oracle.

Independent chains (different seeds) of the same model must give the same equilibrium
expectation values within statistical errors.  The GPU runs 32 replicas in one batched
context; the oracle runs 6 chains.  Each chain's time series is analysed with Wolff's
Gamma method (yukawacolourednoise_amd.analysis.ComputeStatistics), the chain means are
averaged, and GPU and oracle must agree within 5 combined standard errors for <phi^2>,
<|sigma|> and V <sigma^2>.  The same comparison against an oracle run at a shifted m^2
must fail, so the test can tell the two apart."""
import sys

import numpy as np
import pytest

import oracle as O
import yukawacolourednoise_amd.analysis as A

pytestmark = pytest.mark.gpu

NT = NX = 8
V = NT * NX
PHYS = dict(m2=-0.2, lam=0.5, g=0.3, M=0.3, cutFraction=0.5, epsBar=0.05, adaptive=0, cgTol=1e-10,
            cgTolMode=1, cgMaxIter=10000)
THERM, STEPS, EVERY = 500, 3000, 2


def _observables(phi):
    """phi[..., Nt, Nx] -> (<phi^2>, |sigma|, V sigma^2) per configuration."""
    f = phi.reshape(phi.shape[:-2] + (V,))
    s = f.mean(axis=-1)
    return np.stack([np.mean(f * f, axis=-1), np.abs(s), V * s * s], axis=-1)


def _combine(series):
    """series[chain, sample, obs] -> (mean, error) per observable from per-chain Gamma-method
    errors of independent chains."""
    K, _, nobs = series.shape
    out = []
    for o in range(nobs):
        st = [A.ComputeStatistics(series[k, :, o]) for k in range(K)]
        m = np.mean([s.average for s in st])
        e = np.sqrt(np.sum([s.sigmaF ** 2 for s in st])) / K
        out.append((m, e))
    return out


def _oracle_series(m2, seeds):
    O.set_threads(1)
    out = []
    for sd in seeds:
        p = O.params(Nt=NT, Nx=NX, **{**PHYS, "m2": m2}, seed=int(sd))
        phi = O.hot_start(int(sd) + 7, NT, NX)
        obs = []
        for step in range(THERM + STEPS):
            phi, _, st = O.langevin_step(p, phi, step)
            assert st["converged"] == 1
            if step >= THERM and (step - THERM) % EVERY == 0:
                obs.append(_observables(phi))
        out.append(np.array(obs))
    return np.array(out)


def test_long_run_observables_agree_with_oracle(gpu):
    Y = gpu
    R = 32
    p = Y.default_params(Nt=NT, Nx=NX, nReplicas=R, **PHYS)
    print(f"# GPU: {R} replicas of {NT}x{NX}, {THERM}+{STEPS} steps", flush=True)
    with Y.Context(p) as c:
        c.set_replicas(np.full(R, PHYS["m2"]), np.arange(100, 100 + R, dtype=np.uint64))
        c.hot_start(1)
        c.langevin_run(0, THERM)
        obs = []
        for k in range(0, STEPS, EVERY):
            obs.append(_observables(c.get_field()))
            st = c.langevin_run(THERM + k, EVERY)
            assert st["converged"] == 1
    gpu_series = np.transpose(np.array(obs), (1, 0, 2))          # [replica, sample, obs]
    print("# oracle chains", flush=True)
    orc = _oracle_series(PHYS["m2"], range(900, 906))
    shifted = _oracle_series(PHYS["m2"] + 0.3, range(950, 956))
    g, o, s = _combine(gpu_series), _combine(orc), _combine(shifted)
    names = ["<phi^2>", "<|sigma|>", "V<sigma^2>"]
    for n, (mg, eg), (mo, eo), (ms, es) in zip(names, g, o, s):
        z = abs(mg - mo) / np.hypot(eg, eo)
        zs = abs(mg - ms) / np.hypot(eg, es)
        print(f"# {n}: gpu {mg:.5f} +- {eg:.5f}  oracle {mo:.5f} +- {eo:.5f}  (z = {z:.2f}); "
              f"shifted m^2 {ms:.5f} +- {es:.5f} (z = {zs:.1f})", flush=True)
        assert z < 5.0, n
    # the shifted-m^2 run must be told apart on the most precise observable
    zs_phi2 = abs(g[0][0] - s[0][0]) / np.hypot(g[0][1], s[0][1])
    assert zs_phi2 > 5.0, zs_phi2
