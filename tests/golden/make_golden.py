"""Writes the golden fixtures under tests/golden/ from the CPU FP64 oracle.

Run from the repository root: python tests/golden/make_golden.py
Each fixture is a small .npz (float64 arrays + scalar parameters); manifest.json
records the SHA-256 of every file so a silent change to the oracle is caught."""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

COMMON = dict(m2=-0.2, lam=0.5, g=0.3, M=0.3, r=1.0, cutFraction=0.5, driftMode=0, seed=1234,
              cgTol=1e-10, cgTolMode=1, cgMaxIter=10000)


def step_fixture(name, Nt, Nx, adaptive, epsBar, init_seed, step, **over):
    kw = dict(COMMON, Nt=Nt, Nx=Nx, adaptive=adaptive, epsBar=epsBar)
    kw.update(over)
    p = O.params(**kw)
    phi0 = O.hot_start(init_seed, Nt, Nx)
    noise = O.coloured_noise(p, step)
    eta = O.fermion_noise(p.seed, 0, step, Nt * Nx)
    drift_f, st = O.fermion_drift(p, phi0, step)
    phi1, drift_tot, eps, mx = O.bosonic_update(p, phi0, drift_f, noise)
    full, eps2, _ = O.langevin_step(p, phi0, step)
    assert np.array_equal(full, phi1) and eps == eps2
    np.savez(os.path.join(HERE, name), phi0=phi0, noise=noise, eta=eta, drift_fermion=drift_f,
             drift_total=drift_tot, phi1=phi1, eps=eps, cg_iters=st["iters"], step=step,
             init_seed=init_seed, **{k: v for k, v in kw.items()})


def filter_fixture(name, Nt, Nx, s):
    xi = np.random.default_rng(Nt * 1000 + Nx).standard_normal((Nt, Nx))
    np.savez(os.path.join(HERE, name), xi=xi, out=O.coloured_filter(xi, s), cutFraction=s)


def dirac_fixture(name, Nt, Nx):
    p = O.params(Nt=Nt, Nx=Nx, M=0.05, g=0.5, r=1.0)
    rng = np.random.default_rng(99)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    psi = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    np.savez(os.path.join(HERE, name), phi=phi, psi=psi, Dpsi=O.dirac(p, phi, psi, False),
             Ddagpsi=O.dirac(p, phi, psi, True), M=0.05, g=0.5, r=1.0)


def c1_trajectory_fixture(name, steps=2000, every=500):
    """the 32^2 CPU-runnable workload in full: cold start, 2000 steps, snapshots every 500"""
    kw = dict(COMMON, Nt=32, Nx=32, adaptive=0, epsBar=0.01)
    p = O.params(**kw)
    phi = np.zeros((32, 32))
    snaps, iters = [], []
    for step in range(steps):
        phi, _, st = O.langevin_step(p, phi, step)
        iters.append(st["iters"])
        if (step + 1) % every == 0:
            snaps.append(phi.copy())
    np.savez(os.path.join(HERE, name), snapshots=np.array(snaps), every=every, steps=steps,
             cg_iters=np.array(iters, np.int32), **{k: v for k, v in kw.items()})


def main():
    O.set_threads(1)
    c1_trajectory_fixture("c1_32x32_2000_steps.npz")
    step_fixture("step_16x16_fixed.npz", 16, 16, 0, 0.01, 7, 0)
    # adaptive eps depends on max|drift|: fixed CG iteration count so both sides run the
    # same recurrence and eps is reproducible to rounding
    step_fixture("step_8x8_adaptive.npz", 8, 8, 1, 0.01, 3, 5, cgTol=0.0, cgTolMode=0, cgMaxIter=30)
    step_fixture("step_6x32_fixed.npz", 6, 32, 0, 0.005, 11, 2)
    filter_fixture("filter_6x32_s05.npz", 6, 32, 0.5)
    filter_fixture("filter_32x32_s025.npz", 32, 32, 0.25)
    dirac_fixture("dirac_8x12.npz", 8, 12)
    man = {}
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            man[f] = hashlib.sha256(open(os.path.join(HERE, f), "rb").read()).hexdigest()
    json.dump(man, open(os.path.join(HERE, "manifest.json"), "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
