"""CPU runs of the oracle itself (no GPU): the reference's CPU-runnable workload in
full, and the oracle under the host address and undefined-behaviour sanitizers.

- 32^2 (m^2 = -0.2, lambda = 0.5, M = 0.3, g = 0.3, s = 0.5, eps = 0.01 fixed, cold
  start, seed 1234, 2000 steps, every solve to true residual 1e-10): every solve
  converges, the trajectory reproduces the committed snapshots every 500 steps
  (tests/golden/c1_32x32_2000_steps.npz, written by tests/golden/make_golden.py) and
  the steps/s on the host is printed.
- oracle/asan_check.c, built with -fsanitize=address,undefined, runs every oracle entry
  point on 8x8, 6x10, 2x12 and 16x16 and must exit cleanly."""
import os
import subprocess
import time

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "c1_32x32_2000_steps.npz")


def test_c1_full_run_on_cpu():
    g = np.load(GOLDEN)
    kw = {k: g[k].item() for k in ("Nt", "Nx", "m2", "lam", "g", "M", "r", "cutFraction", "driftMode", "epsBar",
                                    "adaptive", "seed", "cgTol", "cgTolMode", "cgMaxIter")}
    p = O.params(**kw)
    threads = O.set_threads(1)
    phi = np.zeros((kw["Nt"], kw["Nx"]))
    every, steps = int(g["every"]), int(g["steps"])
    t0 = time.perf_counter()
    iters = 0
    for step in range(steps):
        phi, _, st = O.langevin_step(p, phi, step)
        assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, (step, st)
        assert st["iters"] == g["cg_iters"][step], step
        iters += st["iters"]
        if (step + 1) % every == 0:
            ref = g["snapshots"][(step + 1) // every - 1]
            assert np.abs(phi - ref).max() <= 1e-13 * np.abs(ref).max(), step
    wall = time.perf_counter() - t0
    print(f"\n32^2 workload in full on the CPU oracle ({threads} thread): {steps} steps in {wall:.1f} s, "
          f"{steps / wall:.0f} steps/s, {iters / steps:.1f} CG iterations per solve")


def test_oracle_under_address_and_ub_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "asan_check")], env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "clean" in r.stdout and "Sanitizer" not in r.stderr, r.stderr
