"""Whole Langevin trajectories at their full lengths, not single steps.

With the same noise, Langevin dynamics contracts: two runs whose drifts differ at the level
of the CG tolerance stay together.  The oracle run at tolerance 1e-10 and at 1e-12 differs
by 2e-12 after 200 steps of the 32^2 workload and by 6e-13 after 2000.  So a GPU trajectory
can be held to the oracle's over its full length, for every solver.

- 32^2, the reference's own CPU-runnable case: m^2 = -0.2, lambda = 0.5, M = 0.3, g = 0.3,
  s = 0.5, eps = 0.01 fixed, cold start, seed 1234, 2000 steps, CG to 1e-10.  Each solver
  (plain FP64, mixed precision, even-odd FP64, even-odd pipelined mixed) is compared with
  the oracle's plain FP64 trajectory every 500 steps, and so are FP64 and even-odd mixed
  with the reference's adaptive step size, and flexible CG with the learned preconditioner
  after 500 steps of online training.
- 512^2 mixed-precision comparison: M = 0.05, g = 0.5, m^2 = -0.4, lambda = 1, s = 0.5,
  eps = 0.005, hot start (init seed 7), 2000 steps.  FP32 inner iterations with FP64
  reliable updates, plain and on the even-odd Schur complement, against the FP64 CG
  trajectory (itself held to the oracle step by step at 256^2 and 1024^2 elsewhere).
- The 64-chain m^2 scan, 500 steps batched (512 systems per CG launch) against four of its
  chains run alone.
- 256^2 at s = 1 and 0.25, 1000 steps: mixed and even-odd mixed against FP64.
- The 1024^2 headline over 5000 steps: even-odd mixed and the learned preconditioner
  against plain FP64."""
import os

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import mk

pytestmark = pytest.mark.gpu

C1 = dict(Nt=32, Nx=32, m2=-0.2, lam=0.5, g=0.3, M=0.3, cutFraction=0.5, epsBar=0.01, adaptive=0, seed=1234,
          cgTol=1e-10, cgTolMode=1, cgMaxIter=10000)
C3 = dict(Nt=512, Nx=512, M=0.05, g=0.5, m2=-0.4, lam=1.0, cutFraction=0.5, epsBar=0.005, adaptive=0, seed=1234,
          cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
STEPS, EVERY = 2000, 500
METHODS = {"fp64": (0, 0), "mixed": (1, 0), "eo-fp64": (0, 1), "eo-mixed": (1, 1)}


def l2rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _set_method(Y, p, method):
    mixed, eo = METHODS[method]
    p.cgPrecision = Y.YCN_CG_MIXED if mixed else p.cgPrecision
    p.evenOdd = eo


_oracle_c1 = {}


def _oracle_snapshots(adaptive=0):
    """the oracle's 32^2 trajectory from a cold start, every EVERY steps (computed once
    per step-size mode)"""
    if adaptive not in _oracle_c1:
        O.set_threads(min(8, os.cpu_count() or 1))
        orc = O.params(**{**C1, "adaptive": adaptive})
        phi = np.zeros((C1["Nt"], C1["Nx"]))
        snaps = {}
        for step in range(STEPS):
            phi, _, st = O.langevin_step(orc, phi, step)
            assert st["converged"] == 1
            if (step + 1) % EVERY == 0:
                snaps[step + 1] = phi.copy()
        _oracle_c1[adaptive] = snaps
    return _oracle_c1[adaptive]


@pytest.mark.parametrize("method", list(METHODS))
def test_32_full_trajectory(gpu, method):
    _trajectory_32(gpu, method, adaptive=0)


@pytest.mark.parametrize("method", ["fp64", "eo-mixed"])
def test_32_adaptive_step_trajectory(gpu, method):
    """the reference's adaptive step (eps = epsBar 30 / max |drift|, clamped) over the
    same 2000 steps: eps depends on the drift's maximum, so the solvers' drifts must agree
    to the CG tolerance at every step for the trajectories to stay together"""
    _trajectory_32(gpu, method, adaptive=1)


def _trajectory_32(Y, method, adaptive):
    print(f"# 32^2 trajectory, {STEPS} steps, {method}, adaptive={adaptive}", flush=True)
    ref = _oracle_snapshots(adaptive)
    p, _ = mk(Y, **{**C1, "adaptive": adaptive})
    _set_method(Y, p, method)
    with Y.Context(p) as c:
        c.set_field(np.zeros((C1["Nt"], C1["Nx"])))
        for k in range(0, STEPS, EVERY):
            st = c.langevin_run(k, EVERY)
            assert st["converged"] == 1, st
            e = l2rel(c.get_field(), ref[k + EVERY])
            print(f"#   step {k + EVERY}: |phi - phi_oracle| / |phi_oracle| = {e:.2e}, "
                  f"{st['iters_mean']:.1f} iterations per solve", flush=True)
            assert e < 1e-9, (k + EVERY, e)


def test_32_learned_preconditioner_trajectory(gpu):
    """Flexible CG with the BF16 CNN preconditioner: trained online during the first 500
    steps (whose solves are plain FP64 CG, so the chain is the oracle's), then 1500 steps
    of flexible CG.  Convergence is judged on the FP64 true residual of the plain operator,
    so the trajectory stays the oracle's whatever the network learned."""
    Y = gpu
    print(f"# 32^2 trajectory, {STEPS} steps, learned preconditioner", flush=True)
    ref = _oracle_snapshots()
    p, _ = mk(Y, **C1)
    p.cgPrecision = Y.YCN_CG_MIXED
    p.precond = Y.YCN_PRECOND_CNN
    p.reliableEvery = 50
    with Y.Context(p) as c:
        c.set_field(np.zeros((C1["Nt"], C1["Nx"])))
        info = c.precond_train(0, EVERY, pairs_per_solve=4, lr=1e-3)
        e = l2rel(c.get_field(), ref[EVERY])
        print(f"#   trained on {info['pairs']} pairs (loss {info['loss_first']:.3f} -> {info['loss_last']:.3f}); "
              f"step {EVERY}: {e:.2e}", flush=True)
        assert e < 1e-9
        for k in range(EVERY, STEPS, EVERY):
            st = c.langevin_run(k, EVERY)
            assert st["converged"] == 1, st
            e = l2rel(c.get_field(), ref[k + EVERY])
            print(f"#   step {k + EVERY}: {e:.2e}, {st['iters_mean']:.1f} iterations per solve", flush=True)
            assert e < 1e-9, (k + EVERY, e)


def test_512_mixed_precision_trajectories(gpu):
    Y = gpu
    N = C3["Nt"]
    phi0 = O.hot_start(7, N, N)
    fields, iters = {}, {}
    for method in ("fp64", "mixed", "eo-mixed"):
        print(f"# 512^2 trajectory, {STEPS} steps, {method}", flush=True)
        p, _ = mk(Y, **C3)
        _set_method(Y, p, method)
        with Y.Context(p) as c:
            c.set_field(phi0)
            st = c.langevin_run(0, STEPS)
            assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
            fields[method] = c.get_field()
            iters[method] = st["iters_mean"]
        print(f"#   {iters[method]:.1f} iterations per solve", flush=True)
    assert np.all(np.isfinite(fields["fp64"]))
    for method in ("mixed", "eo-mixed"):
        e = l2rel(fields[method], fields["fp64"])
        print(f"#   {method} vs fp64 after {STEPS} steps: {e:.2e}", flush=True)
        assert e < 1e-9, method


def test_chain_scan_long_run_batched_equals_alone(gpu):
    """The 64-chain m^2 scan (64 chains of 64^2, m^2 = -1 .. 0.5, seeds 0..63, 8 noise
    vectors per drift) over 500 steps with the even-odd mixed solver, batched (512 systems
    per CG launch, per-system pauses, reliable updates and step state), against four of
    its chains each run alone in a context of one replica: batching must not couple the
    chains, so the fields agree to the CG tolerance after 500 steps."""
    Y = gpu
    R, N, n = 64, 64, 500
    base = dict(Nt=N, Nx=N, lam=0.5, g=0.3, M=0.3, r=1.0, cutFraction=0.5, driftMode=0, epsBar=0.01, adaptive=0,
                cgTol=1e-10, cgTolMode=1, cgMaxIter=20000, nNoise=8)
    m2 = -1.0 + 1.5 * np.arange(R) / 63.0
    seeds = np.arange(R, dtype=np.uint64)
    phi0 = np.stack([O.hot_start(10 + k, N, N) for k in range(R)])

    def ctx(nrep):
        p = Y.default_params(nReplicas=nrep, **base)
        p.evenOdd = 1
        p.cgPrecision = Y.YCN_CG_MIXED
        return Y.Context(p)

    print(f"# chain scan, {R} chains batched, {n} steps", flush=True)
    with ctx(R) as c:
        c.set_replicas(m2, seeds)
        c.set_field(phi0)
        st = c.langevin_run(0, n)
        assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
        g = c.get_field()
    for k in (0, 21, 42, 63):
        with ctx(1) as c:
            c.set_replicas(m2[k:k + 1], seeds[k:k + 1])
            c.set_field(phi0[k])
            st = c.langevin_run(0, n)
            assert st["converged"] == 1, st
            e = l2rel(g[k], c.get_field())
        print(f"#   chain {k} (m^2 = {m2[k]:.3f}): batched vs alone {e:.2e}", flush=True)
        assert e < 1e-9, k


@pytest.mark.parametrize("s", [1.0, 0.25])
def test_256_trajectories(gpu, s):
    """The 256^2 configuration (m^2 = -0.2, lambda = 0.5, M = 0.3, g = 0.3, hot start with
    init seed 7, eps = 0.005) over 1000 steps at the unfiltered (s = 1) and the most
    filtered (s = 0.25) coloured noise: mixed precision and even-odd mixed against the FP64
    trajectory (itself held to the oracle over its first steps elsewhere)."""
    Y = gpu
    N, n = 256, 1000
    base = dict(Nt=N, Nx=N, m2=-0.2, lam=0.5, g=0.3, M=0.3, cutFraction=s, epsBar=0.005, adaptive=0, seed=1234,
                cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    phi0 = O.hot_start(7, N, N)
    fields = {}
    for method in ("fp64", "mixed", "eo-mixed"):
        print(f"# 256^2 s={s} trajectory, {n} steps, {method}", flush=True)
        p, _ = mk(Y, **base)
        _set_method(Y, p, method)
        with Y.Context(p) as c:
            c.set_field(phi0)
            st = c.langevin_run(0, n)
            assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
            fields[method] = c.get_field()
        print(f"#   {st['iters_mean']:.1f} iterations per solve", flush=True)
    for method in ("mixed", "eo-mixed"):
        e = l2rel(fields[method], fields["fp64"])
        print(f"#   {method} vs fp64 after {n} steps: {e:.2e}", flush=True)
        assert e < 1e-9, method


@pytest.mark.timeout(1200)
def test_1024_headline_5000_steps(gpu):
    """The 1024^2 headline workload over 5000 steps from a hot start (init seed 7): the
    even-odd pipelined mixed solver and the learned preconditioner against the
    reference's solver, plain FP64 CG, on the same noise.  The learned chain's first 125
    steps are its online training (about 2000 pairs), whose solves are plain FP64 CG;
    flexible CG with the trained network runs the remaining 4875.  Progress every 500
    steps."""
    Y = gpu
    N, n, every, train = 1024, 5000, 500, 125
    base = dict(Nt=N, Nx=N, M=0.05, g=0.5, m2=-0.4, lam=1.0, cutFraction=0.5, epsBar=0.005, adaptive=0, seed=1234,
                cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    phi0 = O.hot_start(7, N, N)
    import time
    fields, iters, secs = {}, {}, {}
    for method in ("eo-mixed", "learned", "fp64"):
        p, _ = mk(Y, **base)
        if method == "learned":
            p.cgPrecision = Y.YCN_CG_MIXED
            p.precond = Y.YCN_PRECOND_CNN
        else:
            _set_method(Y, p, method)
        tot, wall, k = 0.0, 0.0, 0
        with Y.Context(p) as c:
            c.set_field(phi0)
            if method == "learned":
                info = c.precond_train(0, train, pairs_per_solve=16, lr=1e-3)
                print(f"# learned: trained on {info['pairs']} pairs in the first {train} steps", flush=True)
                k = train
            while k < n:
                m = min(every - k % every, n - k)
                c.sync()
                t0 = time.perf_counter()
                st = c.langevin_run(k, m)
                c.sync()
                dt = time.perf_counter() - t0
                wall += dt
                assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
                tot += st["iters_mean"] * m
                k += m
                print(f"# {method}: {k} steps, {st['iters_mean']:.1f} iterations per solve, "
                      f"{m / dt:.1f} steps/s", flush=True)
            fields[method] = c.get_field()
        steps = n - (train if method == "learned" else 0)
        iters[method], secs[method] = tot / steps, wall / steps
    for method in ("eo-mixed", "learned"):
        e = l2rel(fields[method], fields["fp64"])
        print(f"# after {n} steps: |phi_{method} - phi_fp64| / |phi_fp64| = {e:.2e}; iterations per solve "
              f"{iters[method]:.1f} vs {iters['fp64']:.1f} (plain FP64); "
              f"{1 / secs[method]:.1f} vs {1 / secs['fp64']:.1f} steps/s ({secs['fp64'] / secs[method]:.2f}x)",
              flush=True)
        assert e < 1e-8, method
