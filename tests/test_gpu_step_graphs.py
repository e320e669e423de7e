"""Graph-captured Langevin steps (ycn_langevin_run with the even-odd mixed solver).

Each step after a context's first solve replays a captured graph (the solve's setup, a
fixed chunk schedule of the pipelined FP32 iteration with its FP64 reliable updates sized
from the previous solve's rounds, and the drift) between the coloured noise and the update.
The host waits once per segment of up to 16 steps (ycn_host_syncs counts every wait), and
never inside a step; the reference waits on the device about 14 times per step and copies
the step size to the host each time (code/src/main.cu:305-315, FermionicDrift.cu:61-78,
CGsolver.cu:47,58).

Every run is compared with the same run of single eagerly launched steps
(YCN_STEP_GRAPHS=0, which polls CG convergence inside every step): the fields agree to
1e-9 (the two place their reliable updates differently, so the iterates agree to the
solver tolerance, not bit for bit), every solve
meets the FP64 true-residual tolerance, and the host waits only at segment ends.  A run
whose field is swapped for a harder one after the first solve outlasts the captured
schedule: the host finishes the halted step, and the trajectory is still the eager one."""
import math
import os

import numpy as np
import pytest

from test_gpu_parity import mk

pytestmark = pytest.mark.gpu

SEG = 16   # captured steps per host wait
C3 = dict(M=0.05, g=0.5, m2=-0.4, lam=1.0, cutFraction=0.5, epsBar=0.005, seed=1234, cgTol=1e-10, cgTolMode=1,
          cgMaxIter=20000)


def l2rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _ctx(Y, monkeypatch, graphs, n, reps, adaptive, every=None):
    p, _ = mk(Y, Nt=n, Nx=n, nReplicas=reps, adaptive=adaptive, **C3)
    p.cgPrecision = Y.YCN_CG_MIXED
    p.evenOdd = 1
    if every is not None:
        p.reliableEvery = every
    with monkeypatch.context() as m:
        if graphs:
            m.delenv("YCN_STEP_GRAPHS", raising=False)
        else:
            m.setenv("YCN_STEP_GRAPHS", "0")
        c = Y.Context(p)
    c.set_replicas(np.full(reps, C3["m2"]), np.arange(11, 11 + reps, dtype=np.uint64))
    return c


def _run(c, fields, nsteps):
    """one eager first step from fields[0], then (if given) the field swapped for
    fields[1] and nsteps steps in one run; returns (field, stats, host waits of the run,
    halted steps of the run)"""
    c.set_field(fields[0])
    st = c.langevin_run(0, 1)
    assert st["converged"] == 1, st
    if len(fields) > 1:
        c.set_field(fields[1])
    s0, h0 = c.host_syncs()
    st = c.langevin_run(1, nsteps)
    s1, h1 = c.host_syncs()
    return c.get_field(), st, s1 - s0, h1 - h0


@pytest.mark.parametrize("n,reps,adaptive,nsteps", [(256, 2, 0, 40), (256, 2, 1, 40), (128, 4, 0, 70),
                                                    (1024, 1, 0, 33)])
def test_captured_steps_match_eager_steps(gpu, monkeypatch, n, reps, adaptive, nsteps):
    Y = gpu
    print(f"# {n}^2, {reps} replicas, adaptive={adaptive}, {nsteps} steps", flush=True)
    hot = np.random.default_rng(n + reps).uniform(-1.0, 1.0, (reps, n, n))   # a hot start
    out = {}
    for graphs in (True, False):
        c = _ctx(Y, monkeypatch, graphs, n, reps, adaptive)
        with c:
            out[graphs] = _run(c, [hot], nsteps)
    (fg, sg, wg, hg), (fe, se, we, he) = out[True], out[False]
    e = l2rel(fg, fe)
    print(f"#   |phi_graph - phi_eager| / |phi_eager| = {e:.2e}; host waits {wg} (captured) vs {we} (eager); "
          f"iterations per solve {sg['iters_mean']:.1f} vs {se['iters_mean']:.1f}; halted steps {hg}", flush=True)
    for st in (sg, se):
        assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
    assert e < 1e-9, e
    # one wait per segment of captured steps, none inside a step (the eager steps poll
    # at least once each); a halted step adds the waits of its eager finish
    if hg == 0:
        assert wg == math.ceil(nsteps / SEG), wg
    assert we >= nsteps, we
    assert abs(sg["iters_mean"] - se["iters_mean"]) <= 0.05 * se["iters_mean"] + 2


@pytest.mark.parametrize("adaptive", [0, 1])
def test_halted_step_is_finished_by_the_host(gpu, monkeypatch, adaptive):
    """The first solve runs on a large constant field (mass M + g phi = 2.55: a few
    iterations), so the captured schedule is sized for it; the field is then replaced by
    a hot one with mass around 0.05, whose solves need many times more.  The first
    captured step halts, the later steps of its segment wait (their chunk kernels keep
    iterating the halted solve on its own field), the host finishes it and resumes with
    a schedule sized from it."""
    Y = gpu
    n, reps, nsteps = 128, 2, 40
    easy = np.full((reps, n, n), 5.0)
    hot = np.random.default_rng(5).uniform(-1.0, 1.0, (reps, n, n))
    out = {}
    for graphs in (True, False):
        c = _ctx(Y, monkeypatch, graphs, n, reps, adaptive)
        with c:
            out[graphs] = _run(c, [easy, hot], nsteps)
    (fg, sg, wg, hg), (fe, se, we, he) = out[True], out[False]
    e = l2rel(fg, fe)
    print(f"# adaptive={adaptive}: |phi_graph - phi_eager| / |phi_eager| = {e:.2e}; halted steps {hg}; "
          f"host waits {wg} (captured) vs {we} (eager)", flush=True)
    assert hg >= 1
    assert he == 0
    for st in (sg, se):
        assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
    assert e < 1e-9, e
    assert wg < we


def test_many_reliable_updates_per_solve(gpu, monkeypatch):
    """With an FP64 reliable update every 8 iterations a solve makes many more of them
    than the seven whose iteration counts the solver state keeps; the captured schedule
    continues past the last recorded one in chunks of the update period up to the
    previous solve's iteration count, so the steps do not halt."""
    Y = gpu
    n, reps, nsteps = 128, 2, 40
    hot = np.random.default_rng(9).uniform(-1.0, 1.0, (reps, n, n))
    out = {}
    for graphs in (True, False):
        c = _ctx(Y, monkeypatch, graphs, n, reps, 0, every=8)
        with c:
            out[graphs] = _run(c, [hot], nsteps)
    (fg, sg, wg, hg), (fe, se, we, he) = out[True], out[False]
    e = l2rel(fg, fe)
    print(f"# reliable update every 8: |phi_graph - phi_eager| / |phi_eager| = {e:.2e}; halted steps {hg}; "
          f"host waits {wg} (captured) vs {we} (eager); iterations per solve {sg['iters_mean']:.1f}", flush=True)
    assert sg["iters_mean"] > 8 * 8   # more updates than the state records
    for st in (sg, se):
        assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
    assert e < 1e-9, e
    assert hg <= 2, hg
    assert wg < we
