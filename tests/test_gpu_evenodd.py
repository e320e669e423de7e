"""Even-odd preconditioned fermion force on the GPU against the oracle: fixed-iteration
iterates of the Schur-complement CG (same recurrence on both sides) to 1e-12, converged
drifts against the unpreconditioned oracle drift to the solver accuracy, Langevin steps,
mixed precision, batching, every fused tile shape and the lattice-size sweep shapes."""
import numpy as np
import pytest

import oracle as O
from test_gpu_parity import RTOL, mk, rel

pytestmark = pytest.mark.gpu

EO_TILES64 = [(8, 32), (16, 16), (4, 64), (4, 32), (2, 32)]
EO_TILES32 = [(8, 32), (16, 32), (8, 64), (16, 16), (4, 64), (2, 32), (32, 32), (16, 64),
              (8, 32, "persistent"), (16, 32, "persistent"), (16, 16, "persistent"), (8, 64, "persistent")]


@pytest.mark.parametrize("Nt,Nx", [(16, 16), (32, 32), (6, 32), (8, 64)])
def test_eo_force_fixed_iterations(gpu, Nt, Nx):
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=Nx, cgTol=0.0, cgTolMode=0, cgMaxIter=30)
    p.evenOdd = 1
    orc.evenOdd = 1
    phi = O.hot_start(11, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        d, st = c.fermion_force(5)
    dr, sr = O.fermion_drift(orc, phi, 5)
    assert st["iters"] == sr["iters"] == 30
    assert rel(d, dr) < RTOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("Nt,Nx", [(32, 32), (64, 64)])
def test_eo_force_converged(gpu, precision, Nt, Nx):
    """Converged to |eta - D^dag chi| / |eta| <= 1e-10: equals the oracle's
    unpreconditioned drift to the solver accuracy."""
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.05, g=0.5, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    if precision == "mixed":
        p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(3, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        d, st = c.fermion_force(2)
    dr, sr = O.fermion_drift(orc, phi, 2)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
    assert st["iters"] < sr["iters"]
    assert rel(d, dr) < 1e-8


@pytest.mark.parametrize("adaptive", [0, 1])
def test_eo_langevin_steps_fixed_iterations(gpu, adaptive):
    Y = gpu
    Nt, Nx = 16, 32
    p, orc = mk(Y, Nt=Nt, Nx=Nx, adaptive=adaptive, cgTol=0.0, cgTolMode=0, cgMaxIter=25)
    p.evenOdd = 1
    orc.evenOdd = 1
    phi = O.hot_start(4, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        for step in range(3):
            r = c.langevin_step(step)
            phi, eps, _ = O.langevin_step(orc, phi, step)
            assert rel(c.get_field(), phi) < RTOL, step
            assert abs(r.eps[0] - eps) <= 1e-12 * eps


@pytest.mark.parametrize("tile", range(len(EO_TILES64)))
def test_eo_every_fp64_tile(gpu, tile, monkeypatch):
    monkeypatch.setenv("YCN_TILE_EO", str(tile))
    Y = gpu
    Nt, Nx = 16, 64
    p, orc = mk(Y, Nt=Nt, Nx=Nx, cgTol=0.0, cgTolMode=0, cgMaxIter=20)
    p.evenOdd = 1
    orc.evenOdd = 1
    phi = O.hot_start(8, Nt, Nx)
    with Y.Context(p) as c:
        assert "forced by YCN_TILE_EO" in c.tuning_report()
        c.set_field(phi)
        d, _ = c.fermion_force(1)
    dr, _ = O.fermion_drift(orc, phi, 1)
    assert rel(d, dr) < RTOL, EO_TILES64[tile]


FIN_REPORT = {"separate": "separate finalize", "fused": "prologue finalize"}


@pytest.mark.parametrize("fin", ["separate", "fused"])
@pytest.mark.parametrize("tile", range(len(EO_TILES32)))
def test_eo_every_fp32_tile(gpu, tile, fin, monkeypatch):
    """Every FP32 tile, with the finalize as its own kernel and in the next launch's
    prologue (every block sums the previous launch's partials)."""
    monkeypatch.setenv("YCN_TILE_EO32", str(tile))
    monkeypatch.setenv("YCN_EO_FIN", fin)
    Y = gpu
    Nt, Nx = 32, 128
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.1, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(9, Nt, Nx)
    t = EO_TILES32[tile]
    name = f"{t[0]}x{t[1]}" + ("p" if "persistent" in t else "")
    with Y.Context(p) as c:
        assert f"fp32 pipelined iteration: {name}, {FIN_REPORT[fin]}" in c.tuning_report()
        c.set_field(phi)
        d, st = c.fermion_force(1)
    dr, _ = O.fermion_drift(orc, phi, 1)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10
    assert rel(d, dr) < 1e-8, EO_TILES32[tile]


@pytest.mark.parametrize("fin", ["separate", "fused"])
def test_eo_finalize_many_blocks(gpu, fin, monkeypatch):
    """104 x 96 with 2 x 32 tiles: 156 block partials per system, more than the 64
    threads that sum them, for two noise systems at once; both finalize forms reach the
    oracle's drift."""
    monkeypatch.setenv("YCN_TILE_EO32", "5")
    monkeypatch.setenv("YCN_EO_FIN", fin)
    Y = gpu
    Nt, Nx = 104, 96
    p, orc = mk(Y, Nt=Nt, Nx=Nx, nNoise=2, M=0.1, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(19, Nt, Nx)
    with Y.Context(p) as c:
        assert "2x32, " + FIN_REPORT[fin] in c.tuning_report()
        c.set_field(phi)
        d, st = c.fermion_force(3)
    dr, _ = O.fermion_drift(orc, phi, 3)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10 and st["systems"] == 2
    assert rel(d, dr) < 1e-8


def test_eo_batched_replicas_and_noise(gpu):
    Y = gpu
    Nt, Nx, R, J = 16, 32, 3, 2
    p, _ = mk(Y, Nt=Nt, Nx=Nx, nReplicas=R, nNoise=J, cgTol=1e-11, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    m2 = np.linspace(-0.5, 0.2, R)
    seeds = np.arange(20, 20 + R, dtype=np.uint64)
    phi = np.stack([O.hot_start(30 + k, Nt, Nx) for k in range(R)])
    with Y.Context(p) as c:
        c.set_replicas(m2, seeds)
        c.set_field(phi)
        d, st = c.fermion_force(6)
    assert st["converged"] == 1
    for k in range(R):
        o = O.params(Nt=Nt, Nx=Nx, m2=m2[k], lam=0.5, g=0.3, M=0.3, cutFraction=0.5, seed=int(seeds[k]),
                     cgTol=1e-11, cgTolMode=1, cgMaxIter=20000, nNoise=J)
        dr, _ = O.fermion_drift(o, phi[k], 6)
        assert rel(d[k], dr) < 1e-8, k


@pytest.mark.parametrize("Nt", [2, 10, 22])
def test_eo_small_time_extents(gpu, Nt):
    """Time extents of the lattice-size sweep of runs.py:28-36 (Nt = 2 .. 22 at Nx = 32,
    rescaled drift; Wilson r = 1, where the even-odd diagonal stays away from zero):
    fixed-iteration even-odd steps against the oracle."""
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=32, m2=-0.125, lam=0.3955, g=0.07, M=0.01, r=1.0, driftMode=1, epsBar=0.005,
                adaptive=1, cgTol=0.0, cgTolMode=0, cgMaxIter=20)
    p.evenOdd = 1
    orc.evenOdd = 1
    phi = O.hot_start(12, Nt, 32)
    with Y.Context(p) as c:
        c.set_field(phi)
        for step in range(2):
            c.langevin_step(step)
            phi, _, _ = O.langevin_step(orc, phi, step)
            assert rel(c.get_field(), phi) < RTOL, step


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_eo_second_solve_starts_from_learned_threshold(gpu, precision):
    """Relative mode: a context's next solve starts from the recursive threshold its
    previous solve's true-residual check asked for.  Solving the same system twice, the
    second solve needs no more true-residual checks than the first, both meet the 1e-10
    bar, and the second drift still equals the oracle's to the solver accuracy."""
    Y = gpu
    Nt, Nx = 64, 64
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.05, g=0.5, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    if precision == "mixed":
        p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(3, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        d1, s1 = c.fermion_force(2)
        d2, s2 = c.fermion_force(2)   # same eta and field: the same system
    dr, _ = O.fermion_drift(orc, phi, 2)
    print(f"# {precision}: first {s1['iters']} iterations / {s1['restarts']} checks, "
          f"second {s2['iters']} / {s2['restarts']}", flush=True)
    for s in (s1, s2):
        assert s["converged"] == 1 and s["true_rel_resid"] <= 1e-10, s
    assert s2["restarts"] <= s1["restarts"]
    assert rel(d1, dr) < 1e-8 and rel(d2, dr) < 1e-8


@pytest.mark.parametrize("r", [0.7, 1.0])
def test_eo_mixed_wilson_parameter(gpu, r):
    """The pipelined FP32 kernels have a specialised hop for Wilson r = 1 (projector
    form) and the general one otherwise; both reach the oracle's drift."""
    Y = gpu
    Nt, Nx = 32, 64
    p, orc = mk(Y, Nt=Nt, Nx=Nx, r=r, M=0.1, g=0.5, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(21, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        d, st = c.fermion_force(4)
    dr, _ = O.fermion_drift(orc, phi, 4)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
    assert rel(d, dr) < 1e-8


@pytest.mark.parametrize("Nt", [2, 6, 10])
def test_eo_mixed_small_time_extents(gpu, Nt):
    """The fastest path on the sweep's small time extents (Nt = 2 .. 10 at Nx = 32):
    converged pipelined mixed-precision drift against the oracle's FP64 drift."""
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=32, M=0.1, g=0.3, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(13, Nt, 32)
    with Y.Context(p) as c:
        c.set_field(phi)
        d, st = c.fermion_force(2)
    dr, _ = O.fermion_drift(orc, phi, 2)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
    assert rel(d, dr) < 1e-8


@pytest.mark.parametrize("fin", ["fused", "separate"])
def test_profile_counts_iterations(gpu, fin, monkeypatch):
    """The library's kernel profiling (ycn_profile_begin / ycn_profile_end, the source of
    bench.py's roofline figure).  With the finalize in the next launch's prologue a CG chunk
    is bracketed as a whole and divided by the iterations done in it, so the counted
    launches are exactly the profiled steps' iterations; with the separate finalize every
    iteration launch is bracketed, and the launches that exit at once after a pause are
    left out (at least the iterations are counted).  The bytes per launch are those of one
    pipelined iteration and the profiled time stays within the wall time."""
    import time
    monkeypatch.setenv("YCN_EO_FIN", fin)
    Y = gpu
    N = 64
    p, _ = mk(Y, Nt=N, Nx=N, M=0.05, g=0.5, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.evenOdd = 1
    p.cgPrecision = Y.YCN_CG_MIXED
    with Y.Context(p) as c:
        rep = c.tuning_report()
        assert ("prologue finalize" if fin == "fused" else "separate finalize") in rep.split("|")[-1], rep
        c.hot_start(3)
        c.langevin_run(0, 2)
        c.sync()
        c.profile_begin("cg_normal_op")
        t0 = time.perf_counter()
        st = c.langevin_run(2, 3)
        c.sync()
        wall = time.perf_counter() - t0
        ms, n, b = c.profile_end()
    print(f"# {fin}: {n} launches counted for {st['iters']} iterations, {1e3 * ms / max(n, 1):.2f} us each", flush=True)
    assert st["converged"] == 1
    if fin == "fused":
        assert n == st["iters"]
    else:
        assert st["iters"] <= n <= st["iters"] + 3 * 64
    # FP32 half-lattice vectors (16 B per site of the full lattice) and phi (4 B).  r is
    # rebuilt from the directions: launches read p_{k-2} (launch 0 of a chunk: r), Ap and p
    # and write p and Ap, and every other launch reads and writes x_s, so a chunk's
    # launches average 6 vectors, with either finalize.
    assert b == N * N * (6 * 16 + 4)
    assert 0.0 < ms < 1e3 * wall


def test_eo_mixed_iteration_cap_with_prediction(gpu):
    """An even-odd mixed solve that cannot reach its tolerance stops at the iteration cap
    and reports it (YCN_ERR_NOT_CONVERGED, converged = 0, iterations = cap), also when
    the previous solve's count lets the host enqueue reliable updates without polling."""
    Y = gpu
    Nt, Nx = 32, 64
    cap = 120
    p, _ = mk(Y, Nt=Nt, Nx=Nx, M=0.05, g=0.5, cgTol=1e-30, cgTolMode=1, cgMaxIter=cap)
    p.evenOdd = 1
    p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(11, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        for k in range(3):
            d, st = c.fermion_force(5 + k)
            print(f"# solve {k}: {st}", flush=True)
            assert st["converged"] == 0 and st["iters"] == cap, st
            assert np.all(np.isfinite(d)) and np.isfinite(st["true_rel_resid"])


@pytest.mark.parametrize("cap", [61, 121])
def test_eo_mixed_odd_cap_keeps_the_deferred_update(gpu, monkeypatch, cap):
    """The even launch of each pair of pipelined iterations leaves its x_s update to the
    odd launch after it.  A solve stopped by an odd iteration cap ends after an even
    launch, and the reliable update that closes it must add that launch's term.  CG on
    the normal equations lowers the system residual |b - S x| in every iteration, so the
    true residual at the odd cap must lie clearly below the one a cap one lower reaches
    (the two runs are the same up to that launch, so a dropped term would give exactly
    that value), and above the one at a cap one higher.  With either finalize form (the
    separate one stops in the finalize kernel, the prologue one in the next launch)."""
    Y = gpu
    Nt, Nx = 32, 64
    phi = O.hot_start(11, Nt, Nx)
    res = {}
    for fin in ("fused", "separate"):
        monkeypatch.setenv("YCN_EO_FIN", fin)
        for n in (cap - 1, cap, cap + 1):
            p, _ = mk(Y, Nt=Nt, Nx=Nx, M=0.05, g=0.5, cgTol=1e-30, cgTolMode=1, cgMaxIter=n)
            p.evenOdd = 1
            p.cgPrecision = Y.YCN_CG_MIXED
            with Y.Context(p) as c:
                rep = c.tuning_report()
                assert ("prologue finalize" if fin == "fused" else "separate finalize") in rep.split("|")[-1], rep
                c.set_field(phi)
                _, st = c.fermion_force(5)
            assert st["converged"] == 0 and st["iters"] == n, st
            res[fin, n] = st["true_rel_resid"]
        lo, mid, hi = res[fin, cap - 1], res[fin, cap], res[fin, cap + 1]
        print(f"# cap {cap} ({fin} finalize): true residual {lo:.4e} at {cap - 1}, {mid:.4e} at {cap}, "
              f"{hi:.4e} at {cap + 1}", flush=True)
        assert hi < mid < 0.97 * lo, (fin, lo, mid, hi)
    assert abs(res["fused", cap] - res["separate", cap]) < 0.25 * (res["separate", cap - 1] - res["separate", cap])
