"""HIP path (libycn.so through the C-ABI) against the CPU FP64 oracle on identical inputs.

Tolerance: 1e-12 relative (max-abs error / max-abs value) for every field, the bar the
hot path is held to; CG iteration counts must agree to within one iteration (different
but fixed reduction orders can move the stopping iteration across the threshold).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def mk(Y, **kw):
    base = dict(Nt=16, Nx=16, m2=-0.2, lam=0.5, g=0.3, M=0.3, r=1.0, cutFraction=0.5, driftMode=0,
                epsBar=0.01, adaptive=0, seed=1234, cgTol=1e-10, cgTolMode=1, cgMaxIter=10000)
    base.update(kw)
    ycn = {k: v for k, v in base.items()}
    p = Y.default_params(**ycn)
    fields = [f for f, _ in O.OrcParams._fields_]
    orc = O.params(**{k: v for k, v in base.items() if k in fields})
    return p, orc


SIZES = [(16, 16), (32, 32), (6, 32), (2, 32), (10, 12), (22, 32), (64, 48)]


@pytest.mark.parametrize("Nt,Nx", SIZES)
@pytest.mark.parametrize("s", [1.0, 0.5, 0.25])
def test_coloured_filter(gpu, Nt, Nx, s):
    Y = gpu
    p, _ = mk(Y, Nt=Nt, Nx=Nx, cutFraction=s)
    xi = np.random.default_rng(Nt * 100 + Nx).standard_normal((Nt, Nx))
    with Y.Context(p) as c:
        out = c.coloured_filter(xi)
    ref = O.coloured_filter(xi, s)
    assert rel(out, ref) < RTOL
    if s == 1.0:
        assert rel(out, xi) < 1e-13  # white-noise identity: every mode is kept


@pytest.mark.parametrize("Nt,Nx", [(32, 32), (6, 32), (64, 64)])
def test_coloured_noise_philox(gpu, Nt, Nx):
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=Nx, seed=1234)
    with Y.Context(p) as c:
        for step in (0, 7, 123456789012):
            out = c.coloured_noise(step)
            assert rel(out, O.coloured_noise(orc, step)) < RTOL


@pytest.mark.parametrize("Nt,Nx", [(16, 16), (6, 10), (2, 32), (64, 64)])
@pytest.mark.parametrize("dagger", [False, True])
def test_dirac(gpu, Nt, Nx, dagger):
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.05, g=0.5)
    rng = np.random.default_rng(5)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    psi = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    with Y.Context(p) as c:
        out = c.dirac_apply(psi, phi=phi, dagger=dagger)
    assert rel(out, O.dirac(orc, phi, psi, dagger)) < RTOL


@pytest.mark.parametrize("mtype", [0, 1])
@pytest.mark.parametrize("tolmode", [0, 1])
def test_cg(gpu, mtype, tolmode):
    Y = gpu
    Nt, Nx = 16, 24
    tol = 1e-8 if tolmode == 0 else 1e-10
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.1, g=0.5, cgTol=tol, cgTolMode=tolmode, cgMaxIter=1000)
    rng = np.random.default_rng(11)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    b = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    with Y.Context(p) as c:
        x, st = c.cg_solve(b, phi=phi, mtype=mtype)
    xr, sr = O.cg(orc, phi, b, mtype)
    assert abs(st["iters"] - sr["iters"]) <= 1, (st, sr)
    assert st["converged"] == 1
    assert rel(x, xr) < 1e-9  # solution agrees to the solver tolerance
    if tolmode == 1:
        assert st["true_rel_resid"] <= 1e-10


def test_cg_fixed_iterations(gpu):
    """Same recurrence: with the iteration count fixed both sides give the same iterate."""
    Y = gpu
    Nt, Nx = 16, 16
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.1, g=0.5, cgTol=0.0, cgTolMode=0, cgMaxIter=37)
    rng = np.random.default_rng(3)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    b = rng.standard_normal((Nt * Nx, 4)) + 0j
    with Y.Context(p) as c:
        x, st = c.cg_solve(b, phi=phi, mtype=1)
    xr, sr = O.cg(orc, phi, b, 1)
    assert st["iters"] == sr["iters"] == 37
    assert rel(x, xr) < RTOL


@pytest.mark.parametrize("driftMode", [0, 1])
def test_fermion_force_fixed_iterations(gpu, driftMode):
    """Same eta, same recurrence run for a fixed number of CG iterations: 1e-12."""
    Y = gpu
    Nt, Nx = 16, 16
    p, orc = mk(Y, Nt=Nt, Nx=Nx, driftMode=driftMode, seed=77, cgTol=0.0, cgTolMode=0, cgMaxIter=60)
    phi = O.hot_start(7, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        d, st = c.fermion_force(5)
    dr, sr = O.fermion_drift(orc, phi, 5)
    assert st["iters"] == sr["iters"] == 60
    assert rel(d, dr) < RTOL


def test_fermion_force_converged(gpu):
    """Both solves stop at true residual <= 1e-10; the drift is then fixed only to the
    solver's accuracy (kappa * tol), so the bar here is 1e-9 relative."""
    Y = gpu
    Nt, Nx = 16, 16
    p, orc = mk(Y, Nt=Nt, Nx=Nx, seed=77)
    phi = O.hot_start(7, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        d, st = c.fermion_force(5)
    dr, sr = O.fermion_drift(orc, phi, 5)
    assert abs(st["iters"] - sr["iters"]) <= 1
    assert st["true_rel_resid"] <= 1e-10
    assert rel(d, dr) < 1e-9


def test_hot_start(gpu):
    Y = gpu
    p, _ = mk(Y, Nt=32, Nx=64)
    with Y.Context(p) as c:
        c.hot_start(7)
        phi = c.get_field()
    assert np.array_equal(phi, O.hot_start(7, 32, 64))


@pytest.mark.parametrize("adaptive", [0, 1])
@pytest.mark.parametrize("Nt,Nx", [(16, 16), (32, 32), (6, 32)])
def test_langevin_step(gpu, adaptive, Nt, Nx):
    """Three steps with a fixed CG iteration count (same recurrence on both sides):
    fields to 1e-12, adaptive eps = epsBar*30/max|drift| to 1e-13."""
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=Nx, adaptive=adaptive, epsBar=0.01, cgTol=0.0, cgTolMode=0, cgMaxIter=40)
    phi0 = O.hot_start(7, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi0)
        phi = phi0.copy()
        for step in range(3):
            r = c.langevin_step(step)
            phi, eps, st = O.langevin_step(orc, phi, step)
            g = c.get_field()
            assert rel(g, phi) < RTOL, step
            assert abs(r.eps[0] - eps) <= 1e-13 * eps, (step, r.eps[0], eps)


@pytest.mark.parametrize("Nt", [2, 10, 22])
def test_sweep_configuration(gpu, Nt):
    """The lattice-size sweep of runs.py:28-36: Nt = 2 .. 22 at Nx = 32, Wilson parameter
    0, fermion mass 0.01, Yukawa 0.07, rescaled drift, s = 0.5, m^2 = -0.5 s^2, lambda =
    1.582 s^2 (useMass = true), adaptive step: three fixed-iteration steps vs the oracle."""
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=32, m2=-0.125, lam=0.3955, g=0.07, M=0.01, r=0.0, cutFraction=0.5, driftMode=1,
                epsBar=0.005, adaptive=1, cgTol=0.0, cgTolMode=0, cgMaxIter=30)
    phi0 = O.hot_start(5, Nt, 32)
    with Y.Context(p) as c:
        c.set_field(phi0)
        phi = phi0.copy()
        for step in range(3):
            r = c.langevin_step(step)
            phi, eps, _ = O.langevin_step(orc, phi, step)
            assert rel(c.get_field(), phi) < RTOL, step
            assert abs(r.eps[0] - eps) <= 1e-12 * eps, step


@pytest.mark.parametrize("Nt,Nx", [(3, 32), (7, 32), (6, 32), (22, 32), (10, 12), (13, 18)])
def test_mixed_precision_on_untiled_lattices(gpu, Nt, Nx):
    """Mixed-precision CG on lattices no fused FP32 tile divides (odd Nt of the runs.py
    sweep included) runs the FP32 two-stencil pair: the converged drift matches the
    oracle's FP64 drift to the solver's accuracy, every solve reaches the FP64 true
    residual 1e-10, and 20 Langevin steps stay with an FP64 context's to 1e-9."""
    Y = gpu
    p, orc = mk(Y, Nt=Nt, Nx=Nx, epsBar=0.01)
    phi0 = O.hot_start(11, Nt, Nx)
    ref, _ = O.fermion_drift(orc, phi0, 4)
    fields = {}
    for prec in (Y.YCN_CG_MIXED, Y.YCN_CG_FP64):
        p.cgPrecision = prec
        with Y.Context(p) as c:
            c.set_field(phi0)
            if prec == Y.YCN_CG_MIXED:
                drift, st = c.fermion_force(4)
                assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
                assert rel(drift, ref) < 1e-8
            st = c.langevin_run(0, 20)
            assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
            fields[prec] = c.get_field()
    assert rel(fields[Y.YCN_CG_MIXED], fields[Y.YCN_CG_FP64]) < 1e-9


@pytest.mark.parametrize("adaptive", [0, 1])
def test_langevin_step_converged(gpu, adaptive):
    """Solves to true residual 1e-10: the drift is fixed to the solver accuracy, so the
    adaptive eps (through max|drift|) and the field agree to 1e-10."""
    Y = gpu
    Nt, Nx = 32, 32
    p, orc = mk(Y, Nt=Nt, Nx=Nx, adaptive=adaptive, epsBar=0.01)
    phi0 = O.hot_start(7, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi0)
        r = c.langevin_step(0)
        g = c.get_field()
    ref, eps, st = O.langevin_step(orc, phi0, 0)
    assert r.stats["true_rel_resid"] <= 1e-10
    tol = 1e-12 if not adaptive else 1e-10
    assert rel(g, ref) < tol
    assert abs(r.eps[0] - eps) <= 1e-10 * eps


def test_batched_replicas(gpu):
    Y = gpu
    Nt, Nx, R = 16, 16, 4
    p, orc = mk(Y, Nt=Nt, Nx=Nx, nReplicas=R, adaptive=0)
    m2 = np.linspace(-1.0, 0.5, R)
    seeds = np.arange(R, dtype=np.uint64)
    phi0 = np.stack([O.hot_start(10 + k, Nt, Nx) for k in range(R)])
    with Y.Context(p) as c:
        c.set_replicas(m2, seeds)
        c.set_field(phi0)
        c.langevin_step(0)
        c.langevin_step(1)
        g = c.get_field()
    for k in range(R):
        o = O.params(Nt=Nt, Nx=Nx, m2=m2[k], lam=0.5, g=0.3, M=0.3, cutFraction=0.5, epsBar=0.01,
                     adaptive=0, seed=int(seeds[k]), cgTol=1e-10, cgTolMode=1, cgMaxIter=10000)
        ph, _, _ = O.langevin_step(o, phi0[k], 0)
        ph, _, _ = O.langevin_step(o, ph, 1)
        assert rel(g[k], ph) < RTOL, k


def test_multi_rhs_noise(gpu):
    Y = gpu
    Nt, Nx = 16, 16
    p, orc = mk(Y, Nt=Nt, Nx=Nx, nNoise=3, cgTol=0.0, cgTolMode=0, cgMaxIter=50)
    phi = O.hot_start(3, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        d, st = c.fermion_force(9)
    dr, _ = O.fermion_drift(orc, phi, 9)
    assert st["systems"] == 3
    assert rel(d, dr) < RTOL


@pytest.mark.parametrize("method", ["plain", "eo-fp64", "eo-mixed"])
def test_measure(gpu, method):
    """Time slices, magnetisation, phi^2 and the point-source correlator (main.cu:441-469)
    against the oracle.  With even-odd preconditioning the propagator D^-1 src comes from
    the fermion-force solve of D^dag chi = gamma_5 src (gamma_5 D gamma_5 = D^dag)."""
    Y = gpu
    Nt, Nx = 16, 32
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.5, g=0.3)
    if method != "plain":
        p.evenOdd = 1
    if method == "eo-mixed":
        p.cgPrecision = Y.YCN_CG_MIXED
    phi = O.hot_start(4, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi)
        ob, sl, cr = c.measure(0, 3)
    assert rel(sl, O.time_slices(phi)) < RTOL
    assert abs(ob["magnetisation"] - phi.sum()) < 1e-12 * np.abs(phi).sum()
    assert abs(ob["phi2"] - (phi ** 2).sum()) < 1e-12 * (phi ** 2).sum()
    cref, _ = O.correlator(orc, phi)
    assert rel(cr, cref) < 1e-9


@pytest.mark.parametrize("evenOdd", [0, 1])
def test_measure_batched_replica(gpu, evenOdd):
    """The correlator of replica 1 of a batched context with 2 noise vectors per drift
    (the measurement solves one system on that replica's field) against the oracle on
    that field alone."""
    Y = gpu
    Nt, Nx = 16, 32
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.5, g=0.3, nReplicas=2, nNoise=2)
    p.evenOdd = evenOdd
    p.cgPrecision = Y.YCN_CG_MIXED
    phis = np.stack([O.hot_start(4, Nt, Nx), O.hot_start(9, Nt, Nx)])
    with Y.Context(p) as c:
        c.set_field(phis)
        ob, sl, cr = c.measure(1, 3)
        ob0, _, cr0 = c.measure(0, 3)
    assert rel(sl, O.time_slices(phis[1])) < RTOL
    orc1 = O.params(Nt=Nt, Nx=Nx, M=0.5, g=0.3)
    assert rel(cr, O.correlator(orc1, phis[1])[0]) < 1e-9
    assert rel(cr0, O.correlator(orc1, phis[0])[0]) < 1e-9


@pytest.mark.parametrize("tile", [-1, 0, 1, 2, 3])
@pytest.mark.parametrize("mtype", [0, 1])
def test_cg_every_tile(gpu, tile, mtype, monkeypatch):
    """Each fused normal-operator tile shape (and the two-stencil pair), forced through
    YCN_TILE, gives the oracle's iterate at a fixed iteration count."""
    Y = gpu
    monkeypatch.setenv("YCN_TILE", str(tile))
    Nt, Nx = 32, 64
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.1, g=0.5, cgTol=0.0, cgTolMode=0, cgMaxIter=25)
    rng = np.random.default_rng(7)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    b = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    with Y.Context(p) as c:
        rep = c.tuning_report()
        x, st = c.cg_solve(b, phi=phi, mtype=mtype)
    xr, sr = O.cg(orc, phi, b, mtype)
    assert st["iters"] == sr["iters"] == 25, rep
    assert rel(x, xr) < RTOL, rep


# ---------------------------------------------------------------------------
# mixed precision: FP32 inner CG, FP64 reliable updates, FP64 true-residual stop
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mtype", [0, 1])
def test_mixed_cg(gpu, mtype):
    Y = gpu
    Nt, Nx = 32, 64
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.05, g=0.5, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)
    p.cgPrecision = Y.YCN_CG_MIXED
    p.reliableEvery = 50
    rng = np.random.default_rng(8)
    phi = rng.uniform(-1, 1, (Nt, Nx))
    b = rng.standard_normal((Nt * Nx, 4)) + 1j * rng.standard_normal((Nt * Nx, 4))
    with Y.Context(p) as c:
        x, st = c.cg_solve(b, phi=phi, mtype=mtype)
    xr, sr = O.cg(orc, phi, b, mtype)
    assert st["converged"] == 1 and st["true_rel_resid"] <= 1e-10, st
    # the FP64 true residual of the mixed solution, recomputed by the oracle
    Ax = O.normal_op(orc, phi, x, mtype)
    assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= 1.5e-10
    assert rel(x, xr) < 1e-7   # both solutions are within kappa * tol of the exact one
    # iteration overhead of FP32 inner iterations stays modest
    assert st["iters"] <= 1.5 * sr["iters"] + 100, (st, sr)


def test_mixed_langevin_step(gpu):
    """Field error of one mixed-precision step against the FP64 oracle step."""
    Y = gpu
    Nt, Nx = 32, 32
    p, orc = mk(Y, Nt=Nt, Nx=Nx, M=0.05, g=0.5, m2=-0.4, lam=1.0, epsBar=0.005)
    p.cgPrecision = Y.YCN_CG_MIXED
    phi0 = O.hot_start(7, Nt, Nx)
    with Y.Context(p) as c:
        c.set_field(phi0)
        r = c.langevin_step(0)
        g = c.get_field()
    ref, _, _ = O.langevin_step(orc, phi0, 0)
    assert r.stats["true_rel_resid"] <= 1e-10
    assert rel(g, ref) < 1e-10


# ---------------------------------------------------------------------------
# committed golden fixtures (tests/golden/)
# ---------------------------------------------------------------------------
import json  # noqa: E402
import os  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLD_MAN = json.load(open(os.path.join(GOLD, "manifest.json")))


def _gold(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


@pytest.mark.parametrize("name", sorted(n for n in GOLD_MAN if n.startswith("step_")))
def test_gpu_golden_step(gpu, name):
    Y = gpu
    f = _gold(name)
    keys = ["Nt", "Nx", "m2", "lam", "g", "M", "r", "cutFraction", "driftMode", "epsBar", "adaptive", "seed",
            "cgTol", "cgTolMode", "cgMaxIter"]
    p = Y.default_params(**{k: f[k].item() for k in keys})
    step = int(f["step"])
    with Y.Context(p) as c:
        c.hot_start(int(f["init_seed"]))
        assert np.array_equal(c.get_field(), f["phi0"])
        assert rel(c.coloured_noise(step), f["noise"]) < RTOL
        r = c.langevin_step(step)
        assert rel(c.get_field(), f["phi1"]) < RTOL
        assert abs(r.eps[0] - float(f["eps"])) <= 1e-13 * float(f["eps"])
        assert abs(r.stats["iters"] - int(f["cg_iters"])) <= 1 or float(f["cgTol"]) == 0.0


@pytest.mark.parametrize("name", sorted(n for n in GOLD_MAN if n.startswith("filter_")))
def test_gpu_golden_filter(gpu, name):
    Y = gpu
    f = _gold(name)
    Nt, Nx = f["xi"].shape
    with Y.Context(Y.default_params(Nt=Nt, Nx=Nx, cutFraction=float(f["cutFraction"]))) as c:
        assert rel(c.coloured_filter(f["xi"]), f["out"]) < RTOL


def test_gpu_golden_dirac(gpu):
    Y = gpu
    f = _gold("dirac_8x12.npz")
    with Y.Context(Y.default_params(Nt=8, Nx=12, M=float(f["M"]), g=float(f["g"]), r=float(f["r"]))) as c:
        assert rel(c.dirac_apply(f["psi"], phi=f["phi"]), f["Dpsi"]) < RTOL
        assert rel(c.dirac_apply(f["psi"], phi=f["phi"], dagger=True), f["Ddagpsi"]) < RTOL
