"""The committed golden fixtures (tests/golden/, written by make_golden.py from the
oracle) are intact and the oracle still reproduces them bit-for-bit."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(G, "manifest.json")))


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_manifest(name):
    assert hashlib.sha256(open(os.path.join(G, name), "rb").read()).hexdigest() == MANIFEST[name]


def load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


def orc_params(f):
    keys = [k for k, _ in O.OrcParams._fields_]
    return O.params(**{k: (f[k].item() if hasattr(f[k], "item") else f[k]) for k in keys if k in f})


@pytest.mark.parametrize("name", [n for n in MANIFEST if n.startswith("step_")])
def test_oracle_reproduces_step(name):
    f = load(name)
    p = orc_params(f)
    step = int(f["step"])
    assert np.array_equal(O.hot_start(int(f["init_seed"]), p.Nt, p.Nx), f["phi0"])
    assert np.array_equal(O.coloured_noise(p, step), f["noise"])
    assert np.array_equal(O.fermion_noise(p.seed, 0, step, p.Nt * p.Nx), f["eta"])
    phi1, eps, st = O.langevin_step(p, f["phi0"], step)
    assert np.array_equal(phi1, f["phi1"]) and eps == float(f["eps"])
    assert st["iters"] == int(f["cg_iters"])


@pytest.mark.parametrize("name", [n for n in MANIFEST if n.startswith("filter_")])
def test_oracle_reproduces_filter(name):
    f = load(name)
    assert np.array_equal(O.coloured_filter(f["xi"], float(f["cutFraction"])), f["out"])


def test_oracle_reproduces_dirac():
    f = load("dirac_8x12.npz")
    p = O.params(Nt=8, Nx=12, M=float(f["M"]), g=float(f["g"]), r=float(f["r"]))
    assert np.array_equal(O.dirac(p, f["phi"], f["psi"], False), f["Dpsi"])
    assert np.array_equal(O.dirac(p, f["phi"], f["psi"], True), f["Ddagpsi"])
