"""The C++ host Langevin loop (main.cu's control flow) on the CPU, driven by a
deterministic backend: TOML mapping (main.cu:113-170), burn-in, epsBar calibration
(main.cu:325-352), measurement cadence, output formats (data.csv, traces.csv,
slice.dat, stdout) and the SIGUSR2 early finish (main.cu:65-71,425-428).  The expected
behaviour is restated here independently in Python."""
import math
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "yukawacolourednoise_amd", "host")

INPUT = """[physics]
useMass = "{useMass}"
mass = -0.2
g = 0.5
kappa = 0.18
lambda = 0.02
cutFraction = 0.5

[langevin]
averageEpsilon = 0.02
MaxLangevinTime = {tmax}
ExportTime = 0.5
burnCount = 3
MeasureDriftCount = 2
adaptive = {adaptive}

[io]
configFileName = "test.hdf"
export = "false"
timeSliceFileName = "slice.dat"

[random]
seed = 1234

[fermions]
yukawa_coupling =  1.0
fermion_mass = 1.0
driftMode = 0 # 0 for normal, 1 for rescaled (s^2 times drift)
WilsonParam = 1.0

[lattice]
Nt = 6
Nx = 8
"""


def build_driver(d):
    """Compile the driver loop with the deterministic test backend into directory d."""
    exe = str(d / "driver_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", HOST, "-o", exe,
                    os.path.join(ROOT, "tests", "host", "driver_check.cpp"),
                    os.path.join(HOST, "langevin_driver.cpp"), os.path.join(HOST, "toml_lite.cpp")],
                   check=True)
    return exe


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    return build_driver(tmp_path_factory.mktemp("drv"))


def run(driver, tmp_path, usr2=None, **kw):
    kw.setdefault("useMass", "false")
    kw.setdefault("tmax", 3.0)
    kw.setdefault("adaptive", "true")
    inp = tmp_path / "input.toml"
    inp.write_text(INPUT.format(**kw))
    args = [driver, str(inp), str(tmp_path)] + ([str(usr2)] if usr2 is not None else [])
    r = subprocess.run(args, check=True, capture_output=True, text=True)
    return r.stdout.split("\n"), tmp_path


def expected_flow(adaptive=True, tmax=3.0, usr2=None):
    """Independent restatement of main.cu:302-487 for the fake backend."""
    eps_bar = 0.02
    c = 0

    def step():
        nonlocal c
        f = (0.75 + 0.5 * math.fmod(0.6180339887498949 * c, 1.0)) if adaptive else 1.0
        e = eps_bar * f
        c += 1
        return e

    for _ in range(3):
        t = 0.0
        while t < 0.5:
            t += step()
    elapsed, s, n = 0.0, 0.0, 0
    while elapsed < 2 * 0.5:
        t = 0.0
        while t < 0.5:
            e = step()
            t += e
        elapsed += t
        s += e
        n += 1
    s /= n
    eps_bar *= eps_bar / s
    meas = []
    elapsed = 0.0
    while elapsed < tmax:
        t = 0.0
        early = False
        while t < 0.5:
            if usr2 is not None and c == usr2:
                early = True
            e = step()
            t += e
        elapsed += t
        meas.append(c)
        if early:
            break
    return eps_bar, s, meas


@pytest.mark.parametrize("adaptive", ["true", "false"])
def test_control_flow_and_files(driver, tmp_path, adaptive):
    out, d = run(driver, tmp_path, adaptive=adaptive)
    eps_bar, avg_eps, meas = expected_flow(adaptive == "true")
    assert int(out[0]) == len(meas)
    log = (d / "steps.log").read_text().split("\n")
    measured = [int(l.split()[1]) for l in log if l.startswith("measure")]
    assert measured == meas
    eb = [float(l.split()[1]) for l in log if l.startswith("epsbar")]
    assert len(eb) == 1 and abs(eb[0] - eps_bar) < 1e-12 * eps_bar
    stdout = (d / "stdout.txt").read_text()
    assert "Thermalization done!" in stdout and "Drift measuring done!" in stdout
    assert f"#Average eps during drift measurement = {avg_eps:.6g}" in stdout
    assert "#useMass = \"false\"" in stdout            # input echoed with '#'
    # data.csv: header + Nt rows per measurement
    data = (d / "data.csv").read_text().split("\n")
    assert data[0] == "corr" and len([x for x in data[1:] if x]) == 6 * len(meas)
    # traces.csv: tr,sigma,phi; tr = trace/V, sigma = (sum phi / sqrt(2 kappa))/V
    tr = (d / "traces.csv").read_text().split("\n")
    assert tr[0] == "tr,sigma,phi"
    rows = [list(map(float, x.split(","))) for x in tr[1:] if x]
    assert len(rows) == len(meas)
    sq2k = math.sqrt(2 * 0.18)
    for row, cnt in zip(rows, meas):
        mag = (0.5 * cnt - 40.0) / sq2k
        assert abs(row[0] - 2.0 * cnt / 48) < 1e-4 * max(1, abs(row[0]))
        assert abs(row[1] - mag / 48) < 1e-5 * max(1, abs(row[1]))
        assert abs(row[2] - abs(mag) / 48) < 1e-5 * max(1, abs(row[2]))
    # slice.dat: Nt tab-separated slice means, newline, blank line
    sl = (d / "slice.dat").read_text()
    blocks = [b for b in sl.split("\n\n") if b.strip()]
    assert len(blocks) == len(meas)
    first = [float(v) for v in blocks[0].strip("\n").split("\t") if v]
    assert first == pytest.approx([(t + 0.25 * meas[0]) / 8 for t in range(6)], rel=1e-5)
    assert "#Number of measurements: %d" % len(meas) in stdout
    # stats.csv: time, eps, <phi^2> = phi2 / V, and the backend's CG statistics
    st = (d / "stats.csv").read_text().split("\n")
    assert st[0] == "time,eps,phi2,cg_solves,cg_iters_mean,cg_true_rel_max,cg_not_converged"
    rows = [list(map(float, x.split(","))) for x in st[1:] if x]
    assert len(rows) == len(meas)
    times = [r[0] for r in rows]
    assert all(b > a for a, b in zip(times, times[1:]))
    for row, cnt in zip(rows, meas):
        assert row[2] == pytest.approx(3.0 * cnt / 48, rel=1e-5)
        assert row[3:] == pytest.approx([2, 100.0 + cnt, 1e-11 * cnt, cnt % 2], rel=1e-5)


def test_sigusr2_early_finish(driver, tmp_path):
    _, _, full = expected_flow(True, tmax=3.0)
    usr2 = full[1] + 3          # inside the third export interval
    out, d = run(driver, tmp_path, usr2=usr2)
    _, _, meas = expected_flow(True, tmax=3.0, usr2=usr2)
    assert int(out[0]) == len(meas) == 3
    stdout = (d / "stdout.txt").read_text()
    assert "#Early termination signal received.\n#Wrapping up." in stdout
    assert len([x for x in (d / "traces.csv").read_text().split("\n")[1:] if x]) == 3


def test_parameter_mapping(driver, tmp_path):
    # useMass = "false": m2 = (1 - 2 Lambda)/kappa - 2 nDim, lambda = 6 Lambda/kappa^2,
    # g_device = yukawa / sqrt(2 kappa)  (main.cu:123-129,165-169)
    out, _ = run(driver, tmp_path)
    cfg = out[1].split()
    k, L = 0.18, 0.02
    assert float(cfg[1]) == pytest.approx((1 - 2 * L) / k - 4, rel=1e-5)
    assert float(cfg[2]) == pytest.approx(6 * L / k ** 2, rel=1e-5)
    assert float(cfg[4]) == pytest.approx(1.0 / math.sqrt(2 * k), rel=1e-5)
    assert cfg[5:9] == ["6", "8", "1e-08", "0"]
    assert cfg[9:] == ["0", "2000", "16", "0.001", "0"]     # no preconditioner; training defaults; no even-odd
    # useMass = "true": m2 = mass, lambda = g, kappa = FindKappa (main.cu:58-63,117-122)
    out, _ = run(driver, tmp_path, useMass="true")
    cfg = out[1].split()
    m2, lam = -0.2, 0.5
    kap = (-(4 + m2) + math.sqrt((4 + m2) ** 2 + 4 * lam / 3)) / (2 * lam / 3)
    assert float(cfg[1]) == pytest.approx(m2) and float(cfg[2]) == pytest.approx(lam)
    assert float(cfg[3]) == pytest.approx(kap, rel=1e-5)
    assert float(cfg[4]) == pytest.approx(1.0)


def test_cnn_keys(driver, tmp_path):
    """[cg] precond = "cnn" with its online-training keys."""
    text = INPUT.format(useMass="false", tmax=1.0, adaptive="true")
    text += '\n[cg]\nprecond = "cnn"\ncnnTrainPairs = 64\ncnnPairsPerSolve = 8\ncnnLearningRate = 0.002\n'
    assert 'precond = "cnn"' in text
    inp = tmp_path / "input.toml"
    inp.write_text(text)
    r = subprocess.run([driver, str(inp), str(tmp_path)], check=True, capture_output=True, text=True)
    cfg = r.stdout.split("\n")[1].split()
    assert cfg[9:] == ["1", "64", "8", "0.002", "0"]
    bad = tmp_path / "bad.toml"
    bad.write_text(text.replace("cnnPairsPerSolve = 8", "cnnPairsPerSolve = 0"))
    r = subprocess.run([driver, str(bad), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode != 0


def test_even_odd_key(driver, tmp_path):
    text = INPUT.format(useMass="false", tmax=1.0, adaptive="true") + '\n[cg]\nevenOdd = true\nprecision = "mixed"\n'
    inp = tmp_path / "input.toml"
    inp.write_text(text)
    r = subprocess.run([driver, str(inp), str(tmp_path)], check=True, capture_output=True, text=True)
    assert r.stdout.split("\n")[1].split()[13] == "1"


def test_reference_input_parses(driver, tmp_path):
    """The reference's own input.toml layout (no [lattice] table) maps to the defaults."""
    text = INPUT.format(useMass="false", tmax=1.0, adaptive="true").split("[lattice]")[0]
    text = text.replace("adaptive = true\n", "")
    inp = tmp_path / "input.toml"
    inp.write_text(text)
    r = subprocess.run([driver, str(inp), str(tmp_path)], check=True, capture_output=True, text=True)
    cfg = r.stdout.split("\n")[1].split()
    assert cfg[5:7] == ["16", "16"]          # params.h:29 default lattice
