#!/usr/bin/env python3
"""Langevin steps/s of the 2D lattice Yukawa model with coloured noise on one MI355X.

Workload (the headline configuration the metric is quoted on): Nt = Nx = 1024,
M = 0.05, g = 0.5, m^2 = -0.4, lambda = 1.0, r = 1, cutoff s = 0.5, eps = 0.005 fixed,
one Gaussian noise vector per drift estimate, every CG solve run to true relative
residual <= 1e-10 (recomputed in FP64, restarted while above).  Start: hot start
phi ~ U[-1, 1] (Philox, seed 7) and `therm` untimed thermalisation steps, done once
and shared: every CG method (plain FP64, mixed precision, learned preconditioner)
starts from that same field at the same step counter, runs `warmup` untimed steps and
then the timed region of `steps` full Langevin steps (coloured noise, fermion force,
bosonic drift, update), bracketed by device synchronisation.  With the learned method
the thermalisation steps are also its online training (untimed, reported).  Inputs
and state stay resident in HBM.

Prints one JSON line.  The timed region runs uninstrumented, CG iteration chunks replayed
as captured hipGraphs.  `roofline` is measured live right after it, in a profiled pass
of `prof_steps` steps with HIP events on the library's own stream.  For the pipelined
even-odd iteration with the prologue finalize, a CG chunk is nothing but back-to-back
launches of that kernel: one event pair brackets each chunk, divided by the iterations
done in it (launches after a pause count as time, not as launches).  Otherwise every
launch of the dominant kernel is bracketed (launches that exit at once for converged
systems excluded).
`cpu_baseline` times full FP64 CPU-oracle steps of the same workload from the same start
field on the host cores (about 20 s).  `--cpu-plan` runs the oracle alone on the lattice
sizes of the workload table and exits.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant CG kernel at 1024x1024 (read + write), from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs scaled by what the same counters read
# for a 1 GiB device copy (FETCH_SIZE 0.5 KiB, WRITE_SIZE 1.0 KiB per KiB moved); the
# summaries with their commands are profiles/hbm_traffic_mixed_cg_1024x1024.md (fused
# normal operator) and profiles/hbm_traffic_eo_mixed_cg_1024x1024.md (pipelined
# even-odd iteration: the mean of its two alternating launch forms, with and without the
# x_s update; its r is rebuilt from the directions).  Keyed by (kernel, precision, lattice extent); others report null.
MEASURED_TRAFFIC = {("cg_normal_op", "fp32", 1024): 140720905,
                    ("cg_normal_op-eo", "fp32", 1024): 108734822}

WORKLOAD = dict(Nt=1024, Nx=1024, m2=-0.4, lam=1.0, g=0.5, M=0.05, r=1.0, cutFraction=0.5,
                driftMode=0, epsBar=0.005, adaptive=0, seed=1234, cgTol=1e-10, cgTolMode=1,
                cgMaxIter=20000)


def cpu_baseline(phi0, step0: int, budget_s: float = 20.0):
    """FP64 oracle on the host cores: full Langevin steps of the 1024^2 workload.

    The oracle (noise, plain FP64 CG on D^dag D to the same true residual 1e-10, drift,
    update) steps from the GPU methods' shared start field phi0 at the same step counter,
    so it solves the systems the timed GPU steps solve; steps run until `budget_s` is
    used (at least one).  Nothing is priced or extrapolated."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    cores = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    p = O.params(**WORKLOAD)
    phi = np.ascontiguousarray(phi0, dtype=np.float64)
    n, wall, iters = 0, 0.0, 0
    while n == 0 or wall < budget_s:
        t0 = time.perf_counter()
        phi, _, st = O.langevin_step(p, phi, step0 + n)
        wall += time.perf_counter() - t0
        iters += st["iters"]
        n += 1
    return {
        "value": n / wall,
        "unit": "steps/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle FP64 (plain CG to true residual 1e-10) on {cores} host threads: {n} full {p.Nt}x{p.Nx} "
                   f"Langevin steps from the GPU methods' start field, {wall:.1f} s, {iters / n:.0f} CG "
                   f"iterations per solve"),
        "cpu": cpu_model(),
    }


# The CPU plan of the workload table: the 32^2 case in full, the larger lattices for
# 20 steps each (hot start, seed 7), plain FP64 CG to true residual 1e-10.
CPU_PLAN = {
    "32x32 (m2=-0.2, lambda=0.5, M=0.3, g=0.3, eps=0.01, cold start, 2000 steps)":
        (dict(Nt=32, Nx=32, m2=-0.2, lam=0.5, g=0.3, M=0.3, epsBar=0.01), 2000, None),
    "256x256 (same couplings, eps=0.005, s=0.5, 20 steps)":
        (dict(Nt=256, Nx=256, m2=-0.2, lam=0.5, g=0.3, M=0.3, epsBar=0.005), 20, 7),
    "512x512 (M=0.05, g=0.5, m2=-0.4, lambda=1, eps=0.005, 20 steps)":
        (dict(Nt=512, Nx=512, m2=-0.4, lam=1.0, g=0.5, M=0.05, epsBar=0.005), 20, 7),
    "1024x1024 (the headline couplings, 20 steps)":
        (dict(Nt=1024, Nx=1024, m2=-0.4, lam=1.0, g=0.5, M=0.05, epsBar=0.005), 20, 7),
}


def cpu_plan():
    """Runs CPU_PLAN with the oracle on the host cores; one JSON line."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    cores = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    out = {"cpu_plan": {}, "cores": cores, "cpu": cpu_model()}
    for name, (kw, steps, hot) in CPU_PLAN.items():
        p = O.params(**{**WORKLOAD, **kw, "r": 1.0, "cutFraction": 0.5})
        phi = O.hot_start(hot, p.Nt, p.Nx) if hot is not None else np.zeros((p.Nt, p.Nx))
        iters, worst, t0 = 0, 0.0, time.perf_counter()
        for step in range(steps):
            phi, _, st = O.langevin_step(p, phi, step)
            iters += st["iters"]
            worst = max(worst, st["true_rel_resid"])
        wall = time.perf_counter() - t0
        out["cpu_plan"][name] = {"steps_per_s": steps / wall, "steps": steps, "seconds": wall,
                                 "cg_iters_per_solve": iters / steps, "true_rel_resid_max": worst}
        print(f"# {name}: {steps / wall:.3g} steps/s, {iters / steps:.0f} CG iterations per solve", flush=True)
    print(json.dumps(out))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


THERM_STEPS = 125           # shared thermalisation before any method is timed
TRAIN_PAIRS_PER_SOLVE = 16  # (residual, error) pairs per solve for the learned preconditioner


def make_ctx(Y, size, method):
    """A context of the workload for one CG method.  "eo-mixed-polled" is eo-mixed whose
    Langevin runs launch single steps that poll CG convergence (YCN_STEP_GRAPHS=0) instead
    of replaying captured steps, reported beside it for comparison."""
    if method == "eo-mixed-polled":
        saved = os.environ.get("YCN_STEP_GRAPHS")
        os.environ["YCN_STEP_GRAPHS"] = "0"
        try:
            return make_ctx(Y, size, "eo-mixed")
        finally:
            if saved is None:
                del os.environ["YCN_STEP_GRAPHS"]
            else:
                os.environ["YCN_STEP_GRAPHS"] = saved
    wl = dict(WORKLOAD)
    wl["Nt"] = wl["Nx"] = size
    p = Y.default_params(**wl)
    p.cgPrecision = Y.YCN_CG_FP64 if method in ("fp64", "eo-fp64") else Y.YCN_CG_MIXED
    p.precond = Y.YCN_PRECOND_CNN if method == "learned" else Y.YCN_PRECOND_NONE
    p.evenOdd = 1 if method.startswith("eo-") else 0
    p.reliableEvery = 50
    return Y.Context(p)


def thermalise(Y, size, therm, train, seed=7):
    """Hot start (Philox, `seed`) and `therm` Langevin steps.  With `train` the steps
    are run by the learned method's context, which trains its preconditioner online on
    the fermion solves of those steps (its FP64 CG leaves the (r, e) pairs); otherwise a
    mixed-precision context runs them.  Returns (context or None, phi after the steps,
    training stats)."""
    ctx = make_ctx(Y, size, "learned" if train else "mixed")
    ctx.hot_start(seed)
    t0 = time.perf_counter()
    info = None
    if train:
        info = ctx.precond_train(0, therm, pairs_per_solve=TRAIN_PAIRS_PER_SOLVE, lr=1e-3)
        info["steps"] = therm
    else:
        ctx.langevin_run(0, therm)
    ctx.sync()
    secs = time.perf_counter() - t0
    if info is not None:
        info["seconds"] = secs
    phi = ctx.get_field()
    if not train:
        ctx.close()
        ctx = None
    return ctx, phi, info, secs


def run_method(Y, ctx, size, method, phi0, step0, steps, warmup, profile_kernel, prof_steps):
    """Time `steps` Langevin steps of one CG method.  Every method starts from the same
    thermalised field phi0 at the same step counter step0, so all of them follow the same
    trajectory (same noise; drifts agree to the CG tolerance) and solve the same systems.

    The timed region runs uninstrumented (CG chunks replayed as captured graphs).  A
    profiled pass of `prof_steps` more steps follows it: every launch of
    `profile_kernel` there is bracketed by HIP events on the library's stream (which
    launches kernels one by one instead of replaying graphs); its average duration gives
    the roofline figure."""
    if ctx is None:
        ctx = make_ctx(Y, size, method)
    tuning = ctx.tuning_report()
    ctx.set_field(phi0)
    step = step0
    # warm-up through the same entry point as the timed steps: a first single step (it
    # sizes the solver's chunk schedule), then a run, which captures the step graphs the
    # timed run replays
    for _ in range(min(warmup, 1)):
        ctx.langevin_step(step)
        step += 1
    if warmup > 1:
        ctx.langevin_run(step, warmup - 1)
        step += warmup - 1
    ctx.sync()
    w0, h0 = ctx.host_syncs()
    t0 = time.perf_counter()
    st = ctx.langevin_run(step, steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    w1, h1 = ctx.host_syncs()
    step += steps
    prof = (0.0, 0, 0.0)
    if profile_kernel and prof_steps > 0:
        ctx.profile_begin(profile_kernel)
        ctx.langevin_run(step, prof_steps)
        prof = ctx.profile_end()
    mag = float(ctx.get_field().mean())
    ctx.close()
    return {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
            "cg_iters_per_solve": st["iters"] / max(1, steps), "true_rel_resid_max": st["true_rel_resid"],
            "reliable_updates_or_restarts": st["restarts"], "final_magnetisation": mag,
            "host_waits_per_step": (w1 - w0) / max(1, steps), "captured_steps_finished_by_host": h1 - h0,
            "tuning": tuning, "prof": prof}


def copy_calibration(nbytes: int = 1 << 30):
    """A plain device copy of `nbytes` (torch's copy kernel), bracketed by writes of two
    other buffers of the same size so that neither L2 nor the last-level cache holds the
    source: profiled with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, the copy kernel's
    counters against its known nbytes read + nbytes written give the scale factors that
    turn the counters of the library's kernels into HBM bytes."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    evict = torch.full((n,), 2.0, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    evict.add_(1.0)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    assert float(dst[-1]) == 1.0
    print(json.dumps({"copy_calibration_bytes": nbytes, "kernel": "torch copy_ (float32)"}), flush=True)


C4 = dict(Nt=64, Nx=64, nReplicas=64, nNoise=8, lam=0.5, g=0.3, M=0.3, r=1.0, cutFraction=0.5, driftMode=0,
          epsBar=0.01, adaptive=0, cgTol=1e-10, cgTolMode=1, cgMaxIter=20000)


def bench_c4(Y, steps, warmup, therm, methods=("fp64", "mixed", "eo-fp64", "eo-mixed"), profile_kernel="cg_normal_op",
             prof_steps=2):
    """The batched workload: 64 chains of 64^2 in one context, m^2_k = -1 + 1.5 k / 63,
    seeds 0..63, 8 noise vectors per drift estimate solved together (512 systems per
    CG launch), eps = 0.01 fixed.  Reports chain-steps/s per method from a common
    thermalised start, and the roofline of the fastest method's normal-operator kernel
    from a profiled pass of `prof_steps` steps after its timed region (HIP events on the
    library's stream, as for the headline workload)."""
    m2 = [-1.0 + 1.5 * k / 63 for k in range(64)]
    seeds = list(range(64))
    out = {}
    phi0 = None
    for method in methods:
        p = Y.default_params(**C4)
        p.cgPrecision = Y.YCN_CG_FP64 if method in ("fp64", "eo-fp64") else Y.YCN_CG_MIXED
        p.evenOdd = 1 if method.startswith("eo-") else 0
        p.reliableEvery = 50
        ctx = Y.Context(p)
        ctx.set_replicas(m2, seeds)
        if phi0 is None:
            ctx.hot_start(7)
            ctx.langevin_run(0, therm)
            phi0 = ctx.get_field()
        ctx.set_field(phi0)
        step = therm
        for _ in range(min(warmup, 1)):
            ctx.langevin_step(step)
            step += 1
        if warmup > 1:
            ctx.langevin_run(step, warmup - 1)
            step += warmup - 1
        ctx.sync()
        t0 = time.perf_counter()
        st = ctx.langevin_run(step, steps)
        ctx.sync()
        dt = time.perf_counter() - t0
        tuning = ctx.tuning_report()
        prof = (0.0, 0, 0.0)
        if profile_kernel and prof_steps > 0:
            ctx.profile_begin(profile_kernel)
            ctx.langevin_run(step + steps, prof_steps)
            prof = ctx.profile_end()
        ctx.close()
        out[method] = {"chain_steps_per_s": 64 * steps / dt, "ms_per_batch_step": 1e3 * dt / steps,
                       "cg_iters_per_solve_max": st["iters"] / max(1, steps),
                       "cg_iters_per_solve_mean": st["iters_mean"], "true_rel_resid_max": st["true_rel_resid"],
                       "tuning": tuning, "prof": prof}
        print(f"# c4 {method}: {out[method]['chain_steps_per_s']:.1f} chain-steps/s", file=sys.stderr, flush=True)
    head = max(out, key=lambda m: out[m]["chain_steps_per_s"])
    k_ms, k_n, k_bytes = out[head]["prof"]
    avg_ms = k_ms / max(1, k_n)
    achieved = (k_bytes / 1e9) / (avg_ms / 1e3) if k_n else 0.0
    chunked = head == "eo-mixed" and "prologue finalize" in out[head]["tuning"].split("|")[-1]
    res = {
        "metric": "chain-steps/s, 64 chains x 64^2, 8 noise vectors per drift (true resid <= 1e-10)",
        "value": out[head]["chain_steps_per_s"],
        "unit": "chain-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": out[head]["ms_per_batch_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head in ("fp64", "eo-fp64") else "f32+f64",
        "data": f"synthetic (Philox hot start, seed 7, {therm} thermalisation steps)",
        "config": {"workload": "64 chains of 64x64 (m^2 scan -1 .. 0.5), 512 systems per CG launch, eps 0.01",
                   "cg": head, "parallelism": "single GPU, batched chains"},
        "methods": {m: {k: v for k, v in r.items() if k != "prof"} for m, r in out.items()},
        "roofline": {
            "kernel": profile_kernel + (" even-odd" if head.startswith("eo-") else "")
            + (" (fp64)" if head in ("fp64", "eo-fp64") else " (fp32)") + ", 512 systems per launch",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": avg_ms * 1e3,
            "measured": (f"HIP events around each CG chunk of back-to-back launches in a profiled pass of "
                         f"{prof_steps} steps after the timed region, divided by the iterations done in it"
                         if chunked else
                         f"HIP events around every working launch in a profiled pass of {prof_steps} steps "
                         "after the timed region (interval includes the launch's dispatch latency)"),
            "launches": k_n,
            "algorithmic_bytes_per_launch": k_bytes,
        },
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--size", type=int, default=1024, help="Nt = Nx (default: the 1024^2 workload)")
    ap.add_argument("--therm", type=int, default=THERM_STEPS,
                    help="untimed thermalisation steps shared by every method")
    ap.add_argument("--method", choices=["fp64", "mixed", "learned", "eo-fp64", "eo-mixed", "all"], default="all")
    ap.add_argument("--seed", type=int, default=7, help="hot-start seed of the shared starting field")
    ap.add_argument("--profile-kernel", default="cg_normal_op")
    ap.add_argument("--prof-steps", type=int, default=3,
                    help="steps of the profiled pass after the timed region (roofline)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-plan", action="store_true",
                    help="run the oracle's CPU plan (32^2 in full, 256^2 / 512^2 / 1024^2 for 20 steps) and exit")
    ap.add_argument("--workload", choices=["headline", "c4"], default="headline",
                    help="c4: the batched 64-chain 64^2 workload instead of the 1024^2 headline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.cpu_plan:
        cpu_plan()
        return

    import yukawacolourednoise_amd as Y

    if args.workload == "c4":
        if args.method == "learned":
            sys.exit("the batched workload runs fp64, mixed, eo-fp64 and eo-mixed")
        bench_c4(Y, args.steps, args.warmup, args.therm,
                 ("fp64", "mixed", "eo-fp64", "eo-mixed") if args.method == "all" else (args.method,),
                 args.profile_kernel, min(args.prof_steps, 2))
        return

    methods = ["learned", "fp64", "mixed", "eo-fp64", "eo-mixed"] if args.method == "all" else [args.method]
    if "eo-mixed" in methods:
        methods.append("eo-mixed-polled")
    train = "learned" in methods
    lctx, phi0, train_info, therm_s = thermalise(Y, args.size, args.therm, train, args.seed)
    print(f"# thermalised {args.therm} steps in {therm_s:.1f} s"
          + (f" (training on {train_info['pairs']} pairs, loss {train_info['loss_first']:.3f} -> "
             f"{train_info['loss_last']:.3f})" if train else ""), file=sys.stderr, flush=True)
    res = {}
    for m in methods:
        res[m] = run_method(Y, lctx if m == "learned" else None, args.size, m, phi0, args.therm,
                            args.steps, args.warmup, None if m == "eo-mixed-polled" else args.profile_kernel,
                            args.prof_steps)
        if m == "learned":
            res[m]["training"] = train_info
        print(f"# {m}: {res[m]['steps_per_s']:.3f} steps/s, {res[m]['cg_iters_per_solve']:.0f} CG iters/solve",
              file=sys.stderr, flush=True)
    # headline: the fastest method (every method reaches the same FP64 true residual); the
    # polled form of eo-mixed is reported beside it, not headlined
    head = max((m for m in res if m != "eo-mixed-polled"), key=lambda m: res[m]["steps_per_s"])
    h = res[head]
    k_ms, k_n, k_bytes = h["prof"]
    avg_ms = k_ms / max(1, k_n)
    achieved = (k_bytes / 1e9) / (avg_ms / 1e3) if k_n else 0.0
    # the library brackets whole chunks when the even-odd iteration finalizes in the next
    # launch's prologue (the tuning report's last entry names the form it chose)
    chunked = head == "eo-mixed" and "prologue finalize" in h["tuning"].split("|")[-1]
    cg_desc = {"fp64": "plain FP64 CG", "mixed": "mixed-precision CG (FP32 inner, FP64 reliable update every 50)",
               "learned": "flexible CG with the online-trained BF16 CNN preconditioner (FP32 inner, FP64 reliable "
                          "update every 50)",
               "eo-fp64": "even-odd preconditioned FP64 CG (Schur complement on the even sites)",
               "eo-mixed": "even-odd preconditioned mixed-precision CG (Schur complement on the even sites, FP32 "
                           "inner, FP64 reliable update every 50)"}
    out = {
        "metric": "Langevin steps/s @1024^2 (true resid <= 1e-10)",
        "value": h["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32+f64", "learned": "bf16+f32+f64", "eo-fp64": "f64",
                  "eo-mixed": "f32+f64"}[head],
        "data": f"synthetic (Philox hot start, seed {args.seed}, {args.therm} thermalisation steps)",
        "config": {
            "workload": f"{args.size}x{args.size} near-critical Langevin step, {cg_desc[head]} on D^dag D "
                        "to FP64 true relative residual 1e-10, 1 noise vector, fixed eps 0.005",
            "Nt": args.size, "Nx": args.size, "M": WORKLOAD["M"], "g": WORKLOAD["g"], "m2": WORKLOAD["m2"],
            "lambda": WORKLOAD["lam"], "s": WORKLOAD["cutFraction"], "eps": WORKLOAD["epsBar"],
            "cg": head, "precond": {"learned": "cnn", "eo-fp64": "even-odd", "eo-mixed": "even-odd"}.get(head, "none"),
            "parallelism": "single GPU",
        },
        "cg_iters_per_solve": h["cg_iters_per_solve"],
        "cg_true_rel_resid_max": h["true_rel_resid_max"],
        "methods": {m: {k: v for k, v in r.items() if k != "prof"} for m, r in res.items()},
        "roofline": {
            "kernel": args.profile_kernel + (" even-odd" if head.startswith("eo-") else "")
            + (" (fp64)" if head in ("fp64", "eo-fp64") else " (fp32)"),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get((args.profile_kernel + ("-eo" if head.startswith("eo-") else ""),
                                             "fp64" if head in ("fp64", "eo-fp64") else "fp32", args.size)),
            "avg_launch_us": avg_ms * 1e3,
            "measured": (f"HIP events around each CG chunk of back-to-back launches in a profiled pass of "
                         f"{args.prof_steps} steps after the timed region, divided by the iterations done in it"
                         if chunked else
                         f"HIP events around every working launch in a profiled pass of {args.prof_steps} steps "
                         "after the timed region (interval includes the launch's dispatch latency)"),
            "launches": k_n,
            "algorithmic_bytes_per_launch": k_bytes,
        },
    }
    if "fp64" in res:
        out["speedup_vs_plain_fp64"] = {m: r["steps_per_s"] / res["fp64"]["steps_per_s"] for m, r in res.items()}
    if "learned" in res:
        base = res.get("mixed", res.get("fp64"))
        out["cg_iters_learned_vs_plain"] = {"learned": res["learned"]["cg_iters_per_solve"],
                                            "plain": base["cg_iters_per_solve"] if base else None}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(phi0, args.therm)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
