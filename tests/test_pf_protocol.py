"""CPU check of the two finalize forms of the pipelined even-odd FP32 iteration
(yukawacolourednoise_amd/csrc/evenodd.hip), transcribed here: the separate finalize
(k_fin_pf after every launch) and the prologue finalize (launch i of a chunk decides for
launch i - 1 from the previous slot of block partials, a stopping launch records the
pause, k_fin_pf_chunk closes the chunk).  Launch outputs are modelled as a deterministic
function of the (alpha, beta) a launch uses and of its iteration number, so any
difference in the protocol shows up as different partials.  Both forms must hand every
working launch the same alpha and beta, run the same launches, and leave the same state
(iterations, rho, previous rho, pause) at every chunk end, over random sequences that
include pauses at the stop, at the cap, on a non-positive <p, Ap> and for the drop since
the last reliable update, on the first and on the last launch of a chunk."""
import math
import random

import pytest

TOL_REL_TRUE = 1
RELIABLE_DROP = 1e-5                           # kReliableDrop


class State:
    def __init__(self, rr, stop, max_iter):
        self.iter = 0
        self.iter0 = 0
        self.rr = rr
        self.rr_prev = rr
        self.rr_rel = rr                       # rrRel: rho at the last anchor
        self.alpha = 0.0
        self.beta = 0.0
        self.stop = stop
        self.max_iter = max_iter
        self.done = False

    def key(self):
        return (self.iter, self.rr, self.rr_prev, self.done)


def pf_decide(mu, nu, rho, it, S):
    """pf_decide: stop rule and the next alpha, beta"""
    stop = (it >= S.max_iter) or (math.sqrt(rho) <= S.stop and it % 2 == 0) or \
        (not mu > 0.0 and it % 2 == 0) or (rho < RELIABLE_DROP ** 2 * S.rr_rel and it % 2 == 0) or \
        not math.isfinite(mu) or not math.isfinite(rho)
    alpha = rho / mu if mu > 0.0 else 0.0
    rho_next = alpha * alpha * nu - rho
    if not rho_next > 0.0:
        rho_next = 0.0
    beta = rho_next / rho if rho > 0.0 else 0.0
    return alpha, beta, stop


def launch_output(alpha, beta, it, seed):
    """block-partial sums (mu, nu, rho) of the launch computing iteration `it`: a fixed
    function of what the launch used; rho decays, mu turns non-positive now and then"""
    rng = random.Random(hash((round(alpha, 12), round(beta, 12), it, seed)))
    rate = 0.08 if seed % 2 else 0.5            # fast decay: pauses for the drop since the anchor
    rho = 10.0 ** (-rate * it) * (0.5 + rng.random())
    mu = rho * (0.2 + rng.random()) * (-1.0 if rng.random() < 0.02 else 1.0)
    nu = mu * (1.0 + 3.0 * rng.random())
    return mu, nu, rho


def run_separate(S, n, seed, log):
    """n launches, each followed by k_fin_pf"""
    for i in range(n):
        if S.done:
            continue
        it = S.iter + 1
        log.append((it, S.alpha, S.beta))
        mu, nu, rho = launch_output(S.alpha, S.beta, it, seed)
        S.iter = it
        a, b, stop = pf_decide(mu, nu, rho, S.iter, S)
        S.rr_prev, S.rr = S.rr, rho
        if stop:
            S.done = True
            continue
        S.alpha, S.beta = a, b


def run_prologue(S, n, seed, log):
    """n launches with finMode 1, then k_fin_pf_chunk"""
    slots = [None, None]                       # block partials, slot idx & 1
    for idx in range(n):
        if S.done:                             # blocks exit at once
            continue
        alpha, beta = S.alpha, S.beta          # launch 0 starts from the state
        if idx > 0:
            mu, nu, rho = slots[(idx & 1) ^ 1]
            a, b, stop = pf_decide(mu, nu, rho, S.iter0 + idx, S)
            if stop:                           # block 0 records the pause
                rprev = S.rr if idx == 1 else slots[idx & 1][2]
                S.iter = S.iter0 + idx
                S.rr_prev, S.rr = rprev, rho
                S.done = True
                continue
            alpha, beta = a, b
        it = S.iter0 + idx + 1
        log.append((it, alpha, beta))
        slots[idx & 1] = launch_output(alpha, beta, it, seed)
    if not S.done:                             # k_fin_pf_chunk
        mu, nu, rho = slots[(n - 1) & 1]
        S.iter = S.iter0 + n
        _, _, stop = pf_decide(mu, nu, rho, S.iter, S)
        S.rr_prev, S.rr = slots[n & 1][2], rho
        if stop:
            S.done = True


def reliable_update(S, seed):
    """k_fin_eo_pf_true's resume (the true residual modelled as a fixed function)"""
    rng = random.Random(hash((S.iter, seed, "ru")))
    rn = S.rr * (0.8 + 0.4 * rng.random())
    S.alpha = 0.0
    S.beta = rn / S.rr_prev if S.rr_prev > 0.0 else 0.0
    S.rr = rn
    S.rr_rel = rn
    S.iter0 = S.iter
    S.done = False


@pytest.mark.parametrize("seed", range(40))
def test_prologue_finalize_equals_separate(seed):
    rng = random.Random(seed)
    max_iter = rng.choice([30, 57, 120, 400])
    stop = 10.0 ** (-rng.uniform(1.0, 6.0 if seed % 2 else 16.0))
    chunks = [rng.choice([8, 16, 24, 50]) for _ in range(12)]
    A, B = State(1.0, stop, max_iter), State(1.0, stop, max_iter)
    la, lb = [], []
    for k, n in enumerate(chunks):
        run_separate(A, n, seed, la)
        run_prologue(B, n, seed, lb)
        assert la == lb, (k, n)
        assert A.key() == B.key(), (k, n, A.key(), B.key())
        if A.iter >= max_iter:
            break
        reliable_update(A, seed)
        reliable_update(B, seed)
    assert len(la) > 0


def test_pause_on_first_and_last_launch_of_a_chunk():
    """a chunk whose launch 1 stops (previous rho from the state) and a chunk that ends
    exactly at the cap (the chunk-end kernel pauses)"""
    for max_iter, n in [(2, 8), (8, 8), (16, 8)]:
        A, B = State(1.0, 1e-30, max_iter), State(1.0, 1e-30, max_iter)
        la, lb = [], []
        for _ in range(4):
            run_separate(A, n, 7, la)
            run_prologue(B, n, 7, lb)
            assert la == lb and A.key() == B.key(), (max_iter, n)
            if A.iter >= max_iter:
                break
            reliable_update(A, 7)
            reliable_update(B, 7)
        assert A.iter == max_iter and A.done
