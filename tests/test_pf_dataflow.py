"""CPU check of the data flow of the pipelined even-odd FP32 iteration
(k_eo_normal_fused with PF = true and k_pf_reliable in yukawacolourednoise_amd/csrc/evenodd.hip),
transcribed here in float64 on a symmetric positive definite system A x = b of 300 unknowns.

Per launch idx of a chunk, as the kernel does it:
- r_{idx-1} is read from r at launch 0 and otherwise rebuilt as p_{idx-1} - beta p_{idx-2}
  with the beta launch idx - 1 recorded in ab[(idx - 1) & 1];
- r' = r - alpha Ap_{idx-1}, p' = r' + beta p_{idx-1}, Ap' = A p';
- x_s is left alone by even launches, and odd launches add alpha_e p_{idx-2} + alpha p_{idx-1};
- the directions rotate through three buffers (p_k in slot (k + 1) mod 3), Ap through two;
- each launch records the alpha, beta it used in ab[idx & 1];
- the finalize derives alpha, beta from the partials (mu, nu, rho) by the D'Azevedo, Eijkhout
  and Romine recurrence and stops at the iteration cap.
The reliable update (k_pf_reliable, then the true residual) adds x_s to x, adds the deferred
term alpha_{m-1} p_{m-2} after an odd number m of launches, copies p_{m-2} from slot
(m - 1) mod 3 to slot 0, and restarts with alpha = 0, beta = rho_true / rho_prev.

In exact arithmetic this is plain CG with one iteration repeated after every reliable update
(the restart recomputes the last direction), so after c full chunks of n launches x must be
plain CG's x_{c (n - 1)}, and a stop at an odd cap inside the next chunk must give the iterate
of its last launch.  Planted mistakes (no deferred term at the cap, the wrong restore slot, the
rebuild with the current beta) must fail the same comparison."""
import numpy as np
import pytest


def plain_cg(A, b, n):
    x, r = np.zeros_like(b), b.copy()
    p, rho = r.copy(), r @ r
    xs = [x.copy()]
    for _ in range(n):
        Ap = A @ p
        a = rho / (p @ Ap)
        x = x + a * p
        r = r - a * Ap
        rho_new = r @ r
        p = r + (rho_new / rho) * p
        rho = rho_new
        xs.append(x.copy())
    return xs


def pipelined(A, b, chunk, cap, plant=None):
    """returns x after every reliable update (the last one after the stop at the cap)"""
    nv = len(b)
    x = np.zeros(nv)
    r = b.copy()                              # setup: r = b - A 0
    P = [np.zeros(nv) for _ in range(3)]      # p_{-1} = 0 in slot 0
    Apb = [np.zeros(nv), np.zeros(nv)]
    xs = np.zeros(nv)
    ab = [(0.0, 0.0), (0.0, 0.0)]
    alpha, beta = 0.0, 0.0
    it, out = 0, []
    rho_prev = r @ r
    while True:
        it0, rhos, stopped = it, [], False
        for idx in range(chunk):
            slot = lambda k: P[(k + 1) % 3]
            if idx == 0:
                rin = r
            else:
                bp = beta if plant == "rebuild_with_current_beta" else ab[(idx - 1) & 1][1]
                rin = slot(idx - 1) - bp * slot(idx - 2)
            par = idx & 1
            rn = rin - alpha * Apb[par]
            pn = rn + beta * slot(idx - 1)
            if par:
                xs = xs + ab[(idx - 1) & 1][0] * slot(idx - 2) + alpha * slot(idx - 1)
            ab[idx & 1] = (alpha, beta)
            Ap = A @ pn
            P[(idx + 1) % 3] = pn
            Apb[par ^ 1] = Ap
            mu, nu, rho = pn @ Ap, Ap @ Ap, rn @ rn
            rhos.append(rho)
            # finalize of this launch
            it += 1
            alpha = rho / mu
            beta = max(alpha * alpha * nu - rho, 0.0) / rho
            if it >= cap:
                stopped = True
                break
        # reliable update
        m = it - it0
        x = x + xs
        if m & 1 and plant != "no_deferred_term":
            x = x + ab[(m - 1) & 1][0] * P[(m - 1) % 3]
        src = (m - 2) % 3 if plant == "wrong_slot" else (m - 1) % 3
        P[0] = P[src].copy()
        xs = np.zeros(nv)
        out.append(x.copy())
        if stopped:
            return out
        r = b - A @ x
        rho_true = r @ r
        rho_prev = rhos[-2] if len(rhos) >= 2 else rho_prev
        alpha, beta = 0.0, rho_true / rho_prev
        rho_prev = rho_true


def system(seed, nv=300):
    """A = S^T S with condition number about 1e6: 50 iterations lower the residual only
    about twofold, so consecutive iterates stay several percent apart while the float64
    rounding of the two forms (the pipelined recursion cancels once the residual has
    fallen far) stays far below that"""
    rng = np.random.default_rng(seed)
    S = rng.standard_normal((nv, nv)) / np.sqrt(nv) + 0.3 * np.eye(nv)
    return S.T @ S, rng.standard_normal(nv)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("chunk,cap", [(8, 37), (8, 40), (6, 29), (10, 51)])
def test_pipelined_data_flow_is_cg(seed, chunk, cap):
    A, b = system(seed)
    ref = plain_cg(A, b, cap + 2)
    out = pipelined(A, b, chunk, cap)
    nfull = cap // chunk if cap % chunk else cap // chunk - 1
    for c in range(1, nfull + 1):
        assert rel(out[c - 1], ref[c * (chunk - 1)]) < 1e-8, (c, rel(out[c - 1], ref[c * (chunk - 1)]))
    last = nfull * (chunk - 1) + (cap - nfull * chunk) - 1   # the stopping chunk's launches
    assert len(out) == nfull + 1
    assert rel(out[-1], ref[last]) < 1e-8, rel(out[-1], ref[last])
    assert rel(ref[last - 1], ref[last]) > 1e-3   # iterates well apart


@pytest.mark.parametrize("plant", ["no_deferred_term", "wrong_slot", "rebuild_with_current_beta"])
def test_planted_mistakes_are_caught(plant):
    """chunks of 8, a stop at the odd cap 37 (m = 5 launches in the last chunk)"""
    A, b = system(0)
    ref = plain_cg(A, b, 40)
    good = pipelined(A, b, 8, 37)
    bad = pipelined(A, b, 8, 37, plant)
    targets = [ref[7], ref[14], ref[21], ref[28], ref[32]]
    assert max(rel(g, t) for g, t in zip(good, targets)) < 1e-8
    assert max(rel(g, t) for g, t in zip(bad, targets)) > 1e-4, plant
