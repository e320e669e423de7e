"""CPU check of the index arithmetic of the fused even-odd kernel (k_eo_normal_fused in
yukawacolourednoise_amd/csrc/evenodd.hip), transcribed here: for every tile of its
tables and every lattice the tables accept, including time extents smaller than the
4-site halo (Nt = 2), every global and LDS access stays in bounds, every stage covers
its region's sites of the right parity exactly once, and the stage-4 outputs cover the
even sites of the lattice exactly once."""
import itertools

import pytest

EO64 = [(8, 32), (16, 16), (4, 64), (4, 32), (2, 32)]
EO32 = [(8, 32), (16, 32), (8, 64), (16, 16), (4, 64), (2, 32)]


def wrapm(a, n):
    m = a % n          # Python % is already non-negative; the kernel adds n when negative
    return m


def site_of(w, hh, par, TX):
    c0 = 4 - hh
    half = (TX + 2 * hh) // 2
    lr = c0 + w // half
    lc = c0 + 2 * (w - (lr - c0) * half) + (((par + lr) ^ c0) & 1)
    return lr, lc


def check(Nt, Nx, TT, TX):
    H, W, HF, WF = TT + 8, TX + 8, TT + 6, TX + 6
    Nh, Vh = Nx // 2, Nt * Nx // 2
    out = set()
    for ty, tx in itertools.product(range(Nt // TT), range(Nx // TX)):
        t0, x0 = ty * TT, tx * TX
        # stage 0: even sites of the whole array
        seen = set()
        for w in range(H * (W // 2)):
            lr = w // (W // 2)
            lc = 2 * (w - lr * (W // 2)) + (lr & 1)
            assert 0 <= lr < H and 0 <= lc < W and (lr + lc) % 2 == 0
            gt, gx = wrapm(t0 - 4 + lr, Nt), wrapm(x0 - 4 + lc, Nx)
            assert (gt + gx) % 2 == 0
            j = gt * Nh + (gx >> 1)
            assert 0 <= j < Vh
            seen.add((lr, lc))
        assert len(seen) == H * W // 2
        for w in range(HF * WF):
            lr, lc = w // WF, w % WF
            gt, gx = wrapm(t0 - 3 + lr, Nt), wrapm(x0 - 3 + lc, Nx)
            assert 0 <= gt * Nx + gx < Nt * Nx
        # stages 1-4
        for hh, par in [(3, 1), (2, 0), (1, 1), (0, 0)]:
            n = (TT + 2 * hh) * ((TX + 2 * hh) // 2)
            cells = set()
            for w in range(n):
                lr, lc = site_of(w, hh, par, TX)
                assert (lr + lc) % 2 == par
                assert 4 - hh <= lr < 4 + TT + hh and 4 - hh <= lc < 4 + TX + hh
                for dr, dc in [(1, 0), (-1, 0), (0, 1), (0, -1)]:
                    assert 0 <= lr + dr < H and 0 <= lc + dc < W
                assert 0 <= (lr - 1) * WF + (lc - 1) < HF * WF
                cells.add((lr, lc))
                if hh == 0:
                    gt, gx = wrapm(t0 - 4 + lr, Nt), wrapm(x0 - 4 + lc, Nx)
                    j = gt * Nh + (gx >> 1)
                    assert 0 <= j < Vh
                    assert j not in out
                    out.add(j)
            assert len(cells) == n
    assert out == set(range(Vh))


@pytest.mark.parametrize("TT,TX", sorted(set(EO64 + EO32)))
def test_eo_fused_indexing(TT, TX):
    shapes = [(Nt, Nx) for Nt in (2, 4, 6, 8, 10, 16, 22, 32) for Nx in (16, 32, 64, 96)
              if Nt % TT == 0 and Nx % TX == 0]
    assert shapes
    for Nt, Nx in shapes:
        check(Nt, Nx, TT, TX)
