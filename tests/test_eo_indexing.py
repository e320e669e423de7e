"""CPU check of the index arithmetic of the fused even-odd kernel (k_eo_normal_fused in
yukawacolourednoise_amd/csrc/evenodd.hip), transcribed here, with its division-free
modulo (wrapf) emulated in single precision: for every tile of its
tables and every lattice the tables accept, including time extents smaller than the
4-site halo (Nt = 2), every global and LDS access stays in bounds, the parity-split LDS
slots used for each neighbour hold exactly that neighbour, every stage covers its
region's sites of the right parity once, and stage 4 writes every even site of the
lattice once."""
import itertools

import numpy as np
import pytest

EO64 = [(8, 32), (16, 16), (4, 64), (4, 32), (2, 32)]
EO32 = [(8, 32), (16, 32), (8, 64), (16, 16), (4, 64), (2, 32), (32, 32), (16, 64)]


def wrapf(a, n):
    """The kernel's division-free a mod n: a - n floor(fl(fl(a + 1/2) * fl(1/n))), every
    operation rounded to single precision as on the GPU."""
    a = np.asarray(a, dtype=np.int64)
    inv = np.float32(1.0) / np.float32(n)
    q = np.floor((a.astype(np.float32) + np.float32(0.5)) * inv).astype(np.int64)
    return a - n * q


def wrapm(a, n):
    m = int(wrapf(a, n))
    assert m == a % n
    return m


def test_wrapf_is_exact_modulo():
    """Every extent up to 4096 over the coordinate range the kernel wraps (a tile's halo
    reaches 4 sites past either end, also for extents below 4), and large extents near
    the 2^21 limit the library accepts."""
    for n in range(1, 4097):
        a = np.arange(-4 * n - 8, 2 * n + 40)
        assert np.array_equal(wrapf(a, n), a % n), n
    rng = np.random.default_rng(0)
    for n in rng.integers(4097, 1 << 21, 200):
        a = np.concatenate([np.arange(-8, 40), np.arange(n - 40, n + 40), rng.integers(-(1 << 21), 1 << 22, 4000)])
        assert np.array_equal(wrapf(a, int(n)), a % int(n)), n


def slot_col(row, k, par):
    """array column held by slot (row, k) of the parity-`par` array"""
    return 2 * k + ((row + par) & 1)


def site_of(w, hh, par, TX, K):
    c0 = 4 - hh
    half = (TX + 2 * hh) // 2
    row = c0 + w // half
    col = c0 + 2 * (w - (row - c0) * half) + (((par + row) ^ c0) & 1)
    k = col >> 1
    i = row * K + k
    return row, col, i, i + K, i - K, row * K + ((col + 1) >> 1), row * K + ((col - 1) >> 1)


def check(Nt, Nx, TT, TX):
    H, K = TT + 8, (TX + 8) // 2
    HO = H - 2   # the odd-parity arrays hold rows 1 .. H - 2: odd entry (row, k) at (row - 1) K + k
    Nh, Vh = Nx // 2, Nt * Nx // 2
    out = set()
    for ty, tx in itertools.product(range(Nt // TT), range(Nx // TX)):
        t0, x0 = ty * TT, tx * TX
        for w in range(H * K):
            row, k = w // K, w % K
            gt = wrapm(t0 - 4 + row, Nt)
            colE, colO = 2 * k + (row & 1), 2 * k + 1 - (row & 1)
            assert (row + colE) % 2 == 0 and (row + colO) % 2 == 1
            gxE, gxO = wrapm(x0 - 4 + colE, Nx), wrapm(x0 - 4 + colO, Nx)
            assert (gt + gxE) % 2 == 0
            assert 0 <= gt * Nh + (gxE >> 1) < Vh
            assert 0 <= gt * Nx + gxE < Nt * Nx and 0 <= gt * Nx + gxO < Nt * Nx
            if 1 <= row <= HO:   # stage 0 stores the odd diagonal of rows 1 .. H - 2
                assert 0 <= w - K < HO * K
        for hh, par in [(3, 1), (2, 0), (1, 1), (0, 0)]:
            n = (TT + 2 * hh) * ((TX + 2 * hh) // 2)
            cells = set()
            for w in range(n):
                row, col, i, itf, itb, ixf, ixb = site_of(w, hh, par, TX, K)
                assert (row + col) % 2 == par
                assert slot_col(row, col >> 1, par) == col and i == row * K + (col >> 1)
                for idx in (i, itf, itb, ixf, ixb):
                    assert 0 <= idx < H * K
                # odd-parity accesses (the site itself in stages 1 and 3, the neighbours in
                # stages 2 and 4) land in the odd arrays' rows 1 .. H - 2
                odd = (i,) if par == 1 else (itf, itb, ixf, ixb)
                for idx in odd:
                    assert 0 <= idx - K < HO * K
                # neighbour slots of the other parity hold (row +- 1, col), (row, col +- 1)
                q = 1 - par
                assert slot_col(row + 1, itf % K, q) == col and itf // K == row + 1
                assert slot_col(row - 1, itb % K, q) == col and itb // K == row - 1
                assert slot_col(row, ixf % K, q) == col + 1 and ixf // K == row
                assert slot_col(row, ixb % K, q) == col - 1 and ixb // K == row
                cells.add((row, col))
                if hh == 0:   # the tile's own sites: written without a wrap
                    gt, gx = t0 - 4 + row, x0 - 4 + col
                    assert 0 <= gt < Nt and 0 <= gx < Nx
                    j = gt * Nh + (gx >> 1)
                    assert 0 <= j < Vh and j not in out
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


@pytest.mark.parametrize("nTiles,grid", [(2048, 1024), (2048, 512), (4096, 1024), (64, 64), (24, 8), (5, 5)])
def test_persistent_tile_walk(nTiles, grid):
    """Persistent launches: block b walks tiles b, b + grid, ...; with the XCD remap
    (nTiles % 8 == 0, grid % 8 == 0) every tile is visited once and block b's tiles lie in
    the band of XCD b % 8."""
    remap = nTiles % 8 == 0
    seen = []
    for b in range(grid):
        t = b
        while t < nTiles:
            u = (t % 8) * (nTiles // 8) + t // 8 if remap else t
            if remap:
                assert u // (nTiles // 8) == b % 8
            seen.append(u)
            t += grid
    assert sorted(seen) == list(range(nTiles))
