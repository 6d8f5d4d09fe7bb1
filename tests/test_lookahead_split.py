"""Stage-2 lookahead split (CPU): a sweep's A/B far update is cut into a
lookahead part (applied first, on the main stream, before the next sweep's
generation) and the rest (applied beside the next generation on a second
stream), after the reference's lookahead/bulk split (src/rht2ht.f90:1630-1772
and 1803-1958).  The library (paraht.cpp, stage2) splits
  * the rows above the window: for block ib's sequence icm, the rows within
    Lr = nq+3 blocks of the window's first row go to the lookahead, the rows
    above them to the rest;
  * the columns right of the window at Lc = nq+2 blocks past each block's
    istop;
and applies lookahead rows, lookahead columns, rest rows, rest columns.
Checked here on the index formulas of the kernels:

* the rest of sweep k touches no entry that the generation of sweep k+1
  touches (the smallest safe shifts are found by search);
* the split order gives the one-pass result (all rows, then all columns) bit
  for bit, on random rotation tapes.  A right rotation on columns (y, y+1)
  and a left rotation on rows (x, x+1) sharing entries must each be applied
  whole on one side of the other; a row lookahead shorter than the column
  lookahead cuts such pairs and is caught.
"""
import numpy as np
import pytest


def look_cols(nq):  # paraht.cpp: s2_look_cols
    return nq + 2


def look_rows(nq):  # paraht.cpp: s2_look_rows
    return look_cols(nq) + 1


def gen_mask(n, r, nq, s):
    """Entries (0-based) of A and B the sweep at column s (1-based) touches in
    k_s2_sweep (chase_step in s2_dev.hpp): the deferred mat-vec, the left
    chain's column, the B block and the window rows (leaders and helpers)."""
    ldr = r + nq
    GA = np.zeros((n, n), bool)
    GB = np.zeros((n, n), bool)
    for ic in range(1, nq + 1):
        jc = s + ic - 1
        ib = 0
        while True:
            ib += 1
            if ib == 1:
                if jc > n:
                    break
                col, row0 = jc, jc + 1
            else:
                col = jc + 1 + (ib - 2) * r
                if col > n:
                    break
                row0 = col + r
            nb = min(r, n - row0 + 1)
            base = s + (ib - 1) * r
            L = min(ldr - 1, n - base)
            if ic > 1 and L > 0:
                GA[base:base + L, col - 1] = True
                if row0 + r - 1 <= n:
                    GB[base:base + L, row0 + r - 2] = True
            if nb >= 2:
                istart = max(s + 1, s + 1 + (ib - 1 - nq + ic) * r)
                R0, C0 = row0 - 1, col - 1
                GA[R0:R0 + nb, C0] = True
                GB[R0:R0 + nb, R0:R0 + nb] = True
                aend = min(n, row0 + r + nb - 1)
                GA[istart - 1:aend, R0:R0 + nb] = True
                GB[istart - 1:row0 - 1, R0:R0 + nb] = True
    return GA, GB


def thr(s, r, nq, x, icm):
    return max(s + 1, s + 1 + (x - nq + icm) * r)


def istop(which, s, r, nq, ib):
    if which == 0:
        return s + nq - 1 if ib == 1 else s + (ib - 2) * r + nq
    return s + (ib - 1) * r + r + nq - 1


def nblocks(n, r, s):
    return (n - s) // r + ((n - s) % r != 0)


def row_range(i, s, r, nq, ib, icm, lo_shift=-1, hi_shift=0, row_end=None):
    """S2Range row test of s2_far.hpp (row_in_range)."""
    if row_end is not None and i >= row_end:
        return False
    return i + 1 < thr(s, r, nq, ib - hi_shift, icm) and \
        (lo_shift < 0 or thr(s, r, nq, ib - lo_shift, icm) <= i + 1)


def col_range(j, ist, r, c_lo, c_hi):
    return j + 1 > ist + c_lo * r and (c_hi < 0 or j + 1 <= ist + c_hi * r)


def rest_mask(n, r, nq, s, Lr, Lc):
    """Entries the rest of the sweep's far update touches: for block ib's
    sequence icm the window entries icm..icm+r-1 of the rows
    i+1 < thr(ib - Lr, icm), and the whole window of every column more than
    Lc blocks right of a block's istop."""
    FA = np.zeros((n, n), bool)
    FB = np.zeros((n, n), bool)
    for ib in range(1, nblocks(n, r, s) + 1):
        cw0 = s + (ib - 1) * r
        W = min(r + nq - 1, n - cw0)
        if W <= 0:
            continue
        for icm in range(nq):
            hi = min(n, thr(s, r, nq, ib - Lr, icm) - 1)
            c0, c1 = cw0 + icm, min(cw0 + W, cw0 + icm + r)
            FA[:hi, c0:c1] = True
            FB[:hi, c0:c1] = True
        for which, F in ((0, FA), (1, FB)):
            j0 = max(0, istop(which, s, r, nq, ib) + Lc * r)
            F[cw0:cw0 + W, j0:] = True
    return FA, FB


CONFIGS = [(60, 4, 2), (90, 8, 4), (120, 8, 8), (130, 16, 4), (200, 16, 8), (170, 16, 16),
           (140, 5, 3)]


@pytest.mark.parametrize("n,r,nq", CONFIGS)
def test_rest_is_disjoint_from_next_generation(n, r, nq):
    Lr, Lc = look_rows(nq), look_cols(nq)
    need_r = need_c = 0
    for s in range(1, n - 1, nq):
        if s + nq > n - 2:
            break
        GA, GB = gen_mask(n, r, nq, s + nq)
        FA, FB = rest_mask(n, r, nq, s, Lr, Lc)
        assert not (FA & GA).any() and not (FB & GB).any(), (s, Lr, Lc)
        # the smallest shifts that would do, for the record
        while True:
            FA, FB = rest_mask(n, r, nq, s, need_r, 10 ** 6)
            if not (FA & GA).any() and not (FB & GB).any():
                break
            need_r += 1
        while True:
            FA, FB = rest_mask(n, r, nq, s, 10 ** 6, need_c)
            if not (FA & GA).any() and not (FB & GB).any():
                break
            need_c += 1
    assert need_r <= Lr and need_c <= Lc
    # without a lookahead the rest would collide with the next generation
    assert need_r >= 1 and need_c >= 1


def _rotate_rows(M, rows, cols_pair, c, sn):
    """Right rotation of columns (a, b) = cols_pair on the given rows, in the
    far-update kernels' form (x = col b, y = col a): b <- c x + s y,
    a <- c y - s x."""
    a, b = cols_pair
    x = M[rows, b].copy()
    y = M[rows, a].copy()
    M[rows, b] = c * x + sn * y
    M[rows, a] = c * y - sn * x


def apply_rows(M, n, r, nq, s, cs, lo_shift=-1, hi_shift=0):
    """Sequence icm of block ib (descending ib) on the rows in range:
    rotations ig = r-1..1 on window entries (icm+ig-1, icm+ig)."""
    idx = np.arange(n)
    for ib in range(nblocks(n, r, s), 0, -1):
        cw0 = s + (ib - 1) * r
        W = min(r + nq - 1, n - cw0)
        if W <= 0:
            continue
        for icm in range(nq):
            act = np.array([row_range(i, s, r, nq, ib, icm, lo_shift, hi_shift) for i in idx])
            rows = idx[act]
            if rows.size == 0:
                continue
            for ig in range(r - 1, 0, -1):
                e = icm + ig
                if e >= W:
                    continue
                c, sn = cs[ib - 1, icm, ig]
                _rotate_rows(M, rows, (cw0 + e - 1, cw0 + e), c, sn)


def apply_cols(M, n, r, nq, s, cs, c_lo, c_hi):
    """Left rotations of block ib (descending) on the columns in range
    (rows (icm+ig-1, icm+ig) of the window)."""
    idx = np.arange(n)
    for ib in range(nblocks(n, r, s), 0, -1):
        rw0 = s + (ib - 1) * r
        W = min(r + nq - 1, n - rw0)
        if W <= 0:
            continue
        for which in (0, 1):
            ist = istop(which, s, r, nq, ib)
            cols = idx[[col_range(j, ist, r, c_lo, c_hi) for j in idx]]
            if cols.size == 0:
                continue
            for icm in range(nq):
                for ig in range(r - 1, 0, -1):
                    e = icm + ig
                    if e >= W:
                        continue
                    c, sn = cs[ib - 1, icm, ig]
                    a, b = rw0 + e - 1, rw0 + e
                    x = M[which][a, cols].copy()
                    y = M[which][b, cols].copy()
                    M[which][a, cols] = c * x + sn * y
                    M[which][b, cols] = c * y - sn * x


def _tape(rng, n, r, nq, s):
    """Random rotations, identities where a rotation would reach past the
    matrix (as the generation leaves the tape)."""
    nbl = nblocks(n, r, s)
    th = rng.uniform(0, 2 * np.pi, size=(nbl, nq, r))
    cs = np.stack([np.cos(th), np.sin(th)], axis=-1)
    for ib in range(1, nbl + 1):
        W = min(r + nq - 1, n - (s + (ib - 1) * r))
        for icm in range(nq):
            for ig in range(1, r):
                if icm + ig >= W:
                    cs[ib - 1, icm, ig] = (1.0, 0.0)
    return cs


def _one_pass(M0, n, r, nq, s, cs):
    """The one-stream schedule: every row update, then every column update."""
    A = [m.copy() for m in M0]
    for m in A:
        apply_rows(m, n, r, nq, s, cs)
    apply_cols(A, n, r, nq, s, cs, 0, -1)
    return A


def _split(M0, n, r, nq, s, cs, Lr=None):
    """The split schedule: lookahead rows and columns, then the rest."""
    Lc = look_cols(nq)
    Lr = look_rows(nq) if Lr is None else Lr
    A = [m.copy() for m in M0]
    for m in A:
        apply_rows(m, n, r, nq, s, cs, Lr, 0)
    apply_cols(A, n, r, nq, s, cs, 0, Lc)
    for m in A:
        apply_rows(m, n, r, nq, s, cs, -1, Lr)
    apply_cols(A, n, r, nq, s, cs, Lc, -1)
    return A


@pytest.mark.parametrize("n,r,nq", [(130, 8, 4), (200, 16, 4), (140, 5, 3), (150, 8, 8),
                                    (260, 16, 8)])
def test_split_order_equals_one_pass(n, r, nq):
    rng = np.random.default_rng(n + r + nq)
    for s in (1, 1 + nq, 1 + 3 * nq):
        cs = _tape(rng, n, r, nq, s)
        M0 = [rng.standard_normal((n, n)), rng.standard_normal((n, n))]
        one = _one_pass(M0, n, r, nq, s, cs)
        two = _split(M0, n, r, nq, s, cs)
        for a, b in zip(one, two):
            assert np.array_equal(a, b), s


def test_short_row_lookahead_is_caught():
    """A row lookahead of nq blocks (shorter than the column lookahead) cuts
    a right rotation's pair of rows around a left rotation that shares its
    entries: the combined update is wrong, although the rows alone and the
    columns alone still agree with the one-pass order."""
    n, r, nq = 130, 8, 4
    rng = np.random.default_rng(3)
    s = 1
    cs = _tape(rng, n, r, nq, s)
    M0 = [rng.standard_normal((n, n)), rng.standard_normal((n, n))]
    rows1 = M0[0].copy()
    apply_rows(rows1, n, r, nq, s, cs)
    rows2 = M0[0].copy()
    apply_rows(rows2, n, r, nq, s, cs, nq, 0)
    apply_rows(rows2, n, r, nq, s, cs, -1, nq)
    assert np.array_equal(rows1, rows2)
    one = _one_pass(M0, n, r, nq, s, cs)
    bad = _split(M0, n, r, nq, s, cs, Lr=nq)
    assert np.linalg.norm(one[0] - bad[0]) > 1e-3 * np.linalg.norm(one[0])
