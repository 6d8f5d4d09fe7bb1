"""The workgroup loop of the stage-2 far-update kernels (k_s2_rows_w2 and
k_s2_cols_w2 in paraht_amd/csrc/s2_far.hpp), restated on the CPU at the
level of a workgroup's LDS tile: the launch grid (whole 64-row or 64-column
workgroups, so a workgroup may run past the range's last row), the per-block
skip tests, the carry of the window entries two consecutive blocks share,
the zero padding past the matrix, and the stores of only the entries a
row's active sequences touch.  With the library's stage-2 schedule (the
lookahead rows from the sweep's first column down and the columns within
nq+2 blocks, then the rest) it must reproduce the far update applied in one
pass by the formula (tests/test_lookahead_split.py), bit for bit for the rows
and columns separately.
"""
import numpy as np
import pytest

from test_lookahead_split import (_tape, apply_cols, apply_rows, col_range, istop,
                                  look_cols, look_rows, nblocks, row_range, thr)

CPW = 64  # rows (columns) per workgroup: 128 threads, two lanes each


def rows_kernel(M, n, R, NQ, s, cs, lo_shift, hi_shift, row_end, grid_rows):
    """k_s2_rows_w2<R, NQ, 2, true> over grid ceil(grid_rows / CPW)."""
    NW = R + NQ - 1

    for rf in range(0, grid_rows, CPW):
        rows = range(rf, rf + CPW)
        tile = {i: np.zeros(NW) for i in rows}
        carry = False
        for ib in range(nblocks(n, R, s), 0, -1):
            cw0 = s + (ib - 1) * R
            W = min(NW, n - cw0)
            if W <= 0:
                carry = False
                continue
            meet = rf + 1 < thr(s, R, NQ, ib - hi_shift, NQ - 1) and \
                not (lo_shift >= 0 and rf + CPW < thr(s, R, NQ, ib - lo_shift, 0))
            meet = meet and rf < row_end
            if not meet:
                carry = False
                continue
            for i in rows:
                valid = i < n
                t = tile[i]
                if carry:
                    t[R:R + NQ - 1] = t[0:NQ - 1].copy()
                kf = R if carry else NW
                for k in range(kf):
                    t[k] = M[i, cw0 + k] if (valid and k < W) else 0.0
                icm0, icm1 = NQ, 0
                if valid:
                    for icm in range(NQ - 1, -1, -1):
                        if row_range(i, s, R, NQ, ib, icm, lo_shift, hi_shift, row_end):
                            icm0 = icm
                            if icm1 == 0:
                                icm1 = icm + 1
                for icm in range(icm0, icm1):
                    for ig in range(R - 1, 0, -1):
                        e = icm + ig
                        c, sn = cs[ib - 1, icm, ig]
                        x, y = t[e], t[e - 1]
                        t[e] = c * x + sn * y
                        t[e - 1] = c * y - sn * x
                if icm0 < icm1:
                    for k in range(icm0, min(W, icm1 + R - 1)):
                        M[i, cw0 + k] = t[k]
            carry = True


def cols_kernel(Ms, n, R, NQ, s, cs, c_lo, c_hi, col_begin):
    """k_s2_cols_w2<R, NQ, 2> over the columns [col_begin, n)."""
    NW = R + NQ - 1
    for which in (0, 1):
        M = Ms[which]
        for j0 in range(col_begin, n, CPW):
            cmax = min(CPW, n - j0)
            tile = np.zeros((NW, CPW))
            carry = False
            for ib in range(nblocks(n, R, s), 0, -1):
                ist = istop(which, s, R, NQ, ib)
                meet = not (j0 + CPW <= ist + c_lo * R) and \
                    not (c_hi >= 0 and j0 + 1 > ist + c_hi * R)
                rw0 = s + (ib - 1) * R
                W = min(NW, n - rw0)
                if not meet or W <= 0:
                    carry = False
                    continue
                if carry:
                    tile[R:R + NQ - 1, :] = tile[0:NQ - 1, :].copy()
                kf = R if carry else NW
                for cc in range(CPW):
                    for k in range(kf):
                        tile[k, cc] = M[rw0 + k, j0 + cc] if (cc < cmax and k < W) else 0.0
                for cc in range(cmax):
                    if not col_range(j0 + cc, ist, R, c_lo, c_hi):
                        continue
                    t = tile[:, cc]
                    for icm in range(NQ):
                        for ig in range(R - 1, 0, -1):
                            e = icm + ig
                            c, sn = cs[ib - 1, icm, ig]
                            a, b = t[e - 1], t[e]
                            t[e - 1] = c * a + sn * b
                            t[e] = c * b - sn * a
                    for k in range(W):
                        M[rw0 + k, j0 + cc] = t[k]
                carry = True


def schedule(n, R, NQ, s):
    """The ranges stage2() in paraht.cpp launches for one sweep."""
    nbl = nblocks(n, R, s)
    Lr, Lc = look_rows(NQ), look_cols(NQ)
    rows_all = min(n, s + (nbl - 1) * R)
    far_rows = min(n, max(s, s + (nbl - 1 - Lr) * R))
    col0 = min(n, s + NQ - 1)
    return rows_all, far_rows, col0, Lr, Lc


@pytest.mark.parametrize("n,R,NQ", [(59, 8, 4), (130, 8, 4), (150, 8, 8), (220, 16, 8)])
def test_kernel_loop_reproduces_the_formula(n, R, NQ):
    rng = np.random.default_rng(n)
    for s in (1, 1 + NQ, 1 + 3 * NQ):
        cs = _tape(rng, n, R, NQ, s)
        rows_all, far_rows, col0, Lr, Lc = schedule(n, R, NQ, s)
        M0 = rng.standard_normal((n, n))
        one = M0.copy()
        apply_rows(one, n, R, NQ, s, cs)
        k = M0.copy()
        rows_kernel(k, n, R, NQ, s, cs, Lr, 0, rows_all, rows_all)  # lookahead rows
        rows_kernel(k, n, R, NQ, s, cs, -1, Lr, far_rows, far_rows)  # the rest
        assert np.array_equal(one, k), ("rows", s)
        C0 = [rng.standard_normal((n, n)), rng.standard_normal((n, n))]
        one = [m.copy() for m in C0]
        apply_cols(one, n, R, NQ, s, cs, 0, -1)
        kc = [m.copy() for m in C0]
        cols_kernel(kc, n, R, NQ, s, cs, 0, Lc, col0)
        cols_kernel(kc, n, R, NQ, s, cs, Lc, -1, min(n, col0 + Lc * R))
        assert np.array_equal(one[0], kc[0]) and np.array_equal(one[1], kc[1]), ("cols", s)
