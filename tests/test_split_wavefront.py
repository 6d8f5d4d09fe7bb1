"""CPU emulation of the far updates' split wavefront schedule.

The stage-2 far-update kernels (k_s2_rows_w2 / k_s2_cols_w2 in
paraht_amd/csrc/stage2.hip) apply a bulge block's NQ rotation sequences
(sequence icm applies ig = R-1..1 to window entries (icm+ig-1, icm+ig), the
order of rht2ht.f90:1019-1325) with L lanes per row: lane h takes the
sequences [h NQ/L, (h+1) NQ/L) and all lanes run one instruction stream, a
wavefront with R' = R + (L-1) NQ/L, NQ' = NQ/L over registers shifted by
(2h-(L-1)) NQ/L.  This test restates that schedule (the closed forms of the
first and last touch of each register, the identity padding of the per-lane
coefficient tables, the zero-padded tile and the store-before-load streaming
order) and checks it against the plain sequential application, for the
shipped L = 4 and for L = 2.
"""
import numpy as np
import pytest


def wf_first(R, NQ, e):
    f = 1 << 30
    if max(0, e - R + 1) <= min(NQ - 1, e - 1):
        f = R - 1 - e + 2 * max(0, e - R + 1)
    if max(0, e + 2 - R) <= min(NQ - 1, e):
        f = min(f, R - 2 - e + 2 * max(0, e + 2 - R))
    return f


def wf_last(R, NQ, e):
    last = -1
    if max(0, e - R + 1) <= min(NQ - 1, e - 1):
        last = R - 1 - e + 2 * min(NQ - 1, e - 1)
    if max(0, e + 2 - R) <= min(NQ - 1, e):
        last = max(last, R - 2 - e + 2 * min(NQ - 1, e))
    return last


def _rot(w, hi, lo, c, s):
    x, y = w[hi], w[lo]
    w[hi] = c * x + s * y
    w[lo] = c * y - s * x


def sequential(x0, C, S, R, NQ, icm0):
    x = x0.copy()
    for icm in range(icm0, NQ):
        for ig in range(R - 1, 0, -1):
            _rot(x, icm + ig, icm + ig - 1, C[icm, ig], S[icm, ig])
    return x


def split(x0, C, S, R, NQ, icm0, L=4, store_first=True):
    NQL = NQ // L
    RP, NW, PAD = R + (L - 1) * NQL, R + NQ - 1, (L - 1) * NQL
    tile = np.zeros(NW + 2 * PAD)
    tile[PAD:PAD + NW] = x0
    tb = np.zeros((L, NQL, RP, 2))
    tb[..., 0] = 1.0
    for h in range(L):
        for i in range(NQL):
            for igp in range(RP):
                ig, icm = igp - (L - 1 - h) * NQL, h * NQL + i
                if 1 <= ig <= R - 1:
                    tb[h, i, igp] = (C[icm, ig], S[icm, ig])
    w = [np.full(NW, np.nan) for _ in range(L)]
    first = [wf_first(RP, NQL, e) for e in range(NW)]
    last = [wf_last(RP, NQL, e) for e in range(NW)]

    def store(tau):
        for e in range(NW):
            if last[e] == tau:
                for h in range(L):
                    tile[e + 2 * h * NQL] = w[h][e]

    def load(tau):
        for e in range(NW):
            if first[e] == tau + 1:
                for h in range(L):
                    w[h][e] = tile[e + 2 * h * NQL]

    load(-1)
    for tau in range(RP + NQL - 2):
        for h in range(L):
            for i in range(NQL):
                igp = RP - 1 - tau + i
                if 1 <= igp <= RP - 1:
                    c, s = tb[h, i, igp]
                    if i < icm0 - h * NQL:
                        c, s = 1.0, 0.0
                    _rot(w[h], i + igp, i + igp - 1, c, s)
        if store_first:
            store(tau)
            load(tau)
        else:
            load(tau)
            store(tau)
    assert np.all(tile[:PAD] == 0) and np.all(tile[PAD + NW:] == 0)
    return tile[PAD:PAD + NW]


PAIRS = [(8, 4), (8, 8), (16, 4), (16, 8), (16, 16), (32, 8), (32, 16), (32, 32),
         (64, 8), (64, 16)]


@pytest.mark.parametrize("L", [2, 4])
@pytest.mark.parametrize("R,NQ", PAIRS)
def test_split_matches_sequential(R, NQ, L):
    rng = np.random.default_rng(R * 100 + NQ)
    th = rng.uniform(0, 2 * np.pi, (NQ, R))
    C, S = np.cos(th), np.sin(th)
    x0 = rng.standard_normal(R + NQ - 1)
    for icm0 in range(NQ):
        ref = sequential(x0, C, S, R, NQ, icm0)
        got = split(x0, C, S, R, NQ, icm0, L=L)
        assert np.array_equal(got, ref), (R, NQ, L, icm0)


def test_loads_before_stores_break_the_hand_off():
    # the same schedule with each step's loads issued before its stores
    # reads an entry before the other lane has written it back
    R, NQ = 16, 8
    rng = np.random.default_rng(7)
    th = rng.uniform(0, 2 * np.pi, (NQ, R))
    C, S = np.cos(th), np.sin(th)
    x0 = rng.standard_normal(R + NQ - 1)
    ref = sequential(x0, C, S, R, NQ, 0)
    got = split(x0, C, S, R, NQ, 0, L=2, store_first=False)
    assert not np.allclose(got, ref)
