"""CPU checks of the GPU schedule of the Householder stage 2 (rht2ht_householder,
src/rht2ht.f90:1346-2051), on the numpy emulation in hh_emul.py:
  * the emulation restates the C oracle (1e-12 of ||A||);
  * the pipelined generation (step (ic, ib) after column ic-1's step ib+2,
    helpers beside the next steps when nq <= r+1) is bit-identical to the
    reference's column order, and a lag of 1 or asynchronous helpers at
    nq > r+1 are not (planted);
  * the per-row masked compact-WY row updates equal the reference's dlarf
    edges + dlarfb bulk to rounding, split (lookahead, rest) or in one pass;
  * sweep k+1's generation touches no entry that sweep k's far updates read or
    write (poisoned entries), so the far updates can run beside it.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))

import hh_emul as H  # noqa: E402
import oracle as O  # noqa: E402
from paraht_amd.pencils import random_pencil  # noqa: E402

CASES = [(60, 8, 4), (97, 8, 8), (130, 16, 8), (75, 4, 6), (50, 8, 9), (41, 3, 2)]


def _band(n, r, seed=5):
    A0, B0 = random_pencil(n, seed)
    A = np.asfortranarray(A0.copy())
    B = np.asfortranarray(B0.copy())
    Q = np.asfortranarray(np.eye(n))
    Z = np.asfortranarray(np.eye(n))
    O.ab2rht_simple(n, A, B, Q, Z, r, 2)
    return A, B, Q, Z


def _copy(ms):
    return [m.copy() for m in ms]


def _maxdiff(xs, ys):
    return max(float(np.abs(x - y).max()) for x, y in zip(xs, ys))


@pytest.mark.parametrize("n,r,nq", CASES)
def test_emulation_restates_oracle(n, r, nq):
    ms = _band(n, r)
    ref = [m.copy(order="F") for m in ms]
    O.rht2ht_householder(n, *ref, r, nq)
    em = _copy(ms)
    H.reduce_sweeps(*em, r, nq)
    assert _maxdiff(ref, em) <= 1e-12 * np.linalg.norm(ms[0])


@pytest.mark.parametrize("n,r,nq", CASES)
def test_pipelined_order_is_bit_identical(n, r, nq):
    ms = _band(n, r)
    em = _copy(ms)
    H.reduce_sweeps(*em, r, nq)
    for seed in range(3):
        rnd = _copy(ms)
        H.reduce_sweeps(*rnd, r, nq, rng=np.random.default_rng(seed))
        assert _maxdiff(em, rnd) == 0.0


def test_planted_lag_one_breaks_order():
    n, r, nq = 97, 8, 8
    ms = _band(n, r)
    em = _copy(ms)
    H.reduce_sweeps(*em, r, nq)
    worst = 0.0
    for seed in range(3):
        rnd = _copy(ms)
        A, B, Q, Z = rnd
        s = 1
        while s <= n - 2:
            sw = H.Sweep(A, B, Q, Z, r, nq, s)
            sw.generate(np.random.default_rng(seed), lag=1)
            sw.factors()
            sw.lookahead(B)
            sw.lookahead(A)
            sw.far(B)
            sw.far(A)
            sw.qz()
            s += nq
        worst = max(worst, _maxdiff(em, rnd))
    assert worst > 1e-6


def test_planted_async_helpers_above_r_plus_one_break_order():
    n, r, nq = 75, 4, 6  # nq > r + 1
    ms = _band(n, r)
    em = _copy(ms)
    H.reduce_sweeps(*em, r, nq)
    orig = H.Sweep.__init__

    def forced(self, *a, **k):
        orig(self, *a, **k)
        self.async_helpers = True

    H.Sweep.__init__ = forced
    try:
        rnd = _copy(ms)
        H.reduce_sweeps(*rnd, r, nq, rng=np.random.default_rng(1))
    finally:
        H.Sweep.__init__ = orig
    assert _maxdiff(em, rnd) > 1e-6


@pytest.mark.parametrize("n,r,nq", CASES)
def test_masked_row_updates_match_reference_form(n, r, nq):
    ms = _band(n, r)
    em = _copy(ms)
    H.reduce_sweeps(*em, r, nq)
    mk = _copy(ms)
    H.reduce_sweeps(*mk, r, nq, masked=True)
    assert _maxdiff(em, mk) <= 1e-12 * np.linalg.norm(ms[0])


@pytest.mark.parametrize("n,r,nq", CASES)
def test_one_pass_row_then_column_updates_match(n, r, nq):
    """The single-stream order: per matrix, every block's right reflectors on
    each row over its whole range, then every block's left reflectors on the
    columns right of the window."""
    ms = _band(n, r)
    em = _copy(ms)
    H.reduce_sweeps(*em, r, nq)
    op = _copy(ms)
    H.reduce_sweeps(*op, r, nq, one_pass=True)
    assert _maxdiff(em, op) <= 1e-12 * np.linalg.norm(ms[0])


@pytest.mark.parametrize("n,r,nq", CASES)
def test_far_updates_commute_with_next_generation(n, r, nq):
    ms = _band(n, r)
    em = _copy(ms)
    H.reduce_sweeps(*em, r, nq)
    nf = _copy(ms)
    H.reduce_sweeps(*nf, r, nq, next_first=True)
    assert _maxdiff(em, nf) == 0.0


@pytest.mark.parametrize("n,r,nq", [(97, 8, 8), (130, 16, 8), (75, 4, 6)])
def test_next_generation_never_touches_far_entries(n, r, nq):
    """Poison the entries sweep k's far updates change; sweep k+1's generation
    (and its T factors) must neither read them (no NaN reaches what it writes)
    nor write them."""
    A, B, Q, Z = _band(n, r)
    s = 1
    checked = 0
    while s + nq <= n - 2:
        sw = H.Sweep(A, B, Q, Z, r, nq, s)
        sw.generate()
        sw.factors()
        sw.lookahead(B)
        sw.lookahead(A)
        # the entries the far updates change
        fA, fB = A.copy(), B.copy()
        tmp = H.Sweep(fA, fB, Q.copy(), Z.copy(), r, nq, s)
        tmp.U, tmp.V, tmp.tauU, tmp.tauV = sw.U, sw.V, sw.tauU, sw.tauV
        tmp.TU, tmp.TV = sw.TU, sw.TV
        tmp.far(fB)
        tmp.far(fA)
        farA, farB = fA != A, fB != B
        # next generation on poisoned copies
        pA, pB = A.copy(), B.copy()
        pA[farA] = np.nan
        pB[farB] = np.nan
        nxt = H.Sweep(pA, pB, Q.copy(), Z.copy(), r, nq, s + nq)
        nxt.generate()
        nxt.factors()
        for t in (nxt.U, nxt.V, nxt.tauU, nxt.tauV, nxt.TU, nxt.TV):
            assert np.isfinite(t).all()
        clean = H.Sweep(A.copy(), B.copy(), Q.copy(), Z.copy(), r, nq, s + nq)
        clean.generate()
        wroteA = clean.A != A
        wroteB = clean.B != B
        assert not (wroteA & farA).any() and not (wroteB & farB).any()
        assert np.isfinite(pA[wroteA]).all() and np.isfinite(pB[wroteB]).all()
        checked += int(farA.any() or farB.any())
        # continue the reduction in the reference's order
        sw.far(B)
        sw.far(A)
        sw.qz()
        s += nq
    assert checked > 0
