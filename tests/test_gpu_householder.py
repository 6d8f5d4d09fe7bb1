"""GPU parity of the Householder stage 2 (rht2ht_householder,
src/rht2ht.f90:1346-2051): paraht_rht2ht without PARAHT_STAGE2_GIVENS, through
the C-ABI, against the oracle's restatement of the same variant.

The GPU generates the same reflectors in the same order as the reference (the
pipelined column schedule is checked bit for bit on the CPU in
test_hh_schedule.py), so the reduced pencil is compared elementwise, without
sign normalisation, to 1e-9 relative (Frobenius); the invariants hold to
10 n u with exact structural zeros.  Where the reference's own result is wrong
(nq > r with column n-1 in a sweep, the defect its Givens variant shares), the
GPU is compared with the numpy restatement of the sweep instead
(tests/hh_emul.py, which applies the skipped reflectors).
"""
import numpy as np
import pytest

import paraht_amd as ph
from conftest import U
from paraht_amd.pencils import illcond_pencil, random_pencil

pytestmark = pytest.mark.gpu


def _stage1_pencil(O, n, r, seed, illcond=False):
    A0, B0 = (illcond_pencil if illcond else random_pencil)(n, seed=seed)
    A1, B1 = A0.copy(order="F"), B0.copy(order="F")
    Q1, Z1 = np.eye(n, order="F"), np.eye(n, order="F")
    O.ab2rht_simple(n, A1, B1, Q1, Z1, r, 2)
    B1 = np.asfortranarray(np.triu(B1))
    return A0, B0, A1, B1, Q1, Z1


def _run_gpu(n, mats, r, nq, flags_env=None):
    got = [M.copy(order="F") for M in mats]
    ph.rht2ht_householder(n, *[x for M in got for x in (M, n)], r, nq)
    return got


def _invariants(O, A0, B0, got):
    n = A0.shape[0]
    ea, eb = O.backward_errors(A0, B0, *got)
    assert ea <= 10 * n * U and eb <= 10 * n * U, (ea, eb)
    assert O.orthogonality(got[2]) <= 10 * n * U
    assert O.orthogonality(got[3]) <= 10 * n * U
    assert O.below_band(got[0], 1) == 0.0 and O.below_band(got[1], 0) == 0.0


# (n, r, nq): MFMA far-update pairs, generic ones (r or nq not a multiple of
# 4), nq > r + 1 (no helper workgroups), nq = 1, tails
CASES = [(40, 4, 4), (96, 8, 8), (150, 16, 8), (77, 5, 3), (130, 32, 16), (200, 64, 16),
         (129, 16, 1), (64, 2, 2), (61, 4, 6), (75, 4, 6), (300, 64, 32), (257, 16, 16),
         (180, 32, 8), (133, 8, 4), (211, 64, 8), (150, 12, 6), (90, 8, 12)]


def _reference(O, A0, B0, mats, r, nq):
    """The oracle's rht2ht_householder; where the reference's defect for
    nq > r strikes (a sweep's column n-1 skips the reflectors it owes later
    bulge blocks, rht2ht.f90:1465 with nb < 2, so the oracle's backward error
    is O(1)), the numpy restatement of the same sweep that applies them."""
    n = A0.shape[0]
    ref = [M.copy(order="F") for M in mats]
    O.rht2ht_householder(n, *ref, r, nq)
    ea, eb = O.backward_errors(A0, B0, *ref)
    if ea <= 10 * n * U and eb <= 10 * n * U:
        return ref
    assert nq > r  # the only known defect
    import hh_emul
    ref = [M.copy(order="F") for M in mats]
    hh_emul.reduce_sweeps(*ref, r, nq)
    return ref


@pytest.mark.parametrize("n,r,nq", CASES)
def test_householder_stage2_matches_oracle_elementwise(oracle_mod, n, r, nq):
    O = oracle_mod
    print(f"householder n={n} r={r} nq={nq}", flush=True)
    A0, B0, *mats = _stage1_pencil(O, n, r, 900 + n)
    got = _run_gpu(n, mats, r, nq)
    ref = _reference(O, A0, B0, mats, r, nq)
    scale = [np.linalg.norm(A0), np.linalg.norm(B0), n, n]
    for x, y, sc in zip(got, ref, scale):
        assert np.linalg.norm(x - y) / sc <= 1e-9
    _invariants(O, A0, B0, got)


@pytest.mark.parametrize("n,r,nq", [(700, 64, 16), (400, 16, 8), (333, 32, 32), (200, 8, 4)])
def test_householder_lookahead_modes_and_generic_far(oracle_mod, monkeypatch, n, r, nq):
    """The lookahead split on two streams (1) and on one (2) are bit-identical;
    the one-pass update (0) and the generic far-update kernels agree to
    rounding; all match the oracle."""
    O = oracle_mod
    print(f"householder modes n={n} r={r} nq={nq}", flush=True)
    A0, B0, *mats = _stage1_pencil(O, n, r, 950 + n)
    runs = []
    for flag in ("1", "2", "0"):
        monkeypatch.setenv("PARAHT_S2_LOOKAHEAD", flag)
        runs.append(_run_gpu(n, mats, r, nq))
    monkeypatch.setenv("PARAHT_S2_LOOKAHEAD", "1")
    monkeypatch.setenv("PARAHT_HH_GENERIC", "1")
    runs.append(_run_gpu(n, mats, r, nq))
    monkeypatch.delenv("PARAHT_HH_GENERIC")
    monkeypatch.delenv("PARAHT_S2_LOOKAHEAD")
    ref = _reference(O, A0, B0, mats, r, nq)
    scale = [np.linalg.norm(A0), np.linalg.norm(B0), n, n]
    for k in range(4):
        assert np.array_equal(runs[0][k], runs[1][k])
        assert np.linalg.norm(runs[0][k] - runs[2][k]) / scale[k] <= 1e-12
        assert np.linalg.norm(runs[0][k] - runs[3][k]) / scale[k] <= 1e-12
        assert np.linalg.norm(runs[0][k] - ref[k]) / scale[k] <= 1e-9


def test_householder_illconditioned_b(oracle_mod):
    """B with exact zeros on its diagonal (infinite eigenvalues): the block RQ
    factorizations are singular, so only the invariants are compared."""
    O = oracle_mod
    n, r, nq = 400, 16, 8
    A0, B0, *mats = _stage1_pencil(O, n, r, 77, illcond=True)
    got = _run_gpu(n, mats, r, nq)
    _invariants(O, A0, B0, got)


def test_householder_reduce_device_api_and_variants(oracle_mod):
    """paraht_reduce_dev without the Givens flag runs the Householder stage 2
    (the oracle's ab2rht_simple + rht2ht_householder, sign-normalised only for
    stage 1's signs, which it shares); with the flag the Givens chase, and
    both reach the same HT form up to signs."""
    import torch
    O = oracle_mod
    n, r, p, nq = 320, 16, 4, 8
    A0, B0 = random_pencil(n, seed=4242)
    dev = torch.device("cuda:0")
    cm = lambda M: ph.column_major(torch.from_numpy(np.ascontiguousarray(M)).to(dev))
    hh = [M.cpu().numpy() for M in ph.reduce(cm(A0), cm(B0), r, p, nq)]
    gv = [M.cpu().numpy() for M in ph.reduce(cm(A0), cm(B0), r, p, nq, givens=True)]
    ref = O.reduce(A0, B0, r, p, nq, stage2="householder")
    g, o, v = O.sign_normalize(*hh), O.sign_normalize(*ref), O.sign_normalize(*gv)
    scale = [np.linalg.norm(A0), np.linalg.norm(B0), n, n]
    for x, y, z, sc in zip(g, o, v, scale):
        assert np.linalg.norm(x - y) / sc <= 1e-9
        assert np.linalg.norm(x - z) / sc <= 1e-9
    for M in (hh, gv):
        _invariants(O, A0, B0, [np.asfortranarray(x) for x in M])


def test_householder_unsupported_pairs_refused():
    n = 50
    A = np.eye(n, order="F")
    mats = [A.copy(order="F") for _ in range(4)]
    with pytest.raises(ph.ParahtError):
        ph.rht2ht_householder(n, *[x for M in mats for x in (M, n)], 64, 40)  # r + nq > 96
    with pytest.raises(ph.ParahtError):
        ph.rht2ht_householder(n, *[x for M in mats for x in (M, n)], 16, 33)  # nq > 32
