"""Pins the CPU oracle (the C restatement of the reference) to what the
mathematics and independent library routines fix.  CPU only.

* LAPACK DGGHD3 (host LAPACK) produces the same Hessenberg-triangular form up
  to diagonal signs (generalized implicit-Q theorem: both keep Q e1 = e1).
* With B = I the reduction is the ordinary Hessenberg reduction
  (scipy.linalg.hessenberg, up to signs).
* Stage 1 spans block Krylov subspaces: span Q(:,1:k r) = K_k(A B^-1, E_r).
* Invariants: backward errors, orthogonality, exact structural zeros,
  generalized eigenvalues; special cases n <= 2 and an already reduced A.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from conftest import U
from paraht_amd.pencils import illcond_pencil, random_pencil

CASES = [  # n, r, p, nq  (nq <= r: the reference's working range)
    (40, 4, 3, 3), (60, 8, 4, 4), (100, 16, 8, 8), (57, 5, 2, 5),
    (33, 16, 8, 8), (64, 1, 2, 1), (50, 3, 2, 2), (73, 8, 2, 8),
]


@pytest.mark.parametrize("n,r,p,nq", CASES)
@pytest.mark.parametrize("stage2", ["givens", "householder", "unblocked"])
def test_oracle_matches_dgghd3(oracle_mod, n, r, p, nq, stage2):
    O = oracle_mod
    A0, B0 = random_pencil(n, seed=n + r)
    H, T, Q, Z = O.reduce(A0, B0, r, p, nq, stage2=stage2)
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    assert ea <= 10 * n * U and eb <= 10 * n * U
    assert O.orthogonality(Q) <= 10 * n * U and O.orthogonality(Z) <= 10 * n * U
    assert O.below_band(H, 1) == 0.0
    # the reference does not zero the last stage-1 block of B explicitly
    # (ab2rht.f90:335), so the Givens path may keep rounding-level entries
    assert O.below_band(T, 0) <= 10 * n * U * np.linalg.norm(B0)
    assert np.all(Q[0, 1:] == 0) and Q[0, 0] == 1.0  # Q e1 = e1
    Hl, Tl, Ql, Zl = O.dgghd3(A0, B0)
    Hn, Tn, Qn, Zn = O.sign_normalize(H, T, Q, Z)
    Hm, Tm, Qm, Zm = O.sign_normalize(Hl, Tl, Ql, Zl)
    sa, sb = np.linalg.norm(A0), np.linalg.norm(B0)
    assert np.linalg.norm(Hn - Hm) <= 1e-9 * sa
    assert np.linalg.norm(Tn - Tm) <= 1e-9 * sb
    assert np.linalg.norm(Qn - Qm) <= 1e-9 * n
    assert np.linalg.norm(Zn - Zm) <= 1e-9 * n


def test_plain_mistake_is_caught(oracle_mod):
    """A dropped sign in Q (a plausible bug) must fail the DGGHD3 pin."""
    O = oracle_mod
    A0, B0 = random_pencil(30, seed=3)
    H, T, Q, Z = O.reduce(A0, B0, 4, 3, 3)
    Qbad = Q.copy()
    Qbad[:, 5] *= -1.0  # not a valid sign change: H, T not adjusted
    ea, _ = O.backward_errors(A0, B0, H, T, Qbad, Z)
    assert ea > 1e-3


def test_b_identity_is_hessenberg_reduction(oracle_mod):
    O = oracle_mod
    n = 48
    A0, _ = random_pencil(n, seed=11)
    B0 = np.eye(n, order="F")
    for st in ("givens", "householder"):
        H, T, Q, Z = O.reduce(A0, B0, 6, 3, 4, stage2=st)
        Hn, Tn, Qn, Zn = O.sign_normalize(H, T, Q, Z)
        assert np.allclose(Tn, np.eye(n), atol=1e-12)
        Hs, Qs = sl.hessenberg(A0, calc_q=True)
        Hsn, _, Qsn, _ = O.sign_normalize(Hs, np.eye(n), Qs, Qs)
        assert np.linalg.norm(Hn - Hsn) <= 1e-10 * np.linalg.norm(A0)
        assert np.linalg.norm(Qn - Qsn) <= 1e-10 * n


def test_stage2_leaves_reduced_pencil_alone(oracle_mod):
    O = oracle_mod
    n = 30
    A0, B0 = random_pencil(n, seed=5)
    H0 = np.asfortranarray(np.triu(A0, -1))
    for fn in (O.rht2ht_givens, O.rht2ht_householder):
        A, B = H0.copy(order="F"), B0.copy(order="F")
        Q, Z = np.eye(n, order="F"), np.eye(n, order="F")
        fn(n, A, B, Q, Z, 4, 3)
        assert np.array_equal(A, H0) and np.array_equal(B, B0)
        assert np.array_equal(Q, np.eye(n)) and np.array_equal(Z, np.eye(n))


@pytest.mark.parametrize("n", [0, 1, 2])
def test_tiny_pencils_unchanged(oracle_mod, n):
    O = oracle_mod
    A0, B0 = random_pencil(max(n, 1), seed=1)
    ld = max(1, n)
    A = np.zeros((ld, n), order="F")
    B = np.zeros((ld, n), order="F")
    A[:n, :n], B[:n, :n] = A0[:n, :n], B0[:n, :n]
    Q = np.eye(ld, n, order="F")
    Z = np.eye(ld, n, order="F")
    A1, B1 = A.copy(order="F"), B.copy(order="F")
    O.ab2rht_simple(n, A, B, Q, Z, 2, 2)
    O.rht2ht_givens(n, A, B, Q, Z, 2, 2)
    O.rht2ht_householder(n, A, B, Q, Z, 2, 2)
    assert np.array_equal(A, A1) and np.array_equal(B, B1)


@pytest.mark.parametrize("n,r,p", [(24, 4, 3), (30, 3, 2), (32, 4, 8)])
def test_stage1_spans_block_krylov(oracle_mod, n, r, p):
    O = oracle_mod
    A0, B0 = random_pencil(n, seed=2 * n + r)
    A, B = A0.copy(order="F"), B0.copy(order="F")
    Q, Z = np.eye(n, order="F"), np.eye(n, order="F")
    O.ab2rht_simple(n, A, B, Q, Z, r, p)
    # structure of the r-Hessenberg-triangular form
    assert O.below_band(A, r) == 0.0
    assert O.below_band(B, 0) <= 10 * n * U * np.linalg.norm(B0)
    assert np.array_equal(Q[:, :r], np.eye(n)[:, :r])
    ea, eb = O.backward_errors(A0, B0, A, B, Q, Z)
    assert ea <= 10 * n * U and eb <= 10 * n * U
    # block Arnoldi basis of K_k(A B^-1, E_r), compared with span Q(:, :k r)
    M = A0 @ np.linalg.inv(B0)
    V = np.eye(n)[:, :r]
    basis = V
    for k in range(1, n // r):
        dist = np.linalg.norm(Q[:, k * r:].T @ basis)
        assert dist <= 1e-8, (k, dist)
        W = M @ V
        for _ in range(2):
            W = W - basis @ (basis.T @ W)
        V, _ = np.linalg.qr(W)
        basis = np.hstack([basis, V])


def test_generalized_eigenvalues_preserved(oracle_mod):
    O = oracle_mod
    A0, B0 = random_pencil(80, seed=9)
    H, T, _, _ = O.reduce(A0, B0, 8, 4, 4)
    assert O.chordal_eig_distance(A0, B0, H, T) <= 1e-8


def test_illconditioned_b(oracle_mod):
    O = oracle_mod
    n = 120
    A0, B0 = illcond_pencil(n, seed=4)
    assert np.sum(np.diag(B0) == 0.0) == 6
    for st in ("givens", "householder"):
        H, T, Q, Z = O.reduce(A0, B0, 8, 4, 8, stage2=st)
        ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
        assert ea <= 10 * n * U and eb <= 10 * n * U
        assert O.below_band(H, 1) == 0.0
        assert O.orthogonality(Q) <= 10 * n * U


def test_reference_defect_for_nq_above_r(oracle_mod):
    """The reference's blocked stage 2 skips, for the last column (nb < 2,
    rht2ht.f90:827/1484 'cycle'), the lazily-applied left rotations of the
    second bulge block on column n; with nq > r the last sweep can own such a
    block and A loses that update.  The oracle reproduces the defect; the GPU
    path does not have it (it applies left rotations eagerly)."""
    O = oracle_mod
    n, r, nq = 57, 5, 7
    A0, B0 = random_pencil(n, seed=7)
    for st in ("givens", "householder"):
        H, T, Q, Z = O.reduce(A0, B0, r, 2, nq, stage2=st)
        ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
        assert ea > 1e-6 and eb <= 10 * n * U
        E = Q.T @ A0 @ Z - H
        bad = np.argwhere(np.abs(E) > 1e-10)
        assert set(bad[:, 1]) == {n - 1}


def test_golden_fixtures(oracle_mod):
    import glob
    import os
    O = oracle_mod
    files = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
    assert files, "golden fixtures missing"
    for f in files:
        d = np.load(f)
        n, r, p, nq = (int(x) for x in d["params"])
        A0, B0 = d["A"], d["B"]
        H, T, Q, Z = O.reduce(A0, B0, r, p, nq, stage2="householder")
        ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
        assert ea <= 10 * n * U and eb <= 10 * n * U, f
        if np.min(np.abs(np.diag(d["T"]))) < 1e-8 * np.linalg.norm(B0):
            # singular T: the form is not unique, only the invariants pin it
            continue
        Hn, Tn, Qn, Zn = O.sign_normalize(H, T, Q, Z)
        assert np.linalg.norm(Hn - d["H"]) <= 1e-9 * np.linalg.norm(A0), f
        assert np.linalg.norm(Tn - d["T"]) <= 1e-9 * np.linalg.norm(B0), f
        assert np.linalg.norm(Qn - d["Q"]) <= 1e-9 * n, f
        assert np.linalg.norm(Zn - d["Z"]) <= 1e-9 * n, f
