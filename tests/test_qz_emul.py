"""CPU checks of the QZ iteration's algorithm (tests/qz_emul.py, the numpy
statement of paraht_amd/csrc/qz.hip) before it runs on the GPU:
  * eigenvalues agree with LAPACK's (scipy.linalg.eigvals, i.e. DGGEV) in the
    chordal metric, infinite eigenvalues included;
  * known answers: pencils built with prescribed eigenvalues (real, complex
    pairs, infinite) from a quasi-triangular pair and random orthogonal
    factors; B = I reduces to the standard eigenproblem (numpy's DHSEQR);
  * the Schur form: S quasi-upper-triangular (2x2 blocks only for complex
    pairs, their T block diagonal and positive), P upper triangular, both with
    exact zeros, Q^T A Z = S, Q^T B Z = P within 10 n u;
  * a planted mistake (a right reflector that skips the bulge's fill row)
    fails them.
The reference has no QZ step (its driver stops at the HT form,
src/profilereduction.f90:152-162), so LAPACK is the checker.
"""
import os
import sys

import numpy as np
import pytest
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))

import qz_emul as E  # noqa: E402
from paraht_amd.pencils import illcond_pencil, random_pencil  # noqa: E402

U = 2.0 ** -53


def ht_pencil(O, A0, B0, r=4, p=2, nq=2):
    H, T, Q, Z = O.reduce(A0, B0, r, p, nq)
    return np.array(H), np.triu(T), np.array(Q), np.array(Z)


def check_schur(A0, B0, S, P, Q, Z, ar, ai, be):
    n = A0.shape[0]
    ea = np.linalg.norm(Q.T @ A0 @ Z - S) / np.linalg.norm(A0)
    eb = np.linalg.norm(Q.T @ B0 @ Z - P) / np.linalg.norm(B0)
    assert ea <= 10 * n * U and eb <= 10 * n * U, (ea, eb)
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 10 * n * U
    assert np.abs(Z.T @ Z - np.eye(n)).max() <= 10 * n * U
    assert not np.tril(P, -1).any() and not np.tril(S, -2).any()
    j = 0
    while j < n:
        if j + 1 < n and S[j + 1, j] != 0.0:
            # a complex pair: 2x2 block, T's block diagonal and positive
            assert ai[j] > 0.0 and ai[j + 1] < 0.0
            lam = (ar[j] + 1j * ai[j]) / be[j]
            assert abs((ar[j + 1] + 1j * ai[j + 1]) / be[j + 1] - lam.conjugate()) <= 1e-12 * abs(lam)
            assert P[j, j + 1] == 0.0 and P[j, j] > 0.0 and P[j + 1, j + 1] > 0.0
            assert j + 2 >= n or S[j + 2, j + 1] == 0.0
            j += 2
        else:
            assert ai[j] == 0.0 and be[j] >= 0.0
            assert ar[j] == S[j, j] and be[j] == P[j, j]
            j += 1


@pytest.mark.parametrize("gen", [random_pencil, illcond_pencil])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 40, 120])
@pytest.mark.parametrize("schur", [False, True])
def test_eigenvalues_match_lapack(oracle_mod, gen, n, schur):
    A0, B0 = gen(n, seed=100 + n)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0)
    ref = sl.eigvals(A0, B0, homogeneous_eigvals=True)
    ar, ai, be, info = E.hgeqz(H, T, Q, Z, schur=schur)
    assert info == 0
    assert E.match_chordal(ar + 1j * ai, be, ref[0], ref[1]) <= 1e-11
    if schur:
        check_schur(A0, B0, H, T, Q, Z, ar, ai, be)


def _prescribed(n, seed, n_inf=0):
    """A pencil with known eigenvalues: (U S0 V^T, U P0 V^T) for a
    quasi-triangular S0 (real 1x1 blocks, 2x2 blocks [[a, b], [-b, a]] for
    a +- ib with P0's block = I), P0 upper triangular with n_inf zero rows
    and columns (semisimple infinite eigenvalues); returned with B made
    triangular by a QR."""
    rng = np.random.default_rng(seed)
    # strictly upper parts of norm O(1): random triangular matrices with O(1)
    # entries have exponentially ill-conditioned eigenvalues
    off = 0.5 / np.sqrt(n)
    S0 = np.triu(rng.standard_normal((n, n)), 1) * off
    P0 = np.triu(rng.standard_normal((n, n)), 1) * off
    alpha, beta = [], []
    j = 0
    while j < n:
        if j + 1 < n and j % 3 == 0:
            a, b = rng.uniform(-2, 2), rng.uniform(0.5, 2)
            S0[j:j + 2, j:j + 2] = [[a, b], [-b, a]]
            P0[j:j + 2, j:j + 2] = np.eye(2)
            alpha += [a + 1j * b, a - 1j * b]
            beta += [1.0, 1.0]
            j += 2
        else:
            d = rng.uniform(0.5, 2)
            P0[j, j] = d
            S0[j, j] = rng.uniform(-2, 2)
            alpha.append(S0[j, j])
            beta.append(d)
            j += 1
    inf_at = rng.choice([j for j in range(n) if j % 3 != 0 and j % 3 != 1 or j == n - 1],
                        size=n_inf, replace=False) if n_inf else []
    for j in inf_at:
        # a zero row and column keep the infinite eigenvalue semisimple (a
        # zero diagonal entry alone can make Jordan chains at infinity,
        # whose eigenvalues rounding moves by u^(1/k))
        P0[j, :] = 0.0
        P0[:, j] = 0.0
        S0[j, j] = rng.uniform(0.5, 2)
        alpha[j], beta[j] = S0[j, j], 0.0
    Uq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Vq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A, B = Uq @ S0 @ Vq.T, Uq @ P0 @ Vq.T
    Qb, R = np.linalg.qr(B)
    return Qb.T @ A, R, np.array(alpha), np.array(beta)


@pytest.mark.parametrize("n,n_inf", [(12, 0), (31, 0), (30, 2), (60, 3), (90, 5)])
def test_prescribed_eigenvalues(oracle_mod, n, n_inf):
    A0, B0, alpha, beta = _prescribed(n, seed=n + n_inf, n_inf=n_inf)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0)
    for schur in (False, True):
        h, t = H.copy(), T.copy()
        ar, ai, be, info = E.hgeqz(h, t, Q.copy(), Z.copy(), schur=schur)
        assert info == 0
        # infinite eigenvalues: beta = 0 once deflated as such, otherwise the
        # rounding of the reduction leaves a |beta| of order u ||B||
        assert int((np.abs(be) <= 1e-10 * np.hypot(ar, ai)).sum()) == n_inf
        assert E.match_chordal(ar + 1j * ai, be, alpha, beta) <= 1e-12


def test_identity_b_is_the_standard_problem():
    n = 50
    rng = np.random.default_rng(7)
    H = np.triu(rng.standard_normal((n, n)), -1)
    ar, ai, be, info = E.hgeqz(H.copy(), np.eye(n))
    assert info == 0
    ref = np.linalg.eigvals(H)
    assert E.match_chordal(ar + 1j * ai, be, ref, np.ones(n)) <= 1e-12


def test_planted_fill_row_skip_fails(oracle_mod, monkeypatch):
    n = 40
    A0, B0 = random_pencil(n, seed=9)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0)
    orig = E.QZ.right

    def skip_fill(self, c0, v, tau, lo, m, hr1, tr1):
        return orig(self, c0, v, tau, lo, m, min(hr1, c0 + 2), tr1)

    monkeypatch.setattr(E.QZ, "right", skip_fill)
    ar, ai, be, info = E.hgeqz(H.copy(), T.copy(), Q.copy(), Z.copy(), schur=True)
    ref = sl.eigvals(A0, B0, homogeneous_eigvals=True)
    assert info != 0 or E.match_chordal(ar + 1j * ai, be, ref[0], ref[1]) > 1e-6


@pytest.mark.parametrize("n", [6, 45, 130])
def test_lapack_dhgeqz_agrees(oracle_mod, n):
    """The LAPACK routine the GPU iteration replaces (DHGEQZ, through the
    oracle's binding) gives the same eigenvalues as the numpy statement."""
    A0, B0 = illcond_pencil(n, seed=n)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0)
    *_, lar, lai, lbe, linfo = oracle_mod.dhgeqz(H, T)
    ar, ai, be, info = E.hgeqz(H.copy(), T.copy())
    assert linfo == 0 and info == 0
    assert E.match_chordal(ar + 1j * ai, be, lar + 1j * lai, lbe) <= 1e-12


def _edge_pencils(n, rng):
    """Edge cases of the QZ iteration with known eigenvalues (alpha, beta):
    B = 0 (all infinite), H = 0 (all zero), an already triangular pencil
    (the diagonal pairs), and a block-diagonal pencil with a zero
    subdiagonal in the middle (two independent halves)."""
    H = np.triu(rng.standard_normal((n, n)), -1)
    T = np.triu(rng.standard_normal((n, n))) + 3 * np.eye(n)
    cases = [("B=0", H.copy(), np.zeros((n, n)), None),
             ("H=0", np.zeros((n, n)), T.copy(), None)]
    Ht = np.triu(H)
    cases.append(("triangular", Ht, T.copy(), (np.diag(Ht).astype(complex), np.diag(T))))
    Hs = H.copy()
    Hs[n // 2, n // 2 - 1] = 0.0
    cases.append(("split", Hs, T.copy(), None))
    return cases


@pytest.mark.parametrize("n", [1, 2, 7, 24])
def test_edge_pencils(n):
    rng = np.random.default_rng(n)
    for name, H, T, known in _edge_pencils(n, rng):
        for schur in (False, True):
            ar, ai, be, info = E.hgeqz(H.copy(), T.copy(), np.eye(n), np.eye(n), schur=schur)
            assert info == 0, name
            if name == "B=0":
                assert (be == 0).all(), name
            elif name == "H=0":
                assert (ar == 0).all() and (ai == 0).all() and (be > 0).all(), name
            else:
                if known is None:
                    known = sl.eigvals(H, T, homogeneous_eigvals=True)
                assert E.match_chordal(ar + 1j * ai, be, known[0], known[1]) <= 1e-12, name
