"""The oracle's QR of a general B (oracle.qr_b: LAPACK DGEQRF, DORMQR) pinned
by what the factorisation fixes: Q_B orthogonal, Q_B R = B0, A = Q_B^T A0,
Q = Q0 Q_B, exact zeros below R's diagonal, and R's diagonal equal in
magnitude to that of the Cholesky factor of B0^T B0 (R^T R = B0^T B0 fixes
R up to row signs).  A transposed application to A or Q, or a dropped
triangle, fails one of them.  Also the host-side argument checks of qr_b."""
import numpy as np
import pytest

from conftest import U


@pytest.mark.parametrize("n", [0, 1, 2, 7, 50, 131])
def test_oracle_qr_b(oracle_mod, n):
    rng = np.random.default_rng(40 + n)
    A0 = rng.standard_normal((n, n))
    B0 = rng.standard_normal((n, n))
    Q0 = np.linalg.qr(rng.standard_normal((n, n)))[0] if n else np.zeros((0, 0))
    A, R, Q = oracle_mod.qr_b(A0, B0, Q0)
    if n == 0:
        return
    QB = Q0.T @ Q  # Q = Q0 Q_B
    tol = 10 * max(n, 1) * U
    assert np.linalg.norm(QB.T @ QB - np.eye(n)) <= tol
    assert np.linalg.norm(QB @ R - B0) <= tol * np.linalg.norm(B0)
    assert np.linalg.norm(QB.T @ A0 - A) <= tol * np.linalg.norm(A0)
    assert np.all(np.tril(R, -1) == 0.0)
    L = np.linalg.cholesky(B0.T @ B0)
    rel = np.abs(np.abs(np.diag(R)) - np.abs(np.diag(L))) / np.abs(np.diag(L))
    assert rel.max() <= 1e-8


def test_oracle_qr_b_planted_transpose_fails(oracle_mod):
    """The same checks catch Q_B applied untransposed to A."""
    n = 40
    rng = np.random.default_rng(5)
    A0, B0 = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    A, R, Q = oracle_mod.qr_b(A0, B0)
    wrong = Q @ A0
    assert np.linalg.norm(Q.T @ A0 - wrong) > 1e-3 * np.linalg.norm(A0)
    assert np.linalg.norm(Q.T @ A0 - A) <= 10 * n * U * np.linalg.norm(A0)


def test_qr_b_shape_checks():
    import paraht_amd as ph
    n = 6
    A = np.zeros((n, n), order="F")
    with pytest.raises(ValueError, match="smaller than n x n"):
        ph.qr_b(n, A, n, np.zeros((n, n - 1), order="F"), n, np.eye(n, order="F"), n)
    with pytest.raises(ValueError, match="illegal value of argument 5"):
        ph.qr_b(n, A, n, np.zeros((n, n), order="F"), n + 1, np.eye(n, order="F"), n)
    with pytest.raises(ValueError, match="expected"):
        ph.reduce_general(A, np.zeros((n, n + 1)), 4, 2, 2)
