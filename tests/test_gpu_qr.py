"""GPU tests of the QR of a general B (paraht_qr_b, paraht_qr_b_dev through
the C-ABI) and of the reduction of a general pencil (reduce_general).

* against the oracle's LAPACK QR (oracle.qr_b): R, Q_B^T A and Q Q_B agree
  after the row signs of R are matched (the factorisation is unique up to
  them), to 1e-9 relative (Frobenius; forward error of a QR of an N(0,1)
  matrix, condition number ~ n);
* invariants: ||Q^T B0 - R||_F / ||B0||_F, ||Q^T A0 - A||_F / ||A0||_F and
  ||Q^T Q - I||_F <= 10 n u; exact zeros below R's diagonal; rows past n of
  arrays with ld > n untouched;
* the general-pencil reduction: backward errors of H and T against the
  input pencil and orthogonality <= 10 n u, exact zeros below the structure,
  and the generalized eigenvalues of (H, T) those of (A0, B0) to 1e-6 in
  the chordal metric.
"""
import numpy as np
import pytest

import paraht_amd as ph
from conftest import U

pytestmark = pytest.mark.gpu

# (n, nb, p): nb <= 64 and p nb <= 128 take the register-resident generators,
# (64, 4) the generic ones; n < nb, n = nb, n = nb + 1 and ragged last panels
CASES = [(2, 64, 2), (5, 64, 2), (64, 64, 2), (65, 64, 2), (130, 32, 4), (257, 64, 2),
         (300, 16, 8), (777, 48, 2), (1000, 64, 2), (400, 64, 4)]


def _signs(Rg, Ro):
    d = np.sign(np.diag(Rg)) * np.sign(np.diag(Ro))
    d[d == 0] = 1.0
    return d


@pytest.mark.parametrize("n,nb,p", CASES)
def test_qr_b_matches_oracle(oracle_mod, n, nb, p):
    print(f"qr_b n={n} nb={nb} p={p}", flush=True)
    rng = np.random.default_rng(700 + n + nb + p)
    A0 = rng.standard_normal((n, n))
    B0 = rng.standard_normal((n, n))
    ld = n + 3  # leading dimension past n: the extra rows must stay as they are
    A = np.asfortranarray(rng.standard_normal((ld, n)))
    B = np.asfortranarray(rng.standard_normal((ld, n)))
    Q = np.asfortranarray(rng.standard_normal((ld, n)))
    pad = (A[n:].copy(), B[n:].copy(), Q[n:].copy())
    A[:n], B[:n], Q[:n] = A0, B0, np.eye(n)
    ph.qr_b(n, A, ld, B, ld, Q, ld, nb, p)
    assert np.array_equal(A[n:], pad[0]) and np.array_equal(B[n:], pad[1])
    assert np.array_equal(Q[n:], pad[2])
    Ag, Rg, Qg = A[:n], B[:n], Q[:n]
    tol = 10 * n * U
    assert np.all(np.tril(Rg, -1) == 0.0)
    assert np.linalg.norm(Qg.T @ Qg - np.eye(n)) <= tol
    assert np.linalg.norm(Qg.T @ B0 - Rg) <= tol * np.linalg.norm(B0)
    assert np.linalg.norm(Qg.T @ A0 - Ag) <= tol * np.linalg.norm(A0)
    Ao, Ro, Qo = oracle_mod.qr_b(A0, B0)
    d = _signs(Rg, Ro)
    assert np.linalg.norm(d[:, None] * Rg - Ro) <= 1e-9 * np.linalg.norm(Ro)
    assert np.linalg.norm(d[:, None] * Ag - Ao) <= 1e-9 * np.linalg.norm(Ao)
    assert np.linalg.norm(Qg * d[None, :] - Qo) <= 1e-9 * np.sqrt(n)


def test_qr_b_device_carries_q():
    """paraht_qr_b_dev on torch tensors with a non-identity Q: Q <- Q0 Q_B,
    and a triangular B stays triangular with |R| = |B| row for row."""
    import torch
    n, dev = 513, torch.device("cuda:0")
    print(f"qr_b_dev n={n}", flush=True)
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    A0 = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    B0 = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    Q0 = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))[0]
    A, B, Q = (ph.column_major(M.clone()) for M in (A0, B0, Q0))
    ph.qr_b(n, A, None, B, None, Q, None, 64, 2)
    torch.cuda.synchronize()
    QB = Q0.T @ Q
    I = torch.eye(n, dtype=torch.float64, device=dev)
    tol = 10 * n * U
    assert torch.count_nonzero(torch.tril(B, -1)).item() == 0
    assert torch.linalg.norm(QB.T @ QB - I).item() <= tol
    assert (torch.linalg.norm(QB.T @ B0 - B) / torch.linalg.norm(B0)).item() <= tol
    assert (torch.linalg.norm(QB.T @ A0 - A) / torch.linalg.norm(A0)).item() <= tol
    # an upper triangular B: Q_B is diagonal (+-1) and R = Q_B^T B
    T0 = ph.column_major(torch.triu(B0))
    T, Aw, Qw = T0.clone(), ph.column_major(A0.clone()), ph.column_major(I.clone())
    ph.qr_b(n, Aw, None, T, None, Qw, None, 32, 4)
    torch.cuda.synchronize()
    assert torch.allclose(T.abs(), T0.abs(), rtol=1e-12, atol=1e-12)
    assert torch.allclose(Qw.abs(), I, atol=1e-12)


@pytest.mark.parametrize("n,r,p,nq", [(200, 8, 4, 4), (1000, 64, 2, 16)])
def test_reduce_general_pencil(oracle_mod, n, r, p, nq):
    O = oracle_mod
    print(f"reduce_general n={n} r={r} p={p} nq={nq}", flush=True)
    rng = np.random.default_rng(31 + n)
    A0 = rng.standard_normal((n, n))
    B0 = rng.standard_normal((n, n))
    H, T, Q, Z = ph.reduce_general(A0, B0, r, p, nq)
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    tol = 10 * n * U
    assert ea <= tol and eb <= tol, (ea, eb)
    assert O.orthogonality(Q) <= tol and O.orthogonality(Z) <= tol
    assert O.below_band(H, 1) == 0.0 and O.below_band(T, 0) == 0.0
    if n <= 200:
        assert O.chordal_eig_distance(A0, B0, H, T) <= 1e-6
