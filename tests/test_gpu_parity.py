"""GPU parity: libparaht_hip.so (through the C-ABI) against the CPU oracle on
the same seeded pencils, plus the invariants at sizes the oracle cannot reach.

Tolerances (FP64, u = 2^-53):
* backward errors and orthogonality <= 10 n u;
* structural zeros are exact (== 0.0) on the GPU path;
* stage-1 and full reductions are compared with the oracle after sign
  normalisation (the forms are unique up to diagonal signs) to
  1e-9 relative (Frobenius);
* the Givens stage 2 applies the oracle's rotation sequence, so it is compared
  elementwise without normalisation to 1e-9 relative.
"""
import numpy as np
import pytest

import paraht_amd as ph
from conftest import U
from paraht_amd.pencils import illcond_pencil, random_pencil

pytestmark = pytest.mark.gpu


def _fresh(A0, B0):
    n = A0.shape[0]
    return (A0.copy(order="F"), B0.copy(order="F"), np.eye(n, order="F"),
            np.eye(n, order="F"))


def _rel(X, Y, scale):
    return np.linalg.norm(X - Y) / scale


def _check_invariants(O, A0, B0, H, T, Q, Z, hband=1, exact=True):
    n = A0.shape[0]
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    assert ea <= 10 * n * U and eb <= 10 * n * U, (ea, eb)
    assert O.orthogonality(Q) <= 10 * n * U
    assert O.orthogonality(Z) <= 10 * n * U
    if exact:
        assert O.below_band(H, hband) == 0.0
        assert O.below_band(T, 0) == 0.0


STAGE1 = [(40, 4, 3), (96, 8, 2), (150, 16, 8), (77, 5, 4), (130, 32, 2),
          (200, 64, 2), (129, 16, 3)]


@pytest.mark.parametrize("n,r,p", STAGE1)
def test_stage1_matches_oracle(oracle_mod, n, r, p):
    O = oracle_mod
    print(f"stage1 n={n} r={r} p={p}", flush=True)
    A0, B0 = random_pencil(n, seed=100 + n)
    A, B, Q, Z = _fresh(A0, B0)
    ph.ab2rht_simple(n, A, n, B, n, Q, n, Z, n, r, p)
    _check_invariants(O, A0, B0, A, B, Q, Z, hband=r)
    assert np.array_equal(Q[:, :r], np.eye(n)[:, :r])
    Ao, Bo, Qo, Zo = _fresh(A0, B0)
    O.ab2rht_simple(n, Ao, Bo, Qo, Zo, r, p)
    g = O.sign_normalize(A, B, Q, Z, band=r)
    o = O.sign_normalize(Ao, Bo, Qo, Zo, band=r)
    sa, sb = np.linalg.norm(A0), np.linalg.norm(B0)
    assert _rel(g[0], o[0], sa) <= 1e-9
    assert _rel(g[1], o[1], sb) <= 1e-9
    assert _rel(g[2], o[2], n) <= 1e-9
    assert _rel(g[3], o[3], n) <= 1e-9


STAGE2 = [(40, 4, 3), (96, 8, 8), (150, 16, 8), (77, 5, 4), (130, 32, 16),
          (200, 64, 16), (129, 16, 1), (64, 2, 2)]


@pytest.mark.parametrize("n,r,nq", STAGE2)
def test_stage2_givens_matches_oracle(oracle_mod, n, r, nq):
    O = oracle_mod
    print(f"stage2 n={n} r={r} nq={nq}", flush=True)
    A0, B0 = random_pencil(n, seed=200 + n)
    A1, B1, Q1, Z1 = _fresh(A0, B0)
    O.ab2rht_simple(n, A1, B1, Q1, Z1, r, 2)
    B1 = np.asfortranarray(np.triu(B1))  # exact triangle as the GPU stage 1 leaves it
    A, B, Q, Z = A1.copy(order="F"), B1.copy(order="F"), Q1.copy(order="F"), Z1.copy(order="F")
    ph.rht2ht_givens(n, A, n, B, n, Q, n, Z, n, r, nq)
    Ao, Bo, Qo, Zo = A1.copy(order="F"), B1.copy(order="F"), Q1.copy(order="F"), Z1.copy(order="F")
    O.rht2ht_givens(n, Ao, Bo, Qo, Zo, r, nq)
    sa, sb = np.linalg.norm(A0), np.linalg.norm(B0)
    assert _rel(A, Ao, sa) <= 1e-9
    assert _rel(B, Bo, sb) <= 1e-9
    assert _rel(Q, Qo, n) <= 1e-9
    assert _rel(Z, Zo, n) <= 1e-9
    _check_invariants(O, A0, B0, A, B, Q, Z)


FULL = [(100, 16, 8, 8), (64, 8, 4, 4), (257, 16, 8, 8), (300, 64, 2, 16),
        (33, 16, 8, 8)]


@pytest.mark.parametrize("n,r,p,nq", FULL)
def test_full_reduction_matches_oracle_and_dgghd3(oracle_mod, n, r, p, nq):
    O = oracle_mod
    print(f"full n={n} r={r} p={p} nq={nq}", flush=True)
    A0, B0 = random_pencil(n, seed=300 + n)
    H, T, Q, Z = ph.reduce(A0, B0, r, p, nq)
    _check_invariants(O, A0, B0, H, T, Q, Z)
    assert np.all(Q[0, 1:] == 0) and Q[0, 0] == 1.0
    g = O.sign_normalize(H, T, Q, Z)
    for ref in (O.reduce(A0, B0, r, p, nq, stage2="givens"), O.dgghd3(A0, B0)):
        o = O.sign_normalize(*ref)
        assert _rel(g[0], o[0], np.linalg.norm(A0)) <= 1e-9
        assert _rel(g[1], o[1], np.linalg.norm(B0)) <= 1e-9
        assert _rel(g[2], o[2], n) <= 1e-9
        assert _rel(g[3], o[3], n) <= 1e-9
    if n <= 120:
        assert O.chordal_eig_distance(A0, B0, H, T) <= 1e-8


def test_golden_fixtures_on_gpu(oracle_mod):
    import glob
    import os
    O = oracle_mod
    files = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
    assert files
    for f in files:
        d = np.load(f)
        n, r, p, nq = (int(x) for x in d["params"])
        print("golden", os.path.basename(f), flush=True)
        H, T, Q, Z = ph.reduce(d["A"], d["B"], r, p, nq)
        _check_invariants(O, d["A"], d["B"], H, T, Q, Z)
        if np.min(np.abs(np.diag(d["T"]))) < 1e-8 * np.linalg.norm(d["B"]):
            continue
        g = O.sign_normalize(H, T, Q, Z)
        assert _rel(g[0], d["H"], np.linalg.norm(d["A"])) <= 1e-9
        assert _rel(g[1], d["T"], np.linalg.norm(d["B"])) <= 1e-9
        assert _rel(g[2], d["Q"], n) <= 1e-9
        assert _rel(g[3], d["Z"], n) <= 1e-9


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_tiny_sizes(oracle_mod, n):
    O = oracle_mod
    A0, B0 = random_pencil(n, seed=n)
    H, T, Q, Z = ph.reduce(A0, B0, 2, 2, 2)
    _check_invariants(O, A0, B0, H, T, Q, Z)
    if n <= 2:
        assert np.array_equal(H, A0) and np.array_equal(T, B0)


def test_r1_nq1_and_nq_above_r(oracle_mod):
    """r = 1 (stage 1 alone is a Hessenberg reduction), nq = 1, and nq > r
    (where the reference's blocked stage 2 loses an update; the GPU path
    stays correct)."""
    O = oracle_mod
    for (n, r, p, nq) in [(50, 1, 2, 1), (57, 5, 2, 7), (80, 8, 2, 16), (60, 4, 3, 1)]:
        print("edge", n, r, p, nq, flush=True)
        A0, B0 = random_pencil(n, seed=7)
        H, T, Q, Z = ph.reduce(A0, B0, r, p, nq)
        _check_invariants(O, A0, B0, H, T, Q, Z)
        g = O.sign_normalize(H, T, Q, Z)
        o = O.sign_normalize(*O.dgghd3(A0, B0))
        assert _rel(g[0], o[0], np.linalg.norm(A0)) <= 1e-9


def test_leading_dimension_larger_than_n(oracle_mod):
    O = oracle_mod
    n, ld = 70, 83
    A0, B0 = random_pencil(n, seed=12)
    big = lambda M: np.asfortranarray(np.vstack([M, np.full((ld - n, n), 7.0)]))
    A, B = big(A0), big(B0)
    Q, Z = big(np.eye(n)), big(np.eye(n))
    ph.ab2rht_simple(n, A, ld, B, ld, Q, ld, Z, ld, 8, 3)
    ph.rht2ht_givens(n, A, ld, B, ld, Q, ld, Z, ld, 8, 4)
    for M in (A, B, Q, Z):
        assert np.all(M[n:, :] == 7.0)  # padding rows untouched
    _check_invariants(O, A0, B0, A[:n], B[:n], Q[:n], Z[:n])


def test_already_hessenberg_is_left_alone():
    n = 60
    A0, B0 = random_pencil(n, seed=5)
    H0 = np.asfortranarray(np.triu(A0, -1))
    A, B, Q, Z = _fresh(H0, B0)
    ph.rht2ht_givens(n, A, n, B, n, Q, n, Z, n, 8, 4)
    assert np.array_equal(A, H0) and np.array_equal(B, B0)
    assert np.array_equal(Q, np.eye(n)) and np.array_equal(Z, np.eye(n))


def test_illconditioned_pencil(oracle_mod):
    O = oracle_mod
    n = 400
    A0, B0 = illcond_pencil(n, seed=4)
    H, T, Q, Z = ph.reduce(A0, B0, 16, 4, 8)
    _check_invariants(O, A0, B0, H, T, Q, Z)


def test_device_pointer_api_in_place(oracle_mod):
    import torch
    O = oracle_mod
    n, r, p, nq = 180, 16, 4, 8
    A0, B0 = random_pencil(n, seed=21)
    dev = torch.device("cuda:0")
    cm = lambda M: ph.column_major(torch.from_numpy(np.ascontiguousarray(M)).to(dev))
    A, B = cm(A0), cm(B0)
    H, T, Q, Z = ph.reduce(A, B, r, p, nq)
    assert H.data_ptr() == A.data_ptr()
    h = [M.cpu().numpy().copy(order="F") for M in (H, T, Q, Z)]
    _check_invariants(O, A0, B0, *h)
    g = O.sign_normalize(*h)
    o = O.sign_normalize(*ph.reduce(A0, B0, r, p, nq))
    assert _rel(g[0], o[0], np.linalg.norm(A0)) <= 1e-12


def test_config0_invariants_n1000():
    """The reference's README workload shape (n=1000, r=16, p=8, nq=8), checked
    with the invariants on the GPU (torch matmuls in FP64)."""
    import torch
    n = 1000
    A0, B0 = random_pencil(n)
    H, T, Q, Z = ph.reduce(A0, B0, 16, 8, 8)
    d = torch.device("cuda:0")
    t = lambda M: torch.from_numpy(np.ascontiguousarray(M)).to(d)
    A0t, B0t, Ht, Tt, Qt, Zt = map(t, (A0, B0, H, T, Q, Z))
    ea = torch.linalg.norm(Qt.T @ A0t @ Zt - Ht) / torch.linalg.norm(A0t)
    eb = torch.linalg.norm(Qt.T @ B0t @ Zt - Tt) / torch.linalg.norm(B0t)
    I = torch.eye(n, dtype=torch.float64, device=d)
    assert ea.item() <= 10 * n * U and eb.item() <= 10 * n * U
    assert torch.linalg.norm(Qt.T @ Qt - I).item() <= 10 * n * U
    assert torch.linalg.norm(Zt.T @ Zt - I).item() <= 10 * n * U
    assert torch.count_nonzero(torch.tril(Ht, -2)).item() == 0
    assert torch.count_nonzero(torch.tril(Tt, -1)).item() == 0
