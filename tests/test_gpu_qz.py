"""GPU QZ iteration (paraht_hgeqz*, paraht_amd/csrc/qz.hip) against LAPACK
(scipy.linalg.eigvals, i.e. DGGEV, and DHGEQZ itself through the oracle's
binding) and the numpy statement of the same algorithm (tests/qz_emul.py):
  * eigenvalues in the chordal metric within 1e-11 (LAPACK and the numpy
    statement agree with each other to ~1e-14 on these pencils), infinite
    eigenvalues of singular B included;
  * prescribed eigenvalues (known answers);
  * the generalized real Schur form with Q, Z: exact structural zeros,
    2x2 blocks only for complex pairs with T's block diagonal and positive,
    Q^T A Z = S and Q^T B Z = P within 10 n u, orthogonal Q, Z;
  * many pencils in one launch, host and device entry points, the end-to-end
    reduce + QZ against the eigenvalues of the original pencil.
The reference has no QZ step (its driver stops at the HT form,
src/profilereduction.f90:152-162).
"""
import numpy as np
import pytest
import scipy.linalg as sl

import paraht_amd as ph
import qz_emul as E
from paraht_amd.pencils import illcond_pencil, random_pencil
from test_qz_emul import _prescribed, check_schur, ht_pencil

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _gpu(H, T, Q=None, Z=None, schur=False):
    F = lambda M: None if M is None else np.asfortranarray(M.copy())
    h, t, q, z = F(H), F(T), F(Q), F(Z)
    ar, ai, be, info = ph.hgeqz(h, t, q, z, schur=schur)
    return h, t, q, z, ar, ai, be, info


@pytest.mark.parametrize("gen", [random_pencil, illcond_pencil])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 150, 301])
def test_qz_eigenvalues_match_lapack(oracle_mod, gen, n):
    print(f"qz eig {gen.__name__} n={n}", flush=True)
    A0, B0 = gen(n, seed=300 + n)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0)
    *_, ar, ai, be, info = _gpu(H, T)
    assert info == 0
    ref = sl.eigvals(H, T, homogeneous_eigvals=True)
    assert E.match_chordal(ar + 1j * ai, be, ref[0], ref[1]) <= TOL
    ear, eai, ebe, _ = E.hgeqz(H.copy(), T.copy())
    assert E.match_chordal(ar + 1j * ai, be, ear + 1j * eai, ebe) <= TOL
    *_, lar, lai, lbe, linfo = oracle_mod.dhgeqz(H, T)
    assert linfo == 0
    assert E.match_chordal(ar + 1j * ai, be, lar + 1j * lai, lbe) <= TOL
    assert (be >= 0).all()


@pytest.mark.parametrize("gen", [random_pencil, illcond_pencil])
@pytest.mark.parametrize("n", [2, 3, 9, 80, 257])
def test_qz_schur_form_with_q_z(oracle_mod, gen, n):
    print(f"qz schur {gen.__name__} n={n}", flush=True)
    A0, B0 = gen(n, seed=400 + n)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0)
    S, P, Qs, Zs, ar, ai, be, info = _gpu(H, T, Q, Z, schur=True)
    assert info == 0
    check_schur(A0, B0, S, P, Qs, Zs, ar, ai, be)
    ref = sl.eigvals(A0, B0, homogeneous_eigvals=True)
    assert E.match_chordal(ar + 1j * ai, be, ref[0], ref[1]) <= TOL


@pytest.mark.parametrize("n,n_inf", [(31, 0), (60, 3), (90, 5)])
def test_qz_prescribed_eigenvalues(oracle_mod, n, n_inf):
    print(f"qz prescribed n={n} n_inf={n_inf}", flush=True)
    A0, B0, alpha, beta = _prescribed(n, seed=n + n_inf, n_inf=n_inf)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0)
    for schur in (False, True):
        *_, ar, ai, be, info = _gpu(H, T, Q, Z, schur=schur)
        assert info == 0
        assert int((np.abs(be) <= 1e-10 * np.hypot(ar, ai)).sum()) == n_inf
        assert E.match_chordal(ar + 1j * ai, be, alpha, beta) <= 1e-12


def test_qz_batched_pencils(oracle_mod):
    """40 pencils of orders 1..200 in one launch, eigenvalues only and then
    the Schur form with Q, Z, each pencil checked by itself."""
    import torch
    from paraht_amd.batched import Batch
    rng = np.random.default_rng(11)
    sizes = [1, 2, 3] + [int(x) for x in rng.integers(4, 200, size=37)]
    print(f"qz batched sizes={sizes}", flush=True)
    pencils = []
    for i, n in enumerate(sizes):
        A0, B0 = (illcond_pencil if i % 4 == 3 else random_pencil)(n, seed=500 + i)
        pencils.append((A0, B0) + ht_pencil(oracle_mod, A0, B0))
    for schur in (False, True):
        bt = Batch(sizes)
        for i, (A0, B0, H, T, Q, Z) in enumerate(pencils):
            bt.view(bt.A, i).copy_(torch.from_numpy(H))
            bt.view(bt.B, i).copy_(torch.from_numpy(T))
            bt.view(bt.Q, i).copy_(torch.from_numpy(Q))
            bt.view(bt.Z, i).copy_(torch.from_numpy(Z))
        ar, ai, be, info = ph.qz.hgeqz_batch(bt, schur=schur, vectors=schur)
        assert info == [0] * len(sizes)
        ar, ai, be = ar.cpu().numpy(), ai.cpu().numpy(), be.cpu().numpy()
        off = 0
        for i, (A0, B0, H, T, Q, Z) in enumerate(pencils):
            n = sizes[i]
            a, b, c = ar[off:off + n], ai[off:off + n], be[off:off + n]
            ref = sl.eigvals(A0, B0, homogeneous_eigvals=True)
            assert E.match_chordal(a + 1j * b, c, ref[0], ref[1]) <= TOL, i
            if schur:
                v = lambda pool: bt.view(pool, i).cpu().numpy()
                check_schur(A0, B0, v(bt.A), v(bt.B), v(bt.Q), v(bt.Z), a, b, c)
            off += n


def test_qz_device_entry_point_and_end_to_end(oracle_mod):
    """Column-major torch tensors through paraht_hgeqz_dev after the GPU
    reduction of the same pencil (the whole path on the device), and
    paraht_amd.qz.eigvals against the original pencil's eigenvalues."""
    import torch
    n, r, p, nq = 240, 16, 4, 8
    A0, B0 = random_pencil(n, seed=77)
    dev = torch.device("cuda:0")
    cm = lambda M: ph.column_major(torch.from_numpy(np.ascontiguousarray(M)).to(dev))
    H, T, Q, Z = ph.reduce(cm(A0), cm(B0), r, p, nq)
    ar, ai, be, info = ph.hgeqz(H, T, Q, Z, schur=True)
    assert info == 0
    ar, ai, be = ar.cpu().numpy(), ai.cpu().numpy(), be.cpu().numpy()
    S, P, Qs, Zs = (M.cpu().numpy() for M in (H, T, Q, Z))
    check_schur(A0, B0, S, P, Qs, Zs, ar, ai, be)
    ref = sl.eigvals(A0, B0)
    lam = ph.qz.eigvals(A0, B0, r, p, nq)
    assert E.match_chordal(lam, np.ones(n), ref, np.ones(n)) <= TOL


def test_qz_order_1000_eigenvalues(oracle_mod):
    """A single pencil of order 1000 (the workgroup size for large n)."""
    import time
    n = 1000
    A0, B0 = random_pencil(n, seed=1000)
    H, T, Q, Z = ht_pencil(oracle_mod, A0, B0, r=16, p=4, nq=8)
    t0 = time.perf_counter()
    *_, ar, ai, be, info = _gpu(H, T)
    print(f"qz n=1000 eigenvalues: {time.perf_counter() - t0:.2f} s", flush=True)
    assert info == 0
    ref = sl.eigvals(H, T, homogeneous_eigvals=True)
    assert E.match_chordal(ar + 1j * ai, be, ref[0], ref[1]) <= TOL


@pytest.mark.parametrize("n", [1, 2, 7, 24, 300])
def test_qz_edge_pencils(n):
    """B = 0 (every eigenvalue infinite), H = 0 (every eigenvalue zero), an
    already triangular pencil (its diagonal pairs), a zero subdiagonal in the
    middle; and a NaN entry, which returns info = n at once."""
    from test_qz_emul import _edge_pencils
    rng = np.random.default_rng(n)
    print(f"qz edge pencils n={n}", flush=True)
    for name, H, T, known in _edge_pencils(n, rng):
        for schur in (False, True):
            *_, ar, ai, be, info = _gpu(H, T, np.eye(n), np.eye(n), schur=schur)
            assert info == 0, name
            if name == "B=0":
                assert (be == 0).all(), name
            elif name == "H=0":
                assert (ar == 0).all() and (ai == 0).all() and (be > 0).all(), name
            else:
                if known is None:
                    known = sl.eigvals(H, T, homogeneous_eigvals=True)
                assert E.match_chordal(ar + 1j * ai, be, known[0], known[1]) <= TOL, name
    H = np.triu(rng.standard_normal((n, n)), -1)
    H[n // 2, n // 2] = np.nan
    *_, info = _gpu(H, np.eye(n))
    assert info == n


def test_qz_eigvals_general_b_and_batch():
    """paraht_amd.qz.eigvals on a pencil with a general B (the GPU QR of B
    first) and eigvals_batch on pencils of mixed orders (one-stage and
    two-stage reductions in one launch), against LAPACK on the original
    pencils."""
    rng = np.random.default_rng(21)
    n = 150
    A0 = rng.standard_normal((n, n))
    B0 = rng.standard_normal((n, n))
    ref = sl.eigvals(A0, B0)
    lam = ph.qz.eigvals(A0, B0, 16, 4, 8)
    assert E.match_chordal(lam, np.ones(n), ref, np.ones(n)) <= TOL
    sizes = [5, 40, 64, 65, 130, 257]
    pens = [random_pencil(k, seed=600 + k) for k in sizes]
    got = ph.qz.eigvals_batch([a for a, _ in pens], [b for _, b in pens])
    for (a, b), g in zip(pens, got):
        ref = sl.eigvals(a, b)
        assert E.match_chordal(g, np.ones(len(g)), ref, np.ones(len(g))) <= TOL
