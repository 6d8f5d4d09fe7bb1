"""GPU tests of the benchmark configurations at their real sizes and of the
kernel paths the smaller parity tests do not reach.

* cfg1 (n = 4000, 8000; r = 64, p = 2, nq = 16, the benchmark's pencil):
  backward errors ||Q^T A Z - H||_F / ||A||_F and the same for B
  (src/profilereduction.f90:164-174) <= 10 n u, ||Q^T Q - I||_F and
  ||Z^T Z - I||_F <= 10 n u, exact zeros below the structure, Q e1 = e1.
  The checks run as FP64 GEMMs on the GPU.
* cfg2 (n = 4000, ill-conditioned B with 5% exact zeros on its diagonal):
  the same invariants.
* every (r, nq) pair with its own far-update kernels, stage 2 elementwise
  against the oracle's rht2ht_givens (the (64, 32) pair takes the
  sequential-window kernels);
* blocks of P = p r = 256 rows (the generic stage-1 generators) against the
  oracle's ab2rht_simple, and the generic generators forced at P <= 128;
* paraht_rht2ht_unblocked against the oracle's
  rht2ht_householder_unblocked (src/rht2ht.f90:2053-2180), sign-normalised;
* the lookahead split of stage 2 against the one-stream schedule;
* the error path of the sweep's bounded waits.
"""
import os

import numpy as np
import pytest

import paraht_amd as ph
from conftest import U
from paraht_amd.pencils import random_pencil

pytestmark = pytest.mark.gpu


def _gpu_pencil(n, seed, illcond=False):
    """The benchmark's pencil, drawn on the GPU: A ~ N(0,1), B = R of the QR
    of an N(0,1) matrix (cfg2: B upper triangular, N(0,1) above the
    diagonal, diagonal 10^U(-12,0) with 5% of it exactly 0)."""
    import torch
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    A = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    if illcond:
        B = torch.triu(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g), 1)
        d = 10.0 ** (-12.0 * torch.rand(n, dtype=torch.float64, device=dev, generator=g))
        zero = torch.randperm(n, device=dev, generator=g)[: round(0.05 * n)]
        d[zero] = 0.0
        B = B + torch.diag(d)
    else:
        R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g),
                            mode="r")[1]
        B = torch.triu(R)
    return ph.column_major(A), ph.column_major(B)


def _gpu_invariants(A0, B0, H, T, Q, Z):
    import torch
    n = A0.shape[0]
    ea = (torch.linalg.norm(Q.T @ A0 @ Z - H) / torch.linalg.norm(A0)).item()
    eb = (torch.linalg.norm(Q.T @ B0 @ Z - T) / torch.linalg.norm(B0)).item()
    I = torch.eye(n, dtype=torch.float64, device=A0.device)
    oq = torch.linalg.norm(Q.T @ Q - I).item()
    oz = torch.linalg.norm(Z.T @ Z - I).item()
    zh = int(torch.count_nonzero(torch.tril(H, -2)).item())
    zt = int(torch.count_nonzero(torch.tril(T, -1)).item())
    e1 = bool(torch.all(Q[1:, 0] == 0).item() and torch.all(Q[0, 1:] == 0).item()
              and Q[0, 0].item() == 1.0)
    print(f"n={n}: backward {ea:.2e} {eb:.2e}, orthogonality {oq:.2e} {oz:.2e}, "
          f"nonzeros below structure {zh} {zt}, Q e1 = e1: {e1}", flush=True)
    tol = 10 * n * U
    assert ea <= tol and eb <= tol, (ea, eb, tol)
    assert oq <= tol and oz <= tol, (oq, oz, tol)
    assert zh == 0 and zt == 0
    assert e1


@pytest.mark.parametrize("n", [4000, 8000])
def test_cfg1_invariants(n):
    print(f"cfg1 n={n} r=64 p=2 nq=16", flush=True)
    A0, B0 = _gpu_pencil(n, 1302)
    A, B = A0.clone(), B0.clone()
    H, T, Q, Z = ph.reduce(A, B, 64, 2, 16)
    _gpu_invariants(A0, B0, H, T, Q, Z)


@pytest.mark.timeout(900)
def test_cfg1_invariants_n16000():
    print("cfg1 n=16000 r=64 p=2 nq=16", flush=True)
    A0, B0 = _gpu_pencil(16000, 1302)
    A, B = A0.clone(), B0.clone()
    H, T, Q, Z = ph.reduce(A, B, 64, 2, 16)
    del A, B
    _gpu_invariants(A0, B0, H, T, Q, Z)


def test_cfg2_illconditioned_n4000():
    print("cfg2 n=4000 r=64 p=2 nq=16, ill-conditioned B", flush=True)
    A0, B0 = _gpu_pencil(4000, 7, illcond=True)
    import torch
    assert int((torch.diagonal(B0) == 0).sum().item()) == 200
    A, B = A0.clone(), B0.clone()
    H, T, Q, Z = ph.reduce(A, B, 64, 2, 16)
    _gpu_invariants(A0, B0, H, T, Q, Z)


def _stage1_pencil(O, n, r, seed):
    """An r-Hessenberg-triangular pencil from the oracle's stage 1, B's
    triangle exact (as the GPU stage 1 leaves it)."""
    A0, B0 = random_pencil(n, seed=seed)
    A, B = A0.copy(order="F"), B0.copy(order="F")
    Q, Z = np.eye(n, order="F"), np.eye(n, order="F")
    O.ab2rht_simple(n, A, B, Q, Z, r, 2)
    return A0, B0, A, np.asfortranarray(np.triu(B)), Q, Z


PAIRS = [(8, 4), (8, 8), (16, 4), (16, 8), (16, 16), (32, 8), (32, 16), (32, 32),
         (64, 8), (64, 16), (64, 32)]


@pytest.mark.parametrize("r,nq", PAIRS)
def test_every_far_update_pair_matches_oracle(oracle_mod, r, nq):
    O = oracle_mod
    n = 5 * r + 3 * nq + 7
    print(f"stage2 pair r={r} nq={nq} n={n}", flush=True)
    A0, B0, A1, B1, Q1, Z1 = _stage1_pencil(O, n, r, 400 + r + nq)
    got = [M.copy(order="F") for M in (A1, B1, Q1, Z1)]
    ph.rht2ht_givens(n, *[x for M in got for x in (M, n)], r, nq)
    ref = [M.copy(order="F") for M in (A1, B1, Q1, Z1)]
    O.rht2ht_givens(n, *ref, r, nq)
    scale = [np.linalg.norm(A0), np.linalg.norm(B0), n, n]
    for g, o, sc in zip(got, ref, scale):
        assert np.linalg.norm(g - o) / sc <= 1e-9
    ea, eb = O.backward_errors(A0, B0, *got)
    assert ea <= 10 * n * U and eb <= 10 * n * U


@pytest.mark.parametrize("n,r,p", [(600, 64, 4), (500, 32, 8), (300, 16, 16)])
def test_stage1_generic_generators_match_oracle(oracle_mod, n, r, p):
    """P = p r = 256: the generic panel QR and RQ generators (the
    register-resident ones take P <= 128)."""
    O = oracle_mod
    print(f"stage1 generic n={n} r={r} p={p}", flush=True)
    A0, B0 = random_pencil(n, seed=500 + n)
    A, B, Q, Z = (A0.copy(order="F"), B0.copy(order="F"), np.eye(n, order="F"),
                  np.eye(n, order="F"))
    ph.ab2rht_simple(n, A, n, B, n, Q, n, Z, n, r, p)
    Ao, Bo, Qo, Zo = (A0.copy(order="F"), B0.copy(order="F"), np.eye(n, order="F"),
                      np.eye(n, order="F"))
    O.ab2rht_simple(n, Ao, Bo, Qo, Zo, r, p)
    assert O.below_band(A, r) == 0.0 and O.below_band(B, 0) == 0.0
    g = O.sign_normalize(A, B, Q, Z, band=r)
    o = O.sign_normalize(Ao, Bo, Qo, Zo, band=r)
    for x, y, sc in zip(g, o, [np.linalg.norm(A0), np.linalg.norm(B0), n, n]):
        assert np.linalg.norm(x - y) / sc <= 1e-9


def test_generic_generators_forced_match_fast_ones(oracle_mod, monkeypatch):
    """PARAHT_GENERIC_S1 routes blocks of P <= 128 to the generic kernels."""
    O = oracle_mod
    n, r, p = 200, 16, 4
    A0, B0 = random_pencil(n, seed=9)
    fast = ph.reduce(A0, B0, r, p, 8)
    monkeypatch.setenv("PARAHT_GENERIC_S1", "1")
    gen = ph.reduce(A0, B0, r, p, 8)
    monkeypatch.delenv("PARAHT_GENERIC_S1")
    a = O.sign_normalize(*fast)
    b = O.sign_normalize(*gen)
    for x, y, sc in zip(a, b, [np.linalg.norm(A0), np.linalg.norm(B0), n, n]):
        assert np.linalg.norm(x - y) / sc <= 1e-11


def test_unsupported_block_size_is_refused_before_any_work():
    n = 200
    A0, B0 = random_pencil(n, seed=3)
    A = A0.copy(order="F")
    with pytest.raises(ph.ParahtError, match="unsupported"):
        ph.reduce(A0, B0, 64, 8, 16)  # P = 512 rows of 64 reflectors
    with pytest.raises(ph.ParahtError, match="unsupported"):
        ph.reduce(A0, B0, 128, 2, 16)  # stage 2 takes r <= 64
    assert np.array_equal(A, A0)


@pytest.mark.parametrize("n,r", [(150, 8), (200, 16), (131, 5)])
def test_unblocked_matches_oracle(oracle_mod, n, r):
    O = oracle_mod
    print(f"unblocked n={n} r={r}", flush=True)
    A0, B0, A1, B1, Q1, Z1 = _stage1_pencil(O, n, r, 600 + n)
    got = [M.copy(order="F") for M in (A1, B1, Q1, Z1)]
    ph.rht2ht_householder_unblocked(n, *[x for M in got for x in (M, n)], r)
    ref = [M.copy(order="F") for M in (A1, B1, Q1, Z1)]
    O.rht2ht_householder_unblocked(n, *ref, r)
    ea, eb = O.backward_errors(A0, B0, *got)
    assert ea <= 10 * n * U and eb <= 10 * n * U
    assert O.below_band(got[0], 1) == 0.0 and O.below_band(got[1], 0) == 0.0
    g = O.sign_normalize(*got)
    o = O.sign_normalize(*ref)
    for x, y, sc in zip(g, o, [np.linalg.norm(A0), np.linalg.norm(B0), n, n]):
        assert np.linalg.norm(x - y) / sc <= 1e-9


@pytest.mark.parametrize("n,r,nq", [(700, 64, 16), (1500, 64, 16), (400, 16, 8), (333, 32, 32),
                                    (200, 8, 4)])
def test_lookahead_split_matches_one_stream(oracle_mod, monkeypatch, n, r, nq):
    """The A/B far updates split between the main stream (the part the next
    sweep reads) and a second stream (the rest, beside the next sweep) give
    the one-stream result up to rounding (the left and right rotations of an
    entry change order), and match the oracle."""
    O = oracle_mod
    print(f"lookahead n={n} r={r} nq={nq}", flush=True)
    A0, B0, A1, B1, Q1, Z1 = _stage1_pencil(O, n, r, 800 + n)
    runs = []
    for flag in ("1", "0", "2"):
        monkeypatch.setenv("PARAHT_S2_LOOKAHEAD", flag)
        got = [M.copy(order="F") for M in (A1, B1, Q1, Z1)]
        ph.rht2ht_givens(n, *[x for M in got for x in (M, n)], r, nq)
        runs.append(got)
    monkeypatch.delenv("PARAHT_S2_LOOKAHEAD")
    ref = [M.copy(order="F") for M in (A1, B1, Q1, Z1)]
    O.rht2ht_givens(n, *ref, r, nq)
    scale = [np.linalg.norm(A0), np.linalg.norm(B0), n, n]
    for a, b, c, o, sc in zip(runs[0], runs[1], runs[2], ref, scale):
        assert np.linalg.norm(a - b) / sc <= 1e-12
        assert np.array_equal(a, c)  # the same order, on one stream or two
        assert np.linalg.norm(a - o) / sc <= 1e-9


def test_sweep_wait_error_is_reported(monkeypatch):
    """A wait of the call's first sweep that gives up (a spin limit of one
    poll, set from the host) is reported as a HIP error even though later
    sweeps succeed; the next call on the same context is clean."""
    import torch
    n, r, p, nq = 300, 16, 2, 8
    A0, B0 = random_pencil(n, seed=31)
    dev = torch.device("cuda:0")
    cm = lambda M: ph.column_major(torch.from_numpy(np.ascontiguousarray(M)).to(dev))
    monkeypatch.setenv("PARAHT_DEBUG_SPIN0", "1")
    with pytest.raises(ph.ParahtError, match="HIP"):
        ph.reduce(cm(A0), cm(B0), r, p, nq)
    monkeypatch.delenv("PARAHT_DEBUG_SPIN0")
    H, T, Q, Z = ph.reduce(cm(A0), cm(B0), r, p, nq)
    h = [M.cpu().numpy() for M in (H, T, Q, Z)]
    ea = np.linalg.norm(h[2].T @ A0 @ h[3] - h[0]) / np.linalg.norm(A0)
    assert ea <= 10 * n * U
