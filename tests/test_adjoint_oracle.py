"""Pins the adjoint oracle (oracle/adjoint.py) on the CPU.  The reference has
no adjoint, so the pins are:
* central finite differences (h = 1e-6) of the oracle's C restatement of the
  reduction, for the tangent (forward mode) and the VJP (reverse mode), with
  the perturbed decompositions carried to the signs of the unperturbed one
  (the reduction is smooth only between rotation-branch surfaces);
* the transpose identity <g, tangent(E)> = <transpose_recurrence(g), E>
  between the two separately written recurrences;
* torch FP64 autograd through the restated unblocked reduction
  (oracle/autograd_ref.py, src/rht2ht.f90:2053-2180), itself pinned to the C
  restatement of the same routine;
* a planted mistake in the VJP fails the finite-difference check.
"""
import numpy as np
import pytest
import torch

import adjoint as AD
import autograd_ref as AR
from paraht_amd.pencils import random_pencil

H_FD = 1e-6
FD_RTOL = 1e-6


def _weights(n, seed):
    rng = np.random.default_rng(seed)
    return tuple(rng.standard_normal((n, n)) for _ in range(4))


def _directions(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)), np.triu(rng.standard_normal((n, n)))


def _fwd(oracle_mod, r, p, nq):
    return lambda A, B: oracle_mod.reduce(A, B, r, p, nq, stage2="givens")


@pytest.mark.parametrize("n", [12, 24, 37])
def test_torch_restatement_matches_c_unblocked(oracle_mod, n):
    A0, B0 = random_pencil(n, seed=100 + n)
    A, B = np.asfortranarray(A0.copy()), np.asfortranarray(B0.copy())
    Q, Z = np.eye(n, order="F"), np.eye(n, order="F")
    oracle_mod.rht2ht_householder_unblocked(n, A, B, Q, Z, n - 1)
    with torch.no_grad():
        out = AR.reduce_unblocked(torch.from_numpy(A0), torch.from_numpy(B0))
    for got, want in zip(out, (A, B, Q, Z)):
        assert np.linalg.norm(got.numpy() - want) <= 1e-12 * max(1.0, np.linalg.norm(want))


@pytest.mark.parametrize("n,r,p,nq", [(10, 2, 2, 2), (24, 4, 2, 4), (33, 8, 2, 4)])
def test_tangent_matches_central_difference(oracle_mod, n, r, p, nq):
    A0, B0 = random_pencil(n, seed=n)
    fwd = _fwd(oracle_mod, r, p, nq)
    H, T, Q, Z = fwd(A0, B0)
    dA, dB = _directions(n, seed=n + 1)
    ana = AD.jvp(H, T, Q, Z, dA, dB)
    plus = fwd(A0 + H_FD * dA, B0 + H_FD * dB)
    minus = fwd(A0 - H_FD * dA, B0 - H_FD * dB)
    for k, (a, p_, m_) in enumerate(zip(ana, plus, minus)):
        fd = (p_ - m_) / (2 * H_FD)
        if k == 1:
            fd = np.triu(fd)
        assert np.linalg.norm(a - fd) <= FD_RTOL * np.linalg.norm(fd), k


@pytest.mark.parametrize("n", [3, 9, 30])
def test_transpose_identity(n):
    rng = np.random.default_rng(n)
    H = np.triu(rng.standard_normal((n, n)), -1)
    T = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
    EA, EB = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    gQ = np.tril(rng.standard_normal((n, n)), -1)
    gQ[:, 0] = 0.0
    gZ = np.tril(rng.standard_normal((n, n)), -1)
    wQ, wZ = AD.tangent(H, T, EA, EB)
    XA, XB = AD.transpose_recurrence(H, T, gQ, gZ)
    lhs = np.sum(gQ * wQ) + np.sum(gZ * wZ)
    rhs = np.sum(XA * EA) + np.sum(XB * EB)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.linalg.norm(XA) + np.linalg.norm(XB))
    assert np.all(np.triu(XA, -1) == 0) and np.all(np.triu(XB) == 0)


@pytest.mark.parametrize("n,r,p,nq", [(6, 2, 2, 2), (16, 4, 2, 4), (40, 8, 3, 4)])
def test_vjp_matches_central_difference(oracle_mod, n, r, p, nq):
    A0, B0 = random_pencil(n, seed=7 * n)
    fwd = _fwd(oracle_mod, r, p, nq)
    outs = fwd(A0, B0)
    Wt = _weights(n, seed=n)
    Ab, Bb = AD.vjp(*outs, *Wt)
    for trial in range(2):
        dA, dB = _directions(n, seed=1000 * trial + n)
        fd = AD.central_difference(fwd, A0, B0, dA, dB, Wt, H_FD, base=outs)
        an = np.sum(Ab * dA) + np.sum(Bb * dB)
        assert abs(an - fd) <= FD_RTOL * abs(fd)


def test_fd_across_a_sign_branch(oracle_mod):
    """n = 600, (r, p, nq) = (64, 2, 16).  A step of h = 1e-6 along this
    direction crosses rotation-branch surfaces: column pairs of Q flip sign
    between the two perturbed reductions.  The map is also strongly curved
    at this n (the trailing Krylov columns of Q and Z are sensitive), so the
    directional derivative of an N(0,1) direction, here 15 against a gradient
    of Frobenius norm 3.7e4, is compared within 1e-6 of the gradient's norm
    (the typical size of such a derivative) rather than of itself."""
    n = 600
    A0, B0 = random_pencil(n, seed=5 * n)
    fwd = _fwd(oracle_mod, 64, 2, 16)
    outs = fwd(A0, B0)
    rng = np.random.default_rng(n + 2)
    dA, dB = rng.standard_normal((n, n)), np.triu(rng.standard_normal((n, n)))
    plus = fwd(A0 + H_FD * dA, B0 + H_FD * dB)
    minus = fwd(A0 - H_FD * dA, B0 - H_FD * dB)
    assert np.any(np.sum(plus[2] * minus[2], axis=0) < 0)  # a branch is crossed
    Wt = _weights(n, seed=n + 1)
    Ab, Bb = AD.vjp(*outs, *Wt)
    fd = AD.central_difference(fwd, A0, B0, dA, dB, Wt, H_FD, base=outs)
    an = np.sum(Ab * dA) + np.sum(Bb * dB)
    gnorm = np.sqrt(np.sum(Ab * Ab) + np.sum(Bb * Bb))
    assert abs(an - fd) <= FD_RTOL * gnorm
    unaligned = (AD.weighted_loss(plus, Wt) - AD.weighted_loss(minus, Wt)) / (2 * H_FD)
    assert abs(an - unaligned) > 1e-3 * gnorm  # without the alignment it fails


@pytest.mark.parametrize("n", [8, 20])
def test_vjp_matches_torch_autograd(oracle_mod, n):
    A0, B0 = random_pencil(n, seed=3 * n)
    outs = oracle_mod.reduce(A0, B0, 4, 2, 2, stage2="givens")
    Wt = _weights(n, seed=n + 5)
    Ab, Bb = AD.vjp(*outs, *Wt)
    At = torch.from_numpy(A0).requires_grad_()
    Bt = torch.from_numpy(B0).requires_grad_()
    ref = AR.reduce_unblocked(At, Bt)
    signs = AR.sign_align(ref[2], ref[3], torch.from_numpy(outs[2]), torch.from_numpy(outs[3]))
    loss = AR.aligned_loss(ref, signs, tuple(torch.from_numpy(w) for w in Wt))
    gA, gB = torch.autograd.grad(loss, (At, Bt))
    assert abs(loss.item() - AD.weighted_loss(outs, Wt)) <= 1e-10 * n * n
    assert np.linalg.norm(Ab - gA.numpy()) <= 1e-9 * np.linalg.norm(Ab)
    assert np.linalg.norm(Bb - np.triu(gB.numpy())) <= 1e-9 * np.linalg.norm(Bb)


def test_planted_mistake_fails_fd(oracle_mod):
    """A transposed Tinv in the recurrence (a plausible slip) must fail."""
    n = 16
    A0, B0 = random_pencil(n, seed=11)
    fwd = _fwd(oracle_mod, 4, 2, 4)
    outs = fwd(A0, B0)
    Wt = _weights(n, seed=12)
    real_inv = np.linalg.inv
    try:
        AD.np.linalg.inv = lambda M: real_inv(M).T
        Ab, Bb = AD.vjp(*outs, *Wt)
    finally:
        AD.np.linalg.inv = real_inv
    dA, dB = _directions(n, seed=13)
    fd = AD.central_difference(fwd, A0, B0, dA, dB, Wt, H_FD, base=outs)
    an = np.sum(Ab * dA) + np.sum(Bb * dB)
    assert abs(an - fd) > 1e-3 * abs(fd)
