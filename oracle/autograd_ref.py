"""Torch FP64 autograd restatement of the reference's unblocked reduction --
TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker of the adjoint).

``reduce_unblocked`` follows src/rht2ht.f90:2053-2180
(rht2ht_householder_unblocked) step by step with differentiable torch ops, so
that torch's autograd gives an independent reverse mode of the reduction.
With r = n-1 a full A is "r-Hessenberg" and the routine is the one-stage
Householder reduction of a general pencil (B upper triangular); the chase
loop (src/rht2ht.f90:2129-2173) then has no iterations.

Each reflector is applied as a full n x n matrix (O(n^4) in total): this is
a reference for small n, not a fast path.
"""
from __future__ import annotations

import torch


def _larfg(x):
    """(v, tau, beta) of LAPACK dlarfg: (I - tau v v^T) x = beta e1, v[0] = 1,
    beta = -sign(x[0]) ||x||.  x must have a nonzero tail."""
    alpha = x[0]
    beta = -torch.copysign(torch.linalg.vector_norm(x), alpha)
    tau = (beta - alpha) / beta
    v = torch.cat([torch.ones(1, dtype=x.dtype), x[1:] / (alpha - beta)])
    return v, tau, beta


def _reflector(n, lo, v, tau):
    """I - tau w w^T with w = v placed at rows lo .. lo+len(v)-1."""
    w = torch.cat([torch.zeros(lo, dtype=v.dtype), v,
                   torch.zeros(n - lo - v.shape[0], dtype=v.dtype)])
    return torch.eye(n, dtype=v.dtype) - tau * torch.outer(w, w)


def _rq_first_row(Bc):
    """First row of the orthogonal factor of the RQ factorisation Bc = R Qr
    (dgerq2 + dorgr2, src/rht2ht.f90:2102-2106), up to sign: with
    Bc^T J = Q1 R1 (J the flip), Bc = (J R1^T J)(J Q1^T), so it is the last
    column of Q1.  The reflector dlarfg builds from it does not depend on
    the sign."""
    Q1, _ = torch.linalg.qr(Bc.t().flip(-1))
    return Q1[:, -1]


def _zero(M, rows, col):
    """M with M[rows, col] = 0 (the reference's explicit zeroing)."""
    mask = torch.ones_like(M)
    mask[rows, col] = 0.0
    return M * mask


def reduce_unblocked(A, B, r=None):
    """(H, T, Q, Z) with A = Q H Z^T, B = Q T Z^T, Q e1 = e1, from an
    r-Hessenberg A (default r = n-1: any A) and an upper triangular B.
    Differentiable in A and B (FP64 CPU tensors)."""
    n = A.shape[0]
    r = n - 1 if r is None else r
    Q = torch.eye(n, dtype=A.dtype)
    Z = torch.eye(n, dtype=A.dtype)
    for c in range(n - 2):                     # jc = c + 1
        nb = min(r, n - 1 - c)
        # column c of A: rows c+1 .. c+nb (src/rht2ht.f90:2085-2100)
        v, tau, _ = _larfg(A[c + 1:c + 1 + nb, c])
        P = _reflector(n, c + 1, v, tau)
        A = _zero(P @ A, slice(c + 2, c + 1 + nb), c)
        B = P @ B
        Q = Q @ P
        # restore B's triangularity: reflector from the RQ row
        # (src/rht2ht.f90:2102-2122)
        x = _rq_first_row(B[c + 1:c + 1 + nb, c + 1:c + 1 + nb])
        w, t, _ = _larfg(x)
        P2 = _reflector(n, c + 1, w, t)
        A = A @ P2
        B = _zero(B @ P2, slice(c + 2, c + 1 + nb), c + 1)
        Z = Z @ P2
        # chase the fill (src/rht2ht.f90:2129-2173), jb = b + 1
        for b in range(c + 1, n - r, r):
            nb = min(r, n - b - r)
            v, tau, _ = _larfg(A[b + r:b + r + nb, b])
            P = _reflector(n, b + r, v, tau)
            A = _zero(P @ A, slice(b + r + 1, b + r + nb), b)
            B = P @ B
            Q = Q @ P
            x = _rq_first_row(B[b + r:b + r + nb, b + r:b + r + nb])
            w, t, _ = _larfg(x)
            P2 = _reflector(n, b + r, w, t)
            A = A @ P2
            B = _zero(B @ P2, slice(b + r + 1, b + r + nb), b + r)
            Z = Z @ P2
    return A, B, Q, Z


def sign_align(Qr, Zr, Q, Z):
    """Diagonal signs (dq, dz) with Q ~ Qr diag(dq), Z ~ Zr diag(dz)."""
    dq = torch.sign(torch.einsum("ij,ij->j", Qr.detach(), Q))
    dz = torch.sign(torch.einsum("ij,ij->j", Zr.detach(), Z))
    return dq, dz


def aligned_loss(outs, signs, weights):
    """Weighted loss of (D_Q H D_Z, D_Q T D_Z, Q D_Q, Z D_Z): the reference's
    decomposition carried to the signs of another one."""
    H, T, Q, Z = outs
    dq, dz = signs
    W, V, P, S = weights
    Hs = dq[:, None] * H * dz[None, :]
    Ts = dq[:, None] * T * dz[None, :]
    return ((torch.triu(W, -1) * Hs).sum() + (torch.triu(V) * Ts).sum()
            + (P * (Q * dq[None, :])).sum() + (S * (Z * dz[None, :])).sum())
