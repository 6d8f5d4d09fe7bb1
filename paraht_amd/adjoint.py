"""Reverse mode of the Hessenberg-triangular reduction: cotangents of
(H, T, Q, Z) to cotangents of (A, B), and a torch.autograd.Function around
the GPU reduction.

For H = Q^T A Z, T = Q^T B Z (Q e1 = e1) the skew matrices Q^T dQ and Z^T dZ
are fixed, column by column, by the structure of dH and dT; the cotangents
follow from the transposed recurrence (paraht_adjoint_dev, a HIP kernel) plus
GEMMs.  The result is the gradient of the reduction as computed by
`ht_reduce`'s forward (the decomposition's signs are locally constant).
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import Pencil


def _col_major_pool(X):
    """(b, n, n) or (n, n) -> flat pool holding each matrix column-major."""
    Y = X if X.dim() == 3 else X.unsqueeze(0)
    return Y.transpose(-1, -2).contiguous().reshape(-1)


def _from_pool(pool, b, n):
    return pool.view(b, n, n).transpose(-1, -2)


def transpose_recurrence(H, T, gQ, gZ, ctx=None):
    """(X_A, X_B) of the transposed structure recurrence on the GPU
    (paraht_adjoint_dev): <gQ, w_Q> + <gZ, w_Z> = <X_A, E_A> + <X_B, E_B>
    for the strictly lower parts w_Q, w_Z of Q^T dQ, Z^T dZ.  H upper
    Hessenberg, T upper triangular and invertible, gQ strictly lower with a
    zero first column, gZ strictly lower; FP64 CUDA tensors of shape (b, n, n).
    """
    import torch
    from . import _ctx_for, _torch_ready
    b, n, _ = H.shape
    eye = torch.eye(n, dtype=H.dtype, device=H.device).expand(b, n, n)
    Tinv = torch.linalg.solve_triangular(T, eye, upper=True)
    pools = [_col_major_pool(M) for M in (H, T, Tinv, gQ, gZ)]
    XA = torch.zeros(b * n * n, dtype=H.dtype, device=H.device)
    XB = torch.zeros_like(XA)
    desc = (Pencil * b)(*[Pencil(k * n * n, n, 0, 0, 0) for k in range(b)])
    c = ctx or _ctx_for(H)
    _torch_ready(H)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.check(_lib.lib().paraht_adjoint_dev(c._h, b, desc, *[ptr(t) for t in pools],
                                             ptr(XA), ptr(XB)), "paraht_adjoint_dev")
    return _from_pool(XA, b, n), _from_pool(XB, b, n)


def ht_backward(H, T, Q, Z, Hb=None, Tb=None, Qb=None, Zb=None, ctx=None):
    """Cotangents (Abar, Bbar) from cotangents of (H, T, Q, Z); None means
    zero.  Tensors are FP64 on one CUDA device, (n, n) or batched (b, n, n)
    with one n.  Bbar is given on B's upper triangle (the domain of B)."""
    import torch
    single = H.dim() == 2
    if single:
        H, T, Q, Z = (M.unsqueeze(0) for M in (H, T, Q, Z))
        Hb, Tb, Qb, Zb = (None if M is None else M.unsqueeze(0) for M in (Hb, Tb, Qb, Zb))
    b, n, _ = H.shape
    zero = torch.zeros_like(H)
    Hb = torch.triu(Hb, -1) if Hb is not None else zero
    Tb = torch.triu(Tb) if Tb is not None else zero
    Qb = Qb if Qb is not None else zero
    Zb = Zb if Zb is not None else zero
    Ht, Tt, Qt, Zt = (M.transpose(-1, -2) for M in (H, T, Q, Z))
    GQ = Qt @ Qb - Hb @ Ht - Tb @ Tt
    GZ = Zt @ Zb + Ht @ Hb + Tt @ Tb
    gQ = torch.tril(GQ - GQ.transpose(-1, -2), -1)
    gQ[..., :, 0] = 0.0
    gZ = torch.tril(GZ - GZ.transpose(-1, -2), -1)
    XA, XB = transpose_recurrence(H, T, gQ, gZ, ctx)
    EA = Hb + XA
    EB = Tb + XB
    Ab = Q @ EA @ Zt
    Bb = torch.triu(Q @ EB @ Zt)
    if single:
        return Ab[0], Bb[0]
    return Ab, Bb


def _forward(A, B, r, p, nq):
    """GPU reduction of one pencil (n, n) or a batch (b, n, n); fresh outputs."""
    import torch

    import paraht_amd as ph
    from .batched import Batch, reduce_batch
    if A.dim() == 2:
        H, T = ph.column_major(A.detach().clone()), ph.column_major(B.detach().clone())
        r_, p_, nq_ = (r, p, nq)
        if r_ is None:
            from .batched import defaults
            r_, p_, nq_ = defaults(A.shape[0]) if A.shape[0] <= 512 else (64, 2, 16)
            if r_ == 1:  # the one-stage path exists only in the batched kernel
                r_, p_, nq_ = 8, 2, 4
        # the adjoint needs only (H, T, Q, Z), not the variant that produced
        # them: the faster Givens chase
        H, T, Q, Z = ph.reduce(H, T, r_, p_, nq_, givens=True)
        return H, T, Q, Z
    bsz, n, _ = A.shape
    bt = Batch([n] * bsz, device=A.device)
    bt.load(list(A.detach()), list(B.detach()))
    reduce_batch(bt, params=None if r is None else (r, p, nq))
    out = [torch.stack([bt.view(M, k) for k in range(bsz)]) for M in (bt.A, bt.B, bt.Q, bt.Z)]
    return tuple(out)


def _make_function():
    import torch

    class HTReduce(torch.autograd.Function):
        """(A, B) -> (H, T, Q, Z) on the GPU with the adjoint as backward."""

        @staticmethod
        def forward(fctx, A, B, r, p, nq):
            H, T, Q, Z = _forward(A, B, r, p, nq)
            fctx.save_for_backward(H, T, Q, Z)
            return H, T, Q, Z

        @staticmethod
        def backward(fctx, Hb, Tb, Qb, Zb):
            H, T, Q, Z = fctx.saved_tensors
            Ab, Bb = ht_backward(H, T, Q, Z, Hb, Tb, Qb, Zb)
            return Ab, Bb, None, None, None

    return HTReduce


_HT = None


def ht_reduce(A, B, r=None, p=None, nq=None):
    """Differentiable HT reduction: A, B FP64 CUDA tensors (n, n) or (b, n, n)
    (B upper triangular); returns (H, T, Q, Z).  A batch goes through the
    batched persistent path.  r, p, nq default to the size-bucket table."""
    global _HT
    if _HT is None:
        _HT = _make_function()
    return _HT.apply(A, B, r, p, nq)
