"""QZ iteration on Hessenberg-triangular pencils (paraht_hgeqz* of
include/paraht.h): the step after the reduction, which the reference's driver
leaves to LAPACK (it stops at the HT form, src/profilereduction.f90:152-162).

    hgeqz(H, T, Q=None, Z=None, schur=False)  -> (alphar, alphai, beta, info)
        DHGEQZ(JOB='E' or 'S', COMPQ/COMPZ='N' or 'V') on one pencil, in place;
        NumPy Fortran-ordered arrays (copied to the GPU and back) or
        column-major torch CUDA tensors (the eigenvalues come back as tensors).
    hgeqz_batch(batch, schur=False, vectors=False) -> (alphar, alphai, beta, info)
        every pencil of a paraht_amd.batched.Batch (after reduce_batch) in one
        launch; eigenvalues of pencil i at sum(sizes[:i]).
    eigvals(A, B, r, p, nq)  -> complex eigenvalues alpha/beta (inf where
        beta = 0): the QR of a general B, both reduction stages, then the QZ
        iteration, on the GPU.
    eigvals_batch(As, Bs)  -> the same for many pencils (B triangular) in two
        launches: the batched reduction and the batched QZ.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PARAHT_QZ_SCHUR


def hgeqz(H, T, Q=None, Z=None, schur=False, ctx=None):
    """Generalized eigenvalues of the Hessenberg-triangular pencil (H, T);
    with schur=True also the generalized real Schur form in place of (H, T)
    and Q <- Q Qz, Z <- Z Zz for the Q, Z given.  Returns (alphar, alphai,
    beta, info); info > 0 names the last eigenvalue (1-based) that did not
    converge."""
    from . import _ctx_for, _dev_ld, _is_torch, _np_ptr, _torch_ready
    L = _lib.lib()
    n = H.shape[0]
    for M, nm in ((H, "H"), (T, "T"), (Q, "Q"), (Z, "Z")):
        if M is not None and (len(M.shape) != 2 or M.shape[0] < n or M.shape[1] < n):
            raise ValueError(f"hgeqz: {nm} has shape {tuple(M.shape)}, smaller than {n} x {n}")
    flags = PARAHT_QZ_SCHUR if schur else 0
    info = ctypes.c_int(0)
    if _is_torch(H):
        import torch
        c = ctx or _ctx_for(H)
        _torch_ready(H)
        ev = torch.zeros(3, max(n, 1), dtype=torch.float64, device=H.device)
        args = []
        for M, nm in ((H, "H"), (T, "T"), (Q, "Q"), (Z, "Z")):
            if M is None:
                args += [None, 1]
            else:
                args += [ctypes.c_void_p(M.data_ptr()), _dev_ld(M, nm)]
        _lib.check(L.paraht_hgeqz_dev(c._h, n, *args, ctypes.c_void_p(ev[0].data_ptr()),
                                      ctypes.c_void_p(ev[1].data_ptr()),
                                      ctypes.c_void_p(ev[2].data_ptr()), flags,
                                      ctypes.byref(info)), "hgeqz")
        return ev[0, :n], ev[1, :n], ev[2, :n], info.value
    ar, ai, be = np.zeros(n), np.zeros(n), np.zeros(n)
    args = []
    for M, nm in ((H, "H"), (T, "T"), (Q, "Q"), (Z, "Z")):
        if M is None:
            args += [None, 1]
        else:
            args += [_np_ptr(M, nm), M.shape[0]]
    _lib.check(L.paraht_hgeqz(n, *args, ar.ctypes.data_as(ctypes.c_void_p),
                              ai.ctypes.data_as(ctypes.c_void_p),
                              be.ctypes.data_as(ctypes.c_void_p), flags, ctypes.byref(info)),
               "hgeqz")
    return ar, ai, be, info.value


def hgeqz_batch(batch, schur=False, vectors=False, ctx=None):
    """QZ on every pencil of a reduced Batch (A pool = H, B pool = T); with
    schur=True the pools become (S, P), and with vectors=True the Q and Z
    pools are accumulated too.  Returns device tensors (alphar, alphai, beta)
    of sum(sizes) entries and the list of per-pencil statuses."""
    import torch
    from . import _ctx_for, _torch_ready
    c = ctx or _ctx_for(batch.A)
    _torch_ready(batch.A)
    tot = max(sum(batch.sizes), 1)
    ev = torch.zeros(3, tot, dtype=torch.float64, device=batch.device)
    cnt = len(batch.sizes)
    info = (ctypes.c_int * max(cnt, 1))()
    desc = batch.descriptors((0, 0, 0))
    qz = ctypes.c_void_p(batch.Q.data_ptr()) if vectors else None
    zz = ctypes.c_void_p(batch.Z.data_ptr()) if vectors else None
    _lib.check(_lib.lib().paraht_hgeqz_batched_dev(
        c._h, cnt, desc, ctypes.c_void_p(batch.A.data_ptr()), ctypes.c_void_p(batch.B.data_ptr()),
        qz, zz, ctypes.c_void_p(ev[0].data_ptr()), ctypes.c_void_p(ev[1].data_ptr()),
        ctypes.c_void_p(ev[2].data_ptr()), PARAHT_QZ_SCHUR if schur else 0, info),
        "hgeqz_batch")
    return ev[0], ev[1], ev[2], list(info)[:cnt]


def _ratios(ar, ai, be):
    alpha = np.asarray(ar) + 1j * np.asarray(ai)
    be = np.asarray(be)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(be != 0.0, alpha / np.where(be != 0.0, be, 1.0), np.inf)


def eigvals(A, B, r=16, p=4, nq=8, givens=False):
    """Generalized eigenvalues of (A, B) on the GPU: the QR of B when B is
    not upper triangular, both reduction stages, then the QZ iteration.
    Complex array alpha/beta (inf where beta = 0)."""
    from . import reduce, reduce_general
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    if np.tril(B, -1).any():
        H, T, _, _ = reduce_general(A, B, r, p, nq, givens=givens)
    else:
        H, T, _, _ = reduce(A, B, r, p, nq, givens=givens)
    ar, ai, be, info = hgeqz(H, np.asfortranarray(T))
    if info:
        raise _lib.ParahtError(f"eigvals: QZ did not converge (info={info})")
    return _ratios(ar, ai, be)


def eigvals_batch(As, Bs, ctx=None):
    """Generalized eigenvalues of many pencils (B upper triangular) on the
    GPU: one batched reduction launch (size-bucket parameters) and one
    batched QZ launch.  Returns a list of complex arrays alpha/beta."""
    from .batched import Batch, reduce_batch
    sizes = [int(a.shape[0]) for a in As]
    bt = Batch(sizes)
    bt.load(As, Bs)
    reduce_batch(bt, ctx=ctx)
    ar, ai, be, info = hgeqz_batch(bt, ctx=ctx)
    bad = [i for i, x in enumerate(info) if x]
    if bad:
        raise _lib.ParahtError(f"eigvals_batch: QZ did not converge for pencils {bad[:8]}")
    ar, ai, be = ar.cpu().numpy(), ai.cpu().numpy(), be.cpu().numpy()
    out, off = [], 0
    for n in sizes:
        out.append(_ratios(ar[off:off + n], ai[off:off + n], be[off:off + n]))
        off += n
    return out
