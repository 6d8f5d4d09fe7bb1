"""Batched reduction of many independent small pencils in one persistent
launch (paraht_batched_dev of include/paraht.h), with the size-bucket
autotuner of (r, p, nq).

Pencils are packed into four device pools (A, B, Q, Z), each pencil's n x n
column-major matrices starting at its offset (leading dimension n).
"""
from __future__ import annotations

import ctypes
import time

from . import _lib
from ._lib import PARAHT_BATCH_GRAPH, Pencil

# (r, nq) pairs and p values the batched kernel supports (r <= 16, p*r <= 128)
PAIRS = [(8, 4), (8, 8), (16, 4), (16, 8), (16, 16)]
P_VALUES = [2, 4, 8]
# r = 1: the one-stage reduction (DGGHRD's Givens sequence with the pencil in
# LDS) for pencils of order <= ONE_STAGE_MAX; p and nq are not used
ONE_STAGE = (1, 2, 1)
ONE_STAGE_MAX = 64
# size buckets of n (upper bounds) the defaults and the autotuner use
BUCKETS = [64, 96, 192, 320, 512]


def defaults(n: int):
    """(r, p, nq) of the library's size-bucket table for order n."""
    r, p, nq = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.lib().paraht_batched_defaults(n, ctypes.byref(r), ctypes.byref(p),
                                                  ctypes.byref(nq)), "paraht_batched_defaults")
    return r.value, p.value, nq.value


class Batch:
    """Packed pools for a list of square FP64 pencils on one CUDA device."""

    def __init__(self, sizes, device="cuda:0"):
        import torch
        self.sizes = [int(n) for n in sizes]
        self.offsets = []
        off = 0
        for n in self.sizes:
            self.offsets.append(off)
            off += n * n
        self.total = max(off, 1)
        self.device = torch.device(device)
        z = lambda: torch.zeros(self.total, dtype=torch.float64, device=self.device)
        self.A, self.B, self.Q, self.Z = z(), z(), z(), z()

    def view(self, pool, i):
        """Pencil i's matrix in `pool` as an n x n column-major tensor view."""
        n, off = self.sizes[i], self.offsets[i]
        return pool[off:off + n * n].view(n, n).t()

    def load(self, As, Bs):
        """Copies the pencils in, sets Q = Z = I."""
        import torch
        for i, (a, b) in enumerate(zip(As, Bs)):
            n = self.sizes[i]
            self.view(self.A, i).copy_(torch.as_tensor(a, dtype=torch.float64))
            self.view(self.B, i).copy_(torch.as_tensor(b, dtype=torch.float64))
            self.view(self.Q, i).copy_(torch.eye(n, dtype=torch.float64))
            self.view(self.Z, i).copy_(torch.eye(n, dtype=torch.float64))

    def descriptors(self, params=None):
        """paraht_pencil array; params: None (library defaults), one (r, p, nq)
        for all, or a callable n -> (r, p, nq)."""
        arr = (Pencil * len(self.sizes))()
        for i, n in enumerate(self.sizes):
            if params is None:
                r = p = nq = 0
            elif callable(params):
                r, p, nq = params(n)
            else:
                r, p, nq = params
            arr[i] = Pencil(self.offsets[i], n, r, p, nq)
        return arr


def reduce_batch(batch: Batch, params=None, ctx=None, graph=False, desc=None):
    """Reduces every pencil of `batch` in place (A -> H, B -> T, Q, Z
    accumulated) with one persistent launch."""
    from . import _ctx_for, _torch_ready
    ctx = ctx or _ctx_for(batch.A)
    _torch_ready(batch.A)
    d = desc if desc is not None else batch.descriptors(params)
    flags = PARAHT_BATCH_GRAPH if graph else 0
    _lib.check(_lib.lib().paraht_batched_dev(
        ctx._h, len(batch.sizes), d, ctypes.c_void_p(batch.A.data_ptr()),
        ctypes.c_void_p(batch.B.data_ptr()), ctypes.c_void_p(batch.Q.data_ptr()),
        ctypes.c_void_p(batch.Z.data_ptr()), flags), "paraht_batched_dev")
    return batch


def autotune(sample, seconds_per_candidate=None, ctx=None, candidates=None):
    """Size-bucket autotuner: for each bucket of BUCKETS, times the batched
    launch on the sample pencils of that bucket with every supported
    (r, p, nq) (the one-stage path too in buckets of order <= ONE_STAGE_MAX)
    and returns {bucket_upper_bound: (r, p, nq)} of the fastest.
    sample: list of (A, B) host or device arrays."""
    import torch
    cands = candidates or [(r, p, nq) for (r, nq) in PAIRS for p in P_VALUES if p * r <= 128]
    table = {}
    lo = 0
    for hi in BUCKETS:
        sub = [(a, b) for (a, b) in sample if lo < a.shape[0] <= hi]
        lo = hi
        if not sub:
            continue
        best = None
        bucket = list(cands)
        if candidates is None and hi <= ONE_STAGE_MAX:
            bucket.append(ONE_STAGE)
        for c in bucket:
            bt = Batch([a.shape[0] for a, _ in sub])
            bt.load([a for a, _ in sub], [b for _, b in sub])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reduce_batch(bt, params=c, ctx=ctx)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, c)
        table[hi] = best[1]
    return table
