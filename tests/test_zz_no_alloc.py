"""The allocation-free entry point (paraht_reduce_dev_noalloc, through the
C-ABI) against paraht_reduce_dev on the same pencils: two calls in a row, the
second with a smaller pencil (the shared scratch array keeps the first
call's contents), both on a caller's stream.  The single-stream schedule
applies every rotation and reflector to every entry in the same order as the
multi-stream one, so the results agree bit for bit; the test also checks the
backward errors (<= 10 n u) and refuses a pencil past the scratch array."""
import ctypes

import numpy as np
import pytest

import paraht_amd as ph
from conftest import U
from paraht_amd import _lib

pytestmark = pytest.mark.gpu


def _pencil(n, seed):
    import torch
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    A = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g),
                        mode="r")[1]
    eye = torch.eye(n, dtype=torch.float64, device=dev)
    return [ph.column_major(M) for M in (A, torch.triu(R), eye, eye.clone())]


def _ptrs(mats):
    out = []
    for M in mats:
        out += [ctypes.c_void_p(M.data_ptr()), M.stride(1)]
    return out


@pytest.mark.parametrize("sizes", [((1500, 64, 2, 16), (700, 64, 2, 16)),
                                   ((900, 16, 8, 8), (333, 8, 4, 4))])
def test_noalloc_matches_reduce_dev_twice(sizes):
    import torch
    L = _lib.lib()
    s = torch.cuda.Stream()
    for call, (n, r, p, nq) in enumerate(sizes):
        print(f"noalloc call {call}: n={n} r={r} p={p} nq={nq}", flush=True)
        ref = [M.clone() for M in _pencil(n, 40 + call)]
        got = [M.clone() for M in ref]
        A0, B0 = ref[0].clone(), ref[1].clone()
        torch.cuda.synchronize()
        ph.reduce(ref[0], ref[1], r, p, nq, ref[2], ref[3])
        st = L.paraht_reduce_dev_noalloc(ctypes.c_void_p(s.cuda_stream), n, *_ptrs(got),
                                         r, p, nq, 0)
        _lib.check(st, "paraht_reduce_dev_noalloc")
        for g, e in zip(got, ref):
            assert torch.equal(g, e)
        H, T, Q, Z = got
        ea = (torch.linalg.norm(Q.T @ A0 @ Z - H) / torch.linalg.norm(A0)).item()
        eb = (torch.linalg.norm(Q.T @ B0 @ Z - T) / torch.linalg.norm(B0)).item()
        assert ea <= 10 * n * U and eb <= 10 * n * U, (ea, eb)


def test_noalloc_refuses_beyond_its_scratch():
    import torch
    L = _lib.lib()
    n = 20000  # past the 64 MiB scratch at r = 64, p = 2
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    ptr = ctypes.c_void_p(dummy.data_ptr())
    st = L.paraht_reduce_dev_noalloc(None, n, ptr, n, ptr, n, ptr, n, ptr, n, 64, 2, 16, 0)
    assert st == 4  # PARAHT_ERR_UNSUPPORTED, before any device work
    st = L.paraht_reduce_dev_noalloc(None, 10, ptr, 10, ptr, 10, ptr, 10, ptr, 10, 2, 2, 2, 1)
    assert st == -14  # per-phase timing is not available here
