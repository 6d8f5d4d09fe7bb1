"""GPU: the device entry points on a caller's stream.

The library runs the Q and Z updates (and stage 1's left update of A) on a
second stream of its own, ordered against the context's stream by events
(include/paraht.h, paraht_ctx_set_stream).  These tests run the full
reduction on a caller-created torch stream and on the context's own stream,
twice each in a row, and require bit-identical results (the same kernels in
the same order on the same inputs) that also pass the oracle's invariants.
"""
import numpy as np
import pytest

import paraht_amd as ph
from conftest import U
from paraht_amd.pencils import random_pencil

pytestmark = pytest.mark.gpu

CASES = [(400, 64, 2, 16), (300, 16, 8, 16), (257, 8, 4, 4)]


def _run(torch, A0, B0, r, p, nq, ctx, stream):
    dev = torch.device("cuda:0")
    with torch.cuda.stream(stream):
        A = ph.column_major(torch.from_numpy(A0).to(dev))
        B = ph.column_major(torch.from_numpy(B0).to(dev))
        H, T, Q, Z = ph.reduce(A, B, r, p, nq, ctx=ctx)
        out = [M.cpu().numpy() for M in (H, T, Q, Z)]
    return out


@pytest.mark.parametrize("n,r,p,nq", CASES)
def test_caller_stream_matches_own_stream(oracle_mod, n, r, p, nq):
    import torch
    print(f"streams n={n} r={r} p={p} nq={nq}", flush=True)
    O = oracle_mod
    A0, B0 = random_pencil(n, seed=700 + n)
    own = ph.Context(0)
    ref = _run(torch, A0, B0, r, p, nq, own, torch.cuda.current_stream())
    s = torch.cuda.Stream()
    ctx = ph.Context(0)
    ctx.set_stream(s)
    for rep in range(2):
        got = _run(torch, A0, B0, r, p, nq, ctx, s)
        for g, e in zip(got, ref):
            assert np.array_equal(g, e), rep
    H, T, Q, Z = got
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    assert ea <= 10 * n * U and eb <= 10 * n * U, (ea, eb)
    assert O.below_band(H, 1) == 0.0 and O.below_band(T, 0) == 0.0
    # the context's own stream again, after the caller's stream was used
    ctx.set_stream(None)
    again = _run(torch, A0, B0, r, p, nq, ctx, torch.cuda.current_stream())
    for g, e in zip(again, ref):
        assert np.array_equal(g, e)
