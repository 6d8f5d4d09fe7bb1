"""Autotuner of the single-pencil reduction's block parameters.

For a pencil order n and bandwidth r, times paraht_reduce_dev (both stages,
the pencil resident in HBM) over candidate (p, nq) -- p: the nb-row blocks
per stage-1 reflector block, nq: the columns per stage-2 sweep -- and keeps
the fastest.  The table lives in tuned.json beside this file:
{"<n>,<r>": {"p": p, "nq": nq, "s": seconds, "candidates": {"p,nq": seconds}}}.

    python -m paraht_amd.tune 4000 8000        # r = 64, writes tuned.json

``tuned_params(n, r)`` returns the table's (p, nq) for the nearest tuned n
with the same r, or None.
"""
from __future__ import annotations

import json
import os
import sys
import time

_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
# (p, nq) the library takes at r = 64: p r rows per generator block (<= 256,
# the generic generators' LDS) and (64, nq) far-update pairs
CANDIDATES_R64 = [(p, nq) for p in (2, 3, 4) for nq in (8, 16, 32)]


def load_table(path: str = _TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def tuned_params(n: int, r: int, path: str = _TABLE):
    """(p, nq) for the tuned order nearest to n (same r), or None."""
    table = load_table(path)
    best = None
    for key, v in table.items():
        kn, kr = (int(x) for x in key.split(","))
        if kr != r:
            continue
        if best is None or abs(kn - n) < abs(best[0] - n):
            best = (kn, v["p"], v["nq"])
    return None if best is None else (best[1], best[2])


def time_reduction(n: int, r: int, p: int, nq: int, reps: int = 1, seed: int = 1302) -> float:
    """Best of `reps` wall times of one device-resident reduction (after one
    untimed run) of the benchmark's pencil: A ~ N(0,1), B = R of a QR of an
    N(0,1) matrix, Q = Z = I."""
    import torch

    import paraht_amd as ph
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    A0 = ph.column_major(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g),
                        mode="r")[1]
    B0 = ph.column_major(torch.triu(R))
    del R
    eye = ph.column_major(torch.eye(n, dtype=torch.float64, device=dev))
    A, B, Q, Z = (torch.empty_like(A0) for _ in range(4))
    ctx = ph.Context(0)
    best = None
    for i in range(reps + 1):
        A.copy_(A0)
        B.copy_(B0)
        Q.copy_(eye)
        Z.copy_(eye)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ph.reduce(A, B, r, p, nq, Q, Z, ctx=ctx)
        t = time.perf_counter() - t0
        if i > 0:
            best = t if best is None else min(best, t)
    ctx.close()
    return best


def autotune(n: int, r: int = 64, candidates=None, reps: int = 1, path: str = _TABLE) -> dict:
    """Times every candidate (p, nq) at order n, records the fastest in the
    table and returns its entry."""
    cands = candidates or (CANDIDATES_R64 if r == 64 else [(p, nq) for p in (2, 4, 8)
                                                           for nq in (4, 8, 16)])
    times = {}
    for p, nq in cands:
        t = time_reduction(n, r, p, nq, reps)
        times[f"{p},{nq}"] = t
        print(f"n={n} r={r} p={p} nq={nq}: {t:.3f} s", flush=True)
    key = min(times, key=times.get)
    p, nq = (int(x) for x in key.split(","))
    table = load_table(path)
    table[f"{n},{r}"] = {"p": p, "nq": nq, "s": times[key], "candidates": times}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return table[f"{n},{r}"]


if __name__ == "__main__":
    for arg in sys.argv[1:] or ["8000"]:
        print(autotune(int(arg)), flush=True)
