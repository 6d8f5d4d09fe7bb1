"""Seeded test pencils (A, B) for the HT reduction.

The reference's driver draws A and B uniformly from [0,1) with the Fortran
runtime's PRNG (seed 1302) and replaces B by the R factor of its QR
(src/profilereduction.f90:38-44,114-123).  That stream depends on the
compiler, so pencils here come from NumPy's PCG64 with the same default seed
and the N(0,1) distribution of the benchmark workloads:

* ``random_pencil``      A ~ N(0,1); B = R of a Householder QR of an N(0,1)
                         matrix (LAPACK sign convention), Q = Z = I.
* ``illcond_pencil``     A ~ N(0,1); B upper triangular, off-diagonal N(0,1),
                         diagonal 10**U(-12,0) with 5% of it set exactly to 0
                         (infinite eigenvalues).
* ``uniform_pencil``     the reference driver's distribution (U[0,1) entries,
                         B = R of QR) with this generator.
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 1302


def _r_factor(M: np.ndarray) -> np.ndarray:
    R = np.linalg.qr(M, mode="r")
    return np.asfortranarray(np.triu(R))


def random_pencil(n: int, seed: int = DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    B = _r_factor(rng.standard_normal((n, n)))
    return A, B


def uniform_pencil(n: int, seed: int = DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.random((n, n)))
    B = _r_factor(rng.random((n, n)))
    return A, B


def illcond_pencil(n: int, seed: int = DEFAULT_SEED, zero_frac: float = 0.05):
    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    B = np.triu(rng.standard_normal((n, n)))
    d = 10.0 ** rng.uniform(-12.0, 0.0, size=n)
    nz = int(round(zero_frac * n))
    d[rng.permutation(n)[:nz]] = 0.0
    B[np.arange(n), np.arange(n)] = d
    return A, np.asfortranarray(B)


def batch_sizes(count: int = 2048, lo: int = 32, hi: int = 512,
                seed: int = DEFAULT_SEED):
    """Sizes of the batched workload: n ~ U{lo..hi}."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=count)
