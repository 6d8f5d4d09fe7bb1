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

Pencil files, the format the Fortran driver reads (pencil=FILE) and writes
(out=FILE): the 8 bytes "PHTPENC1" (a reduced pencil: "PHTREDU1"), the order n
as a little-endian int64, then the matrices (A, B; or H, T, Q, Z), each n x n
column-major little-endian FP64.  ``save_pencil`` / ``load_pencil`` and
``save_reduced`` / ``load_reduced`` read and write them.
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


_PENCIL_MAGIC = b"PHTPENC1"
_REDUCED_MAGIC = b"PHTREDU1"


def _write(path, magic, mats):
    n = mats[0].shape[0]
    for M in mats:
        if M.shape != (n, n):
            raise ValueError(f"every matrix must be {n} x {n}, got {M.shape}")
    with open(path, "wb") as f:
        f.write(magic)
        f.write(np.int64(n).astype("<i8").tobytes())
        for M in mats:
            f.write(np.asarray(M, dtype="<f8").tobytes(order="F"))


def _read(path, magic, count):
    with open(path, "rb") as f:
        head = f.read(16)
        if len(head) != 16 or head[:8] != magic:
            raise ValueError(f"{path}: not a {magic.decode()} file")
        n = int(np.frombuffer(head[8:], dtype="<i8")[0])
        if n < 0:
            raise ValueError(f"{path}: negative order {n}")
        data = np.frombuffer(f.read(), dtype="<f8")
    if data.size != count * n * n:
        raise ValueError(f"{path}: {data.size} values, expected {count} x {n} x {n}")
    return [np.asfortranarray(data[k * n * n:(k + 1) * n * n].reshape((n, n), order="F"))
            for k in range(count)]


def save_pencil(path, A, B):
    """Writes (A, B) as a pencil file."""
    _write(path, _PENCIL_MAGIC, [A, B])


def load_pencil(path):
    """(A, B) of a pencil file, Fortran-ordered float64."""
    return tuple(_read(path, _PENCIL_MAGIC, 2))


def save_reduced(path, H, T, Q, Z):
    _write(path, _REDUCED_MAGIC, [H, T, Q, Z])


def load_reduced(path):
    """(H, T, Q, Z) of a reduced-pencil file (the driver's out=FILE)."""
    return tuple(_read(path, _REDUCED_MAGIC, 4))
