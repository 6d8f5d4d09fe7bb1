"""CPU oracle for the two-stage HT reduction -- TEST INFRASTRUCTURE ONLY.

Only the test-suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker.  The product path
(``paraht_amd`` over ``libparaht_hip.so``) never imports it.

Contents
--------
* ctypes bindings of ``liboracle.so``, the C restatement of the reference:
  ``ab2rht_simple`` (src/ab2rht.f90:8-495), ``rht2ht_householder``
  (src/rht2ht.f90:1346-2051), ``rht2ht_givens`` (src/rht2ht.f90:696-1344) and
  ``rht2ht_householder_unblocked`` (src/rht2ht.f90:2053-2180), with the
  reference's argument lists and in-place semantics on column-major arrays;
* LAPACK DGGHD3 from the host's LAPACK (scipy's OpenBLAS) -- an independent
  library routine that produces the same Hessenberg-triangular form up to
  diagonal signs (generalized implicit-Q theorem), used to pin the oracle;
* ``qr_b``: B <- R of LAPACK DGEQRF with exact zeros below the diagonal
  (src/profilereduction.f90:117-123), the Householder factor carried into A
  and Q (DORMQR, DORGQR), as LAPACK's DGGES does before DGGHD3;
* the reference's result metrics (src/profilereduction.f90:164-174) plus the
  checks the reference lacks: orthogonality, exact structural zeros, sign
  normalisation and generalized eigenvalues in the chordal metric.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ci = ctypes.c_int
        sig = [ci, dp, ci, dp, ci, dp, ci, dp, ci, ci, ci]
        for name in ("oracle_ab2rht_simple", "oracle_rht2ht_householder",
                     "oracle_rht2ht_givens"):
            f = getattr(L, name)
            f.argtypes = sig
            f.restype = ci
        L.oracle_rht2ht_householder_unblocked.argtypes = sig[:-1]
        L.oracle_rht2ht_householder_unblocked.restype = ci
        L.oracle_num_threads.restype = ci
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _call(name, n, A, B, Q, Z, *ints):
    info = getattr(lib(), name)(n, _ptr(A), A.shape[0], _ptr(B), B.shape[0],
                                _ptr(Q), Q.shape[0], _ptr(Z), Z.shape[0], *ints)
    if info != 0:
        raise ValueError(f"{name}: illegal argument {-info}")


def ab2rht_simple(n, A, B, Q, Z, nb, p):
    """In place, like src/ab2rht.f90:8 (arrays are Fortran-ordered float64)."""
    _call("oracle_ab2rht_simple", n, A, B, Q, Z, nb, p)


def rht2ht_householder(n, A, B, Q, Z, r, nq):
    """In place, like src/rht2ht.f90:1346."""
    _call("oracle_rht2ht_householder", n, A, B, Q, Z, r, nq)


def rht2ht_givens(n, A, B, Q, Z, r, nq):
    """In place, like src/rht2ht.f90:696."""
    _call("oracle_rht2ht_givens", n, A, B, Q, Z, r, nq)


def rht2ht_householder_unblocked(n, A, B, Q, Z, r):
    """In place, like src/rht2ht.f90:2053."""
    _call("oracle_rht2ht_householder_unblocked", n, A, B, Q, Z, r)


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def reduce(A0, B0, r, p, nq, stage2="givens"):
    """Both stages on copies: returns H, T, Q, Z with Q^T A0 Z = H."""
    n = A0.shape[0]
    A = np.asfortranarray(A0, dtype=np.float64).copy(order="F")
    B = np.asfortranarray(B0, dtype=np.float64).copy(order="F")
    Q = np.eye(n, order="F")
    Z = np.eye(n, order="F")
    ab2rht_simple(n, A, B, Q, Z, r, p)
    if stage2 == "givens":
        rht2ht_givens(n, A, B, Q, Z, r, nq)
    elif stage2 == "householder":
        rht2ht_householder(n, A, B, Q, Z, r, nq)
    elif stage2 == "unblocked":
        rht2ht_householder_unblocked(n, A, B, Q, Z, r)
    else:
        raise ValueError(stage2)
    return A, B, Q, Z


# --------------------------------------------------------------------------
# Host LAPACK DGGHD3 (an independent library routine used as a pin)
# --------------------------------------------------------------------------
_LAPACK = None


def _lapack():
    global _LAPACK
    if _LAPACK is None:
        import scipy
        d = os.path.join(os.path.dirname(scipy.__file__), os.pardir, "scipy.libs")
        cands = sorted(glob.glob(os.path.join(d, "libscipy_openblas-*.so")))
        if not cands:
            raise RuntimeError("scipy's OpenBLAS not found")
        _LAPACK = ctypes.CDLL(cands[0])
    return _LAPACK


def dgghd3(A0, B0):
    """LAPACK DGGHD3('I','I',n,1,n,...) on copies; returns H, T, Q, Z."""
    L = _lapack()
    f = L.scipy_dgghd3_
    n = A0.shape[0]
    A = np.array(A0, dtype=np.float64, order="F")
    B = np.array(B0, dtype=np.float64, order="F")
    Q = np.zeros((n, n), order="F")
    Z = np.zeros((n, n), order="F")
    ci = ctypes.c_int
    nn, ilo, ihi, ld = ci(n), ci(1), ci(n), ci(max(1, n))
    info = ci(0)
    lwork = ci(-1)
    wq = np.zeros(1)
    cq = ctypes.c_char(b"I")
    args_tail = (ctypes.c_size_t(1), ctypes.c_size_t(1))
    f(ctypes.byref(cq), ctypes.byref(cq), ctypes.byref(nn), ctypes.byref(ilo),
      ctypes.byref(ihi), _ptr(A), ctypes.byref(ld), _ptr(B), ctypes.byref(ld),
      _ptr(Q), ctypes.byref(ld), _ptr(Z), ctypes.byref(ld),
      wq.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(lwork),
      ctypes.byref(info), *args_tail)
    lw = max(1, int(wq[0]))
    work = np.zeros(lw)
    lwork = ci(lw)
    f(ctypes.byref(cq), ctypes.byref(cq), ctypes.byref(nn), ctypes.byref(ilo),
      ctypes.byref(ihi), _ptr(A), ctypes.byref(ld), _ptr(B), ctypes.byref(ld),
      _ptr(Q), ctypes.byref(ld), _ptr(Z), ctypes.byref(ld),
      work.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
      ctypes.byref(lwork), ctypes.byref(info), *args_tail)
    if info.value != 0:
        raise RuntimeError(f"dgghd3 info={info.value}")
    return A, B, Q, Z


def dgghrd(A0, B0):
    """LAPACK DGGHRD('I','I',n,1,n,...) on copies (B0 upper triangular): the
    unblocked Givens reduction to HT form; returns H, T, Q, Z.  The pin of
    the batched path's one-stage kernel for tiny pencils, which applies the
    same rotations in the same order."""
    L = _lapack()
    f = L.scipy_dgghrd_
    n = A0.shape[0]
    A = np.array(A0, dtype=np.float64, order="F")
    B = np.array(B0, dtype=np.float64, order="F")
    Q = np.zeros((max(n, 1), max(n, 1)), order="F")
    Z = np.zeros((max(n, 1), max(n, 1)), order="F")
    ci = ctypes.c_int
    nn, ilo, ihi, ld = ci(n), ci(1), ci(n), ci(max(1, n))
    info = ci(0)
    cq = ctypes.c_char(b"I")
    f(ctypes.byref(cq), ctypes.byref(cq), ctypes.byref(nn), ctypes.byref(ilo),
      ctypes.byref(ihi), _ptr(A), ctypes.byref(ld), _ptr(B), ctypes.byref(ld),
      _ptr(Q), ctypes.byref(ld), _ptr(Z), ctypes.byref(ld), ctypes.byref(info),
      ctypes.c_size_t(1), ctypes.c_size_t(1))
    if info.value != 0:
        raise RuntimeError(f"dgghrd info={info.value}")
    return A, B, Q[:n, :n], Z[:n, :n]


def dhgeqz(H0, T0, schur=False, Q0=None, Z0=None):
    """LAPACK DHGEQZ(JOB='E' or 'S', COMPQ/COMPZ='N' or 'V', ILO=1, IHI=n) on
    copies of a Hessenberg-triangular pencil; returns (H, T, Q, Z, alphar,
    alphai, beta, info).  The library routine the GPU QZ iteration takes the
    place of (the reference's driver stops before it,
    src/profilereduction.f90:152-162)."""
    L = _lapack()
    f = L.scipy_dhgeqz_
    n = H0.shape[0]
    H = np.array(H0, dtype=np.float64, order="F")
    T = np.array(T0, dtype=np.float64, order="F")
    vq, vz = schur and Q0 is not None, schur and Z0 is not None
    Q = np.array(Q0, dtype=np.float64, order="F") if vq else np.zeros((1, 1), order="F")
    Z = np.array(Z0, dtype=np.float64, order="F") if vz else np.zeros((1, 1), order="F")
    ar, ai, be = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
    ci = ctypes.c_int
    job = ctypes.c_char(b"S" if schur else b"E")
    cq = ctypes.c_char(b"V" if vq else b"N")
    cz = ctypes.c_char(b"V" if vz else b"N")
    nn, ilo, ihi, ld = ci(n), ci(1), ci(n), ci(max(1, n))
    ldq, ldz = ci(max(1, n) if vq else 1), ci(max(1, n) if vz else 1)
    info = ci(0)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    tail = (ctypes.c_size_t(1), ctypes.c_size_t(1), ctypes.c_size_t(1))
    for lwork in (-1, None):
        if lwork is None:
            lw = max(1, n, int(work[0]))
            work = np.zeros(lw)
            lwv = ci(lw)
        else:
            work = np.zeros(1)
            lwv = ci(-1)
        f(ctypes.byref(job), ctypes.byref(cq), ctypes.byref(cz), ctypes.byref(nn),
          ctypes.byref(ilo), ctypes.byref(ihi), _ptr(H), ctypes.byref(ld), _ptr(T),
          ctypes.byref(ld), dp(ar), dp(ai), dp(be), _ptr(Q), ctypes.byref(ldq), _ptr(Z),
          ctypes.byref(ldz), dp(work), ctypes.byref(lwv), ctypes.byref(info), *tail)
    return H, T, (Q if vq else None), (Z if vz else None), ar[:n], ai[:n], be[:n], info.value


def qr_b(A0, B0, Q0=None):
    """QR of a general B through the host's LAPACK (scipy): B0 = Q_B R with
    DGEQRF; returns (Q_B^T A0, R with exact zeros below the diagonal,
    Q0 Q_B).  The reference's driver keeps only R (its A is random:
    src/profilereduction.f90:117-123); carrying Q_B into A and Q keeps
    (A, B) equivalent to the input pencil."""
    from scipy.linalg import lapack
    n = B0.shape[0]
    A = np.array(A0, dtype=np.float64, order="F")
    Q = np.eye(n, order="F") if Q0 is None else np.array(Q0, dtype=np.float64, order="F")
    if n == 0:
        return A, np.array(B0, dtype=np.float64, order="F"), Q
    qr, tau, _, info = lapack.dgeqrf(np.array(B0, dtype=np.float64, order="F"))
    assert info == 0
    R = np.triu(qr)
    A, _, info = lapack.dormqr("L", "T", qr, tau, A, max(1, n))
    assert info == 0
    Q, _, info = lapack.dormqr("R", "N", qr, tau, Q, max(1, n))
    assert info == 0
    return np.asfortranarray(A), np.asfortranarray(R), np.asfortranarray(Q)


def dgghd3_mkl(A0, B0):
    """The same DGGHD3 call through Intel MKL (libmkl_rt), when the host has
    it (the reference builds against MKL, make.inc.example:12-16); None
    otherwise.  Used only as a timed CPU baseline."""
    cands = [os.path.join(p, "libmkl_rt.so") for p in
             (os.environ.get("MKLROOT", ""), "/opt/conda/lib")]
    path = next((c for c in cands if c and os.path.exists(c)), None)
    if path is None:
        return None
    f = ctypes.CDLL(path).dgghd3_
    n = A0.shape[0]
    A = np.array(A0, dtype=np.float64, order="F")
    B = np.array(B0, dtype=np.float64, order="F")
    Q = np.zeros((n, n), order="F")
    Z = np.zeros((n, n), order="F")
    ci = ctypes.c_int
    nn, ilo, ihi, ld, info = ci(n), ci(1), ci(n), ci(max(1, n)), ci(0)
    cq = ctypes.c_char(b"I")
    wq = np.zeros(1)
    for lwork in (ci(-1), None):
        work = wq if lwork is not None else np.zeros(max(1, int(wq[0])))
        lw = lwork if lwork is not None else ci(work.size)
        f(ctypes.byref(cq), ctypes.byref(cq), ctypes.byref(nn), ctypes.byref(ilo),
          ctypes.byref(ihi), _ptr(A), ctypes.byref(ld), _ptr(B), ctypes.byref(ld),
          _ptr(Q), ctypes.byref(ld), _ptr(Z), ctypes.byref(ld),
          work.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(lw),
          ctypes.byref(info))
    if info.value != 0:
        raise RuntimeError(f"mkl dgghd3 info={info.value}")
    return A, B, Q, Z


# --------------------------------------------------------------------------
# Metrics
# --------------------------------------------------------------------------
def backward_errors(A0, B0, H, T, Q, Z):
    """||Q^T A0 Z - H||_F/||A0||_F and the same for B (profilereduction.f90:164-174)."""
    ea = np.linalg.norm(Q.T @ A0 @ Z - H) / np.linalg.norm(A0)
    eb = np.linalg.norm(Q.T @ B0 @ Z - T) / np.linalg.norm(B0)
    return float(ea), float(eb)


def orthogonality(Q):
    n = Q.shape[0]
    return float(np.linalg.norm(Q.T @ Q - np.eye(n)))


def below_band(M, band):
    """max |M(i,j)| over i > j + band (band=1: Hessenberg, 0: triangular)."""
    n = M.shape[0]
    mask = np.tril(np.ones((n, n), dtype=bool), -band - 1)
    return float(np.max(np.abs(M[mask]))) if mask.any() else 0.0


def sign_normalize(H, T, Q, Z, band=1):
    """Fix the diagonal-sign freedom of an (r-)Hessenberg-triangular form with
    Q(:,1:band) fixed: choose D, E = diag(+-1) so that T(j,j) >= 0 and
    H(j+band, j) >= 0; returns D H E, D T E, Q D, Z E."""
    n = H.shape[0]
    d = np.ones(n)
    e = np.ones(n)
    for j in range(n):
        t = T[j, j]
        if t == 0.0:
            t = H[j, j] if H[j, j] != 0.0 else 1.0
        e[j] = d[j] * (1.0 if t > 0 else -1.0)
        if j + band < n:
            h = H[j + band, j]
            d[j + band] = e[j] * (1.0 if h >= 0 else -1.0)
    return (d[:, None] * H * e[None, :], d[:, None] * T * e[None, :],
            Q * d[None, :], Z * e[None, :])


def chordal_eig_distance(A0, B0, H, T):
    """Largest chordal distance between matched generalized eigenvalues of
    (A0,B0) and (H,T) (infinite eigenvalues included)."""
    import scipy.linalg as sl
    from scipy.optimize import linear_sum_assignment

    w1 = sl.eigvals(A0, B0, homogeneous_eigvals=True)
    w2 = sl.eigvals(H, T, homogeneous_eigvals=True)

    def unit(w):
        a, b = w
        nrm = np.sqrt(np.abs(a) ** 2 + np.abs(b) ** 2)
        nrm[nrm == 0] = 1.0
        return a / nrm, b / nrm

    a1, b1 = unit(w1)
    a2, b2 = unit(w2)
    # chordal distance |a1 b2 - b1 a2| for unit-normalised pairs
    D = np.abs(a1[:, None] * b2[None, :] - b1[:, None] * a2[None, :])
    ri, ci = linear_sum_assignment(D)
    return float(D[ri, ci].max()) if len(ri) else 0.0
