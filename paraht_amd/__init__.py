"""paraht_amd -- ParaHT's two-stage Hessenberg-triangular reduction on MI355X.

Host-side mirror of the reference's Fortran interface, over the C-ABI of
libparaht_hip.so (include/paraht.h).  The routines keep the reference's names,
argument lists and in-place semantics:

    ab2rht_simple(n, A, lda, B, ldb, Q, ldq, Z, ldz, nb, p, time_components)
        src/ab2rht.f90:8      (A,B) -> r-Hessenberg-triangular, Q<-Q*Q1, Z<-Z*Z1
    rht2ht_householder(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, time_components)
        src/rht2ht.f90:1346   r-Hessenberg-triangular -> Hessenberg-triangular
    rht2ht_givens(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq)
        src/rht2ht.f90:696
    rht2ht_householder_unblocked(n, A, lda, B, ldb, Q, ldq, Z, ldz, r)
        src/rht2ht.f90:2053

Arrays are either NumPy float64 arrays in Fortran order of shape (ld, >=n)
(copied to the GPU and back, like the Fortran drop-in), or torch float64 CUDA
tensors with column-major layout (stride(0) == 1; the work stays in HBM).
An illegal argument raises ValueError naming its position in the reference's
argument list; device failures raise ParahtError.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PARAHT_STAGE2_GIVENS, PARAHT_TIME_COMPONENTS, ParahtError

__all__ = [
    "ab2rht_simple", "rht2ht_householder", "rht2ht_givens",
    "rht2ht_householder_unblocked", "reduce", "qr_b", "reduce_general", "Context",
    "ParahtError",
    "build", "column_major", "batched", "adjoint", "qz", "hgeqz",
]

build = _lib.build


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def column_major(t):
    """A column-major copy of a 2-D torch tensor (stride(0) == 1)."""
    return t.t().contiguous().t()


def _np_ptr(a: np.ndarray, name: str):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64
            and a.flags.f_contiguous):
        raise TypeError(f"{name} must be a Fortran-ordered float64 ndarray")
    return ctypes.c_void_p(a.ctypes.data)


def _dev_ld(t, name: str) -> int:
    import torch
    if t.dtype != torch.float64 or not t.is_cuda or t.dim() != 2:
        raise TypeError(f"{name} must be a 2-D float64 CUDA tensor")
    if t.stride(0) != 1:
        raise TypeError(f"{name} must be column-major (stride(0) == 1)")
    return t.stride(1)


class Context:
    """A device context (stream + workspaces) for the device-pointer API."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().paraht_ctx_create(ctypes.byref(self._h), device),
                   "paraht_ctx_create")
        self.device = device

    def set_stream(self, stream) -> None:
        """Run on a torch/HIP stream (its raw handle) instead of our own."""
        raw = getattr(stream, "cuda_stream", stream)
        _lib.check(_lib.lib().paraht_ctx_set_stream(self._h, ctypes.c_void_p(raw)),
                   "paraht_ctx_set_stream")

    def last_times(self):
        buf = (ctypes.c_double * 5)()
        k = _lib.lib().paraht_ctx_last_times(self._h, buf, 5)
        keys = ["stage1_ms", "stage1_genleft_ms", "stage1_genright_ms",
                "stage2_ms", "stage2_gen_ms"]
        return dict(zip(keys, list(buf)[:k]))

    KERNELS = ("wy_apply", "s1_panel_qr", "s1_rq_gen", "s2_gen", "s2_rows",
               "s2_cols", "other")

    def profile(self, enable: bool = True) -> None:
        """Bracket every kernel launch with HIP events (resets the sums)."""
        _lib.check(_lib.lib().paraht_ctx_profile(self._h, int(enable)),
                   "paraht_ctx_profile")

    def kernel_stats(self):
        """{kernel: (total_ms, launches, algorithmic_flops, algorithmic_bytes)}"""
        out = {}
        for k, name in enumerate(self.KERNELS):
            ms, cnt = ctypes.c_double(), ctypes.c_longlong()
            fl, by = ctypes.c_double(), ctypes.c_double()
            _lib.check(_lib.lib().paraht_ctx_kernel_stats(
                self._h, k, ctypes.byref(ms), ctypes.byref(cnt),
                ctypes.byref(fl), ctypes.byref(by)), "paraht_ctx_kernel_stats")
            out[name] = (ms.value, cnt.value, fl.value, by.value)
        return out

    def close(self) -> None:
        if self._h:
            _lib.lib().paraht_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT_CTX = {}


def _ctx_for(t) -> Context:
    dev = t.device.index if t.device.index is not None else 0
    if dev not in _DEFAULT_CTX:
        _DEFAULT_CTX[dev] = Context(dev)
    return _DEFAULT_CTX[dev]


def _torch_ready(t) -> None:
    """Waits for torch's queued work on t's device: the library runs on its
    context's own non-blocking stream, which is not ordered with torch's."""
    import torch
    torch.cuda.current_stream(t.device).synchronize()


def _check_shapes(n, mats, names, what):
    """Every matrix must hold an n x n block: at least n rows (the leading
    dimension is checked separately) and n columns, or the kernels would
    read and write past its storage."""
    if not isinstance(n, (int, np.integer)) or n < 0:
        raise ValueError(f"{what}: illegal value of argument 1 (n={n})")
    for (m, _), nm in zip(mats, names):
        shape = tuple(m.shape)
        if len(shape) != 2 or shape[0] < n or shape[1] < n:
            raise ValueError(f"{what}: {nm} has shape {shape}, smaller than n x n = {n} x {n}")


def _dispatch(host_name, dev_name, what, n, mats, ints, flags, ctx=None):
    L = _lib.lib()
    names = ("A", "B", "Q", "Z")
    _check_shapes(n, mats, names, what)
    if all(_is_torch(m) for m, _ in mats):
        lds = []
        for (m, ld), nm in zip(mats, names):
            sld = _dev_ld(m, nm)
            if ld is not None and ld != sld:
                raise ValueError(f"ld{nm.lower()}={ld} does not match the tensor's stride {sld}")
            lds.append(sld)
        c = ctx or _ctx_for(mats[0][0])
        _torch_ready(mats[0][0])
        args = [c._h, n]
        for (m, _), ld in zip(mats, lds):
            args += [ctypes.c_void_p(m.data_ptr()), ld]
        st = getattr(L, dev_name)(*args, *ints, *([] if flags is None else [flags]))
        _lib.check(st, what)
        return
    args = [n]
    for pos, ((m, ld), nm) in enumerate(zip(mats, names)):
        if ld is not None and ld != m.shape[0]:
            # argument 3, 5, 7, 9 of the reference's list (lda, ldb, ldq, ldz)
            raise ValueError(f"{what}: illegal value of argument {3 + 2 * pos} "
                             f"(ld{nm.lower()}={ld}, the array's leading dimension is "
                             f"{m.shape[0]})")
        args += [_np_ptr(m, nm), m.shape[0]]
    st = getattr(L, host_name)(*args, *ints, flags) if flags is not None else \
        getattr(L, host_name)(*args, *ints)
    _lib.check(st, what)


def ab2rht_simple(n, A, lda, B, ldb, Q, ldq, Z, ldz, nb, p,
                  time_components=False, ctx=None):
    """Stage 1 in place (src/ab2rht.f90:8-13)."""
    flags = PARAHT_TIME_COMPONENTS if time_components else 0
    _dispatch("paraht_ab2rht", "paraht_ab2rht_dev", "ab2rht_simple", n,
              [(A, lda), (B, ldb), (Q, ldq), (Z, ldz)], [nb, p], flags, ctx)


def rht2ht_householder(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq,
                       time_components=False, ctx=None):
    """Stage 2 in place, the reference driver's entry point
    (src/rht2ht.f90:1346-1351)."""
    flags = PARAHT_TIME_COMPONENTS if time_components else 0
    _dispatch("paraht_rht2ht", "paraht_rht2ht_dev", "rht2ht_householder", n,
              [(A, lda), (B, ldb), (Q, ldq), (Z, ldz)], [r, nq], flags, ctx)


def rht2ht_givens(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, ctx=None):
    """Stage 2 in place with Givens rotations (src/rht2ht.f90:696-700)."""
    _dispatch("paraht_rht2ht", "paraht_rht2ht_dev", "rht2ht_givens", n,
              [(A, lda), (B, ldb), (Q, ldq), (Z, ldz)], [r, nq],
              PARAHT_STAGE2_GIVENS, ctx)


def rht2ht_householder_unblocked(n, A, lda, B, ldb, Q, ldq, Z, ldz, r):
    """Unblocked stage 2 (src/rht2ht.f90:2053-2057); host arrays only."""
    _dispatch("paraht_rht2ht_unblocked", None, "rht2ht_householder_unblocked",
              n, [(A, lda), (B, ldb), (Q, ldq), (Z, ldz)], [r], None)


def reduce(A, B, r, p, nq, Q=None, Z=None, time_components=False, ctx=None, givens=False):
    """Both stages (src/profilereduction.f90:136,144).  Returns (H, T, Q, Z).

    Stage 2 is rht2ht_householder, the reference driver's variant, or with
    givens=True rht2ht_givens.  NumPy inputs are not modified (copies are
    reduced); torch CUDA inputs (column-major) are reduced in place and
    returned."""
    flags = PARAHT_TIME_COMPONENTS if time_components else 0
    if givens:
        flags |= PARAHT_STAGE2_GIVENS
    n = A.shape[0]
    for M, nm in ((A, "A"), (B, "B"), (Q, "Q"), (Z, "Z")):
        if M is not None and tuple(M.shape) != (n, n):
            raise ValueError(f"reduce: {nm} has shape {tuple(M.shape)}, expected ({n}, {n})")
    if _is_torch(A):
        import torch
        if Q is None:
            Q = column_major(torch.eye(n, dtype=torch.float64, device=A.device))
        if Z is None:
            Z = column_major(torch.eye(n, dtype=torch.float64, device=A.device))
        _dispatch(None, "paraht_reduce_dev", "reduce", n,
                  [(A, None), (B, None), (Q, None), (Z, None)], [r, p, nq],
                  flags, ctx)
        return A, B, Q, Z
    H = np.array(A, dtype=np.float64, order="F", copy=True)
    T = np.array(B, dtype=np.float64, order="F", copy=True)
    Q = np.eye(n, order="F") if Q is None else np.array(Q, dtype=np.float64, order="F", copy=True)
    Z = np.eye(n, order="F") if Z is None else np.array(Z, dtype=np.float64, order="F", copy=True)
    _dispatch("paraht_reduce", None, "reduce", n,
              [(H, n), (T, n), (Q, n), (Z, n)], [r, p, nq], flags)
    return H, T, Q, Z


def qr_b(n, A, lda, B, ldb, Q, ldq, nb=64, p=2, ctx=None):
    """B <- R = Q_B^T B (upper triangular), A <- Q_B^T A, Q <- Q Q_B, in
    place, on the GPU: the step the reference's driver does with LAPACK
    before the reduction (src/profilereduction.f90:117-123).  nb columns per
    panel, panel blocks of p*nb rows."""
    _dispatch("paraht_qr_b", "paraht_qr_b_dev", "qr_b", n,
              [(A, lda), (B, ldb), (Q, ldq)], [nb, p], None, ctx)


def reduce_general(A, B, r, p, nq, Q=None, Z=None, qr_nb=64, qr_p=2,
                   time_components=False, ctx=None, givens=False):
    """Reduction of a pencil with a general (not triangular) B: the QR of B
    (qr_b), then both stages.  Returns (H, T, Q, Z) with Q^T A0 Z = H and
    Q^T B0 Z = T for the input pencil (A0, B0) and identity Q, Z; NumPy
    inputs are copied, torch CUDA inputs (column-major) reduced in place."""
    n = A.shape[0]
    for M, nm in ((A, "A"), (B, "B"), (Q, "Q"), (Z, "Z")):
        if M is not None and tuple(M.shape) != (n, n):
            raise ValueError(f"reduce_general: {nm} has shape {tuple(M.shape)}, "
                             f"expected ({n}, {n})")
    if _is_torch(A):
        import torch
        if Q is None:
            Q = column_major(torch.eye(n, dtype=torch.float64, device=A.device))
        qr_b(n, A, None, B, None, Q, None, qr_nb, qr_p, ctx=ctx)
        return reduce(A, B, r, p, nq, Q, Z, time_components, ctx, givens)
    H = np.array(A, dtype=np.float64, order="F", copy=True)
    T = np.array(B, dtype=np.float64, order="F", copy=True)
    Q = np.eye(n, order="F") if Q is None else np.array(Q, dtype=np.float64, order="F", copy=True)
    qr_b(n, H, n, T, n, Q, n, qr_nb, qr_p)
    return reduce(H, T, r, p, nq, Q, Z, time_components, givens=givens)


from . import batched  # noqa: E402  (batched reduction of many small pencils)
from . import adjoint  # noqa: E402  (reverse mode: ht_backward, ht_reduce)
from . import qz  # noqa: E402  (QZ iteration on the reduced pencil)
from .qz import hgeqz  # noqa: E402
