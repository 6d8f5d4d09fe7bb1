"""ctypes binding of libparaht_hip.so (the C-ABI declared in include/paraht.h).

The library is built in-tree by ``paraht_amd/csrc/Makefile`` (``build()``);
loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparaht_hip.so")
_lib = None

PARAHT_TIME_COMPONENTS = 0x1
PARAHT_STAGE2_GIVENS = 0x2
PARAHT_BATCH_GRAPH = 0x4
PARAHT_QZ_SCHUR = 0x8


class Pencil(ctypes.Structure):
    """paraht_pencil of include/paraht.h: one pencil of a batch."""
    _fields_ = [("offset", ctypes.c_longlong), ("n", ctypes.c_int), ("r", ctypes.c_int),
                ("p", ctypes.c_int), ("nq", ctypes.c_int)]

# (name, argtypes) of every entry point of include/paraht.h
_P = ctypes.c_void_p
_I = ctypes.c_int
_DP = ctypes.c_void_p
_SIG_HOST12 = [_I, _DP, _I, _DP, _I, _DP, _I, _DP, _I, _I, _I, _I]
SIGNATURES = {
    "paraht_ctx_create": ([ctypes.POINTER(_P), _I], _I),
    "paraht_ctx_destroy": ([_P], _I),
    "paraht_ctx_set_stream": ([_P, _P], _I),
    "paraht_ctx_last_times": ([_P, ctypes.POINTER(ctypes.c_double), _I], _I),
    "paraht_ctx_profile": ([_P, _I], _I),
    "paraht_ctx_kernel_stats": ([_P, _I, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double)], _I),
    "paraht_ab2rht": (_SIG_HOST12, _I),
    "paraht_rht2ht": (_SIG_HOST12, _I),
    "paraht_rht2ht_unblocked": (_SIG_HOST12[:10], _I),
    "paraht_reduce": (_SIG_HOST12 + [_I], _I),
    "paraht_ab2rht_dev": ([_P] + _SIG_HOST12, _I),
    "paraht_rht2ht_dev": ([_P] + _SIG_HOST12, _I),
    "paraht_reduce_dev": ([_P] + _SIG_HOST12 + [_I], _I),
    "paraht_reduce_dev_noalloc": ([_P] + _SIG_HOST12 + [_I], _I),
    "paraht_qr_b": (_SIG_HOST12[:7] + [_I, _I], _I),
    "paraht_qr_b_dev": ([_P] + _SIG_HOST12[:7] + [_I, _I], _I),
    "paraht_batched_dev": ([_P, _I, ctypes.POINTER(Pencil), _DP, _DP, _DP, _DP, _I], _I),
    "paraht_batched_defaults": ([_I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                 ctypes.POINTER(_I)], _I),
    "paraht_adjoint_dev": ([_P, _I, ctypes.POINTER(Pencil), _DP, _DP, _DP, _DP, _DP, _DP, _DP],
                           _I),
    "paraht_hgeqz": (_SIG_HOST12[:9] + [_DP, _DP, _DP, _I, ctypes.POINTER(_I)], _I),
    "paraht_hgeqz_dev": ([_P] + _SIG_HOST12[:9] + [_DP, _DP, _DP, _I, ctypes.POINTER(_I)], _I),
    "paraht_hgeqz_batched_dev": ([_P, _I, ctypes.POINTER(Pencil), _DP, _DP, _DP, _DP, _DP, _DP,
                                  _DP, _I, ctypes.POINTER(_I)], _I),
    "paraht_version": ([], _I),
    "paraht_status_string": ([_I], ctypes.c_char_p),
}


def build(verbose: bool = False) -> str:
    """Compile libparaht_hip.so for gfx950 with hipcc: every source, every
    time (make -B), so a call checks that the tree builds, not only that a
    library exists."""
    cmd = ["make", "-s", "-B", "-j8", "-C", os.path.join(_HERE, "csrc")]
    subprocess.run(cmd, check=True,
                   stdout=None if verbose else subprocess.DEVNULL)
    return LIB_PATH


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # PyTorch-ROCm ships its own copy of the HIP runtime (same soname
        # libamdhip64.so.7).  Loading torch first lets this library bind to
        # that copy; loading ours first would put two HIP runtimes in the
        # process and torch would then find no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with paraht_amd._lib.build() "
                "(hipcc, gfx950). There is no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class ParahtError(RuntimeError):
    pass


def check(status: int, what: str) -> None:
    if status == 0:
        return
    if status < 0:
        raise ValueError(f"{what}: illegal value of argument {-status}")
    msg = lib().paraht_status_string(status).decode()
    raise ParahtError(f"{what}: {msg} (status {status})")
