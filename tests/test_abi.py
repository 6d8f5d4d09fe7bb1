"""The C-ABI library loads without a GPU and exports every entry point that
include/paraht.h declares (no compute calls: this runs on the CPU)."""
import os
import re

from paraht_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "paraht.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(paraht_\w+)\s*\(", text)))


def test_header_declares_the_reference_boundary():
    names = declared_symbols()
    for must in ("paraht_ab2rht", "paraht_rht2ht", "paraht_rht2ht_unblocked",
                 "paraht_reduce", "paraht_ab2rht_dev", "paraht_rht2ht_dev"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_version_and_status_strings():
    L = _lib.lib()
    assert L.paraht_version() >= 100
    assert L.paraht_status_string(0) == b"success"
    assert b"illegal" in L.paraht_status_string(-3)


def test_argument_errors_are_reported_without_a_device():
    """Argument checks run before any device work (LAPACK-style -i)."""
    import numpy as np
    import paraht_amd as ph
    import pytest
    A = np.zeros((4, 4), order="F")
    with pytest.raises(ValueError, match="argument 3"):
        ph.ab2rht_simple(4, A, 2, A, 4, A, 4, A, 4, 2, 2)
    with pytest.raises(ValueError, match="argument 11"):
        ph.ab2rht_simple(4, A, 4, A, 4, A, 4, A, 4, 2, 1)
    with pytest.raises(ValueError, match="argument 10"):
        ph.rht2ht_givens(4, A, 4, A, 4, A, 4, A, 4, 0, 2)


def test_shapes_smaller_than_n_are_rejected_before_any_call():
    """A matrix with fewer than n rows or columns would let the kernels (and
    the host copies) run past its storage: the wrapper refuses it."""
    import numpy as np
    import paraht_amd as ph
    import pytest
    n = 6
    A = np.zeros((n, n), order="F")
    thin = np.zeros((n, n - 2), order="F")
    short = np.zeros((n - 1, n), order="F")
    with pytest.raises(ValueError, match="smaller than n x n"):
        ph.ab2rht_simple(n, A, n, thin, n, A, n, A, n, 2, 2)
    with pytest.raises(ValueError, match="smaller than n x n"):
        ph.rht2ht_givens(n, A, n, A, n, short, n - 1, A, n, 2, 2)
    with pytest.raises(ValueError, match="argument 7"):
        ph.rht2ht_givens(n, A, n, A, n, A, n + 3, A, n, 2, 2)
    with pytest.raises(ValueError, match="expected"):
        ph.reduce(A, thin, 2, 2, 2)
    with pytest.raises(ValueError, match="expected"):
        ph.reduce(np.zeros((n, n + 1)), A, 2, 2, 2)


def test_torch_shapes_smaller_than_n_are_rejected():
    import pytest
    import torch
    import paraht_amd as ph
    n = 5
    A = torch.zeros(n, n, dtype=torch.float64)
    with pytest.raises(ValueError, match="smaller than n x n"):
        ph._check_shapes(n, [(A, None), (torch.zeros(n, n - 1), None)], ("A", "B"), "x")
    ph._check_shapes(n, [(A, None), (torch.zeros(n + 2, n), None)], ("A", "B"), "x")


def test_qz_argument_errors_are_reported_without_a_device():
    """paraht_hgeqz checks its arguments before any device work, counting
    them like DHGEQZ's LAPACK-style -i (in the order of include/paraht.h)."""
    import ctypes
    import numpy as np
    from paraht_amd import _lib
    L = _lib.lib()
    H = np.zeros((4, 4), order="F")
    p = ctypes.c_void_p(H.ctypes.data)
    info = ctypes.c_int(7)
    ev = [ctypes.c_void_p(np.zeros(4).ctypes.data) for _ in range(3)]
    assert L.paraht_hgeqz(-1, p, 4, p, 4, None, 1, None, 1, *ev, 0, ctypes.byref(info)) == -1
    assert L.paraht_hgeqz(4, p, 3, p, 4, None, 1, None, 1, *ev, 0, ctypes.byref(info)) == -3
    assert L.paraht_hgeqz(4, p, 4, p, 2, None, 1, None, 1, *ev, 0, ctypes.byref(info)) == -5
    assert L.paraht_hgeqz(4, p, 4, p, 4, p, 2, None, 1, *ev, _lib.PARAHT_QZ_SCHUR,
                          ctypes.byref(info)) == -7
    assert L.paraht_hgeqz(4, p, 4, p, 4, None, 1, None, 1, *ev, 0x40, ctypes.byref(info)) == -13
    assert L.paraht_hgeqz(4, p, 4, p, 4, None, 1, None, 1, *ev, 0, None) == -14
    # n = 0 is a quick return with info = 0
    assert L.paraht_hgeqz(0, None, 1, None, 1, None, 1, None, 1, None, None, None, 0,
                          ctypes.byref(info)) == 0 and info.value == 0


def test_batched_size_buckets():
    """The size-bucket table: the one-stage path (r = 1) up to order 64, the
    two-stage parameters above (no device needed)."""
    from paraht_amd.batched import ONE_STAGE_MAX, defaults
    assert defaults(1)[0] == 1 and defaults(ONE_STAGE_MAX)[0] == 1
    assert defaults(ONE_STAGE_MAX + 1) == (16, 8, 16)
    assert defaults(512) == (16, 8, 16)
