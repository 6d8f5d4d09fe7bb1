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
