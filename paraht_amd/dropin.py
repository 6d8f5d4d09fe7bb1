"""Build and locate the Fortran drop-in (modules ab2rht and rht2ht over the
C-ABI, and the profilereduction driver) in paraht_amd/fortran."""
from __future__ import annotations

import os
import shutil
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
FORTRAN_DIR = os.path.join(_HERE, "fortran")
BUILD_DIR = os.path.join(FORTRAN_DIR, "_build")
DRIVER = os.path.join(BUILD_DIR, "profilereduction")
MODULE_OBJECTS = {  # object -> module procedures it must define (Flang mangling)
    "ab2rht.o": ["_QMab2rhtPab2rht_simple"],
    "rht2ht.o": ["_QMrht2htPrht2ht_householder", "_QMrht2htPrht2ht_givens",
                 "_QMrht2htPrht2ht_householder_unblocked"],
}


def compiler() -> str | None:
    """amdflang if it is on PATH (or in /opt/rocm/bin), else None."""
    return shutil.which("amdflang") or (
        "/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


def build(verbose: bool = False) -> str:
    """Compile the modules and the driver with amdflang (after the HIP library)."""
    fc = compiler()
    if fc is None:
        raise RuntimeError("amdflang not found: the Fortran drop-in cannot be built")
    subprocess.run(["make", "-s", "-C", FORTRAN_DIR, f"FC={fc}"], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)
    return DRIVER
