"""The Fortran drop-in on the CPU: the modules define the reference's
procedure names (src/ab2rht.f90:8, src/rht2ht.f90:696,1346,2053 as mangled by
Flang), the driver keeps the reference's command line and result line
(src/profilereduction.f90:25-69,176-185), and the GPU path fails loudly when
no device is present."""
import os
import re
import subprocess

import numpy as np
import pytest

from paraht_amd import dropin

pytestmark = pytest.mark.skipif(dropin.compiler() is None, reason="amdflang not installed")

LINE = re.compile(r"n:\s*(\d+) ;runtime:\s*([-\d.]+) ;error:\s*(\S+)\s+(\S+)")


@pytest.fixture(scope="module")
def driver():
    dropin.build()
    return dropin.DRIVER


def test_module_symbols(driver):
    for obj, syms in dropin.MODULE_OBJECTS.items():
        out = subprocess.run(["nm", os.path.join(dropin.BUILD_DIR, obj)], check=True,
                             capture_output=True, text=True).stdout
        defined = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
        for s in syms:
            assert s in defined, (obj, s)


def test_driver_usage(driver):
    out = subprocess.run([driver], capture_output=True, text=True, timeout=60)
    assert "usage" in out.stdout


def test_driver_lapack_path(driver):
    # algorithm 2 (host DGGHD3) has the reference's result line and small errors
    out = subprocess.run([driver, "120", "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    m = LINE.search(out.stdout)
    assert m and int(m.group(1)) == 120, out.stdout
    u = 2.0 ** -53
    assert float(m.group(3)) < 10 * 120 * u and float(m.group(4)) < 10 * 120 * u


def _gpu_visible() -> bool:
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-device error path")
def test_driver_without_device_fails_loudly(driver):
    out = subprocess.run([driver, "64", "1", "8", "2", "4"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode != 0
    assert "ab2rht_simple: paraht status 3" in out.stdout


def test_pencil_file_round_trip(tmp_path):
    """save_pencil / load_pencil (and the reduced-pencil pair) keep every bit."""
    from paraht_amd.pencils import (load_pencil, load_reduced, random_pencil, save_pencil,
                                    save_reduced)
    A, B = random_pencil(37, seed=5)
    f = tmp_path / "p.pht"
    save_pencil(f, A, B)
    A2, B2 = load_pencil(f)
    assert np.array_equal(A, A2) and np.array_equal(B, B2)
    assert A2.flags.f_contiguous
    g = tmp_path / "r.pht"
    save_reduced(g, A, B, B, A)
    assert all(np.array_equal(x, y) for x, y in zip(load_reduced(g), (A, B, B, A)))
    with pytest.raises(ValueError):
        load_reduced(f)  # a pencil is not a reduced pencil
    with pytest.raises(ValueError):
        save_pencil(f, A, B[:5, :5])


def test_driver_reads_and_writes_pencil_files(driver, tmp_path, oracle_mod):
    """The driver reduces an external pencil (pencil=FILE, n = 0 takes the
    file's order) and writes H, T, Q, Z (out=FILE): algorithm 2 (host DGGHD3)
    on the CPU, checked against the oracle's DGGHD3 after sign normalisation
    and against the pencil in the file."""
    from paraht_amd.pencils import illcond_pencil, load_reduced, save_pencil
    O = oracle_mod
    n = 90
    A0, B0 = illcond_pencil(n, seed=3, zero_frac=0.0)
    f, g = tmp_path / "in.pht", tmp_path / "out.pht"
    save_pencil(f, A0, B0)
    out = subprocess.run([driver, "0", "2", f"pencil={f}", f"out={g}"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    m = LINE.search(out.stdout)
    assert m and int(m.group(1)) == n, out.stdout
    H, T, Q, Z = load_reduced(g)
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    u = 2.0 ** -53
    assert ea <= 10 * n * u and eb <= 10 * n * u
    assert abs(ea - float(m.group(3))) <= 1e-3 * max(ea, 1e-300) + 1e-30
    assert O.below_band(H, 1) == 0.0 and O.below_band(T, 0) == 0.0
    bad = subprocess.run([driver, "91", "2", f"pencil={f}"], capture_output=True, text=True,
                         timeout=60)
    assert bad.returncode != 0 and "n differs" in (bad.stdout + bad.stderr)


def test_driver_general_b_from_a_file(driver, tmp_path, oracle_mod):
    """A pencil file whose B is not triangular: the driver's algorithm 2
    takes B's QR first (LAPACK, carried into A and Q) and reports the errors
    against the pencil in the file; the generalized eigenvalues are kept."""
    from paraht_amd.pencils import load_reduced, save_pencil
    O = oracle_mod
    n = 70
    rng = np.random.default_rng(12)
    A0, B0 = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    f, g = tmp_path / "gen.pht", tmp_path / "gen_out.pht"
    save_pencil(f, A0, B0)
    out = subprocess.run([driver, "0", "2", f"pencil={f}", f"out={g}"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    H, T, Q, Z = load_reduced(g)
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    u = 2.0 ** -53
    assert ea <= 10 * n * u and eb <= 10 * n * u, (ea, eb)
    assert O.below_band(H, 1) == 0.0 and O.below_band(T, 0) == 0.0
    assert O.chordal_eig_distance(A0, B0, H, T) <= 1e-6


QZ_LINE = re.compile(r"qz: time:\s*([-\d.]+) ;info:\s*(\d+) ;infinite:\s*(\d+)")


def test_driver_qz_lapack_path(driver):
    """qz=1 after algorithm 2: LAPACK DHGEQZ on the reduced pencil, a second
    result line with its time, info and the count of infinite eigenvalues."""
    out = subprocess.run([driver, "150", "2", "qz=1"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert LINE.search(out.stdout)
    m = QZ_LINE.search(out.stdout)
    assert m and int(m.group(2)) == 0 and int(m.group(3)) == 0, out.stdout
