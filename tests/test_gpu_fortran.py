"""The Fortran drop-in driver on the GPU: algorithm 1 runs both stages through
the ab2rht / rht2ht modules and the C-ABI, and reports backward errors within
10 n u in the reference's result line (src/profilereduction.f90:176-185)."""
import re
import subprocess

import pytest

from paraht_amd import dropin

pytestmark = pytest.mark.gpu

LINE = re.compile(r"n:\s*(\d+) ;runtime:\s*([-\d.]+) ;error:\s*(\S+)\s+(\S+) ;phase 1 time:\s*([-\d.]+)"
                  r" ;phase 2 time:\s*([-\d.]+) ;r:\s*(\d+) ;p:\s*(\d+) ;nq:\s*(\d+)")


@pytest.mark.parametrize("n,r,p,nq", [(256, 16, 4, 8), (300, 8, 2, 4)])
def test_driver_paraht(n, r, p, nq):
    print(f"fortran driver n={n} r={r} p={p} nq={nq}", flush=True)
    out = subprocess.run([dropin.DRIVER, str(n), "1", str(r), str(p), str(nq)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    m = LINE.search(out.stdout)
    assert m, out.stdout
    assert (int(m.group(1)), int(m.group(7)), int(m.group(8)), int(m.group(9))) == (n, r, p, nq)
    u = 2.0 ** -53
    assert float(m.group(3)) <= 10 * n * u and float(m.group(4)) <= 10 * n * u, out.stdout


def test_driver_time_components():
    out = subprocess.run([dropin.DRIVER, "200", "1", "16", "2", "8", "1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "genlefttime" in out.stdout and "gen:" in out.stdout, out.stdout
    assert LINE.search(out.stdout), out.stdout


def test_driver_paraht_on_a_pencil_file(tmp_path):
    """Algorithm 1 (both GPU stages through the modules) on an external pencil
    read from a file, the reduced pencil written back and checked."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
    import oracle as O
    from paraht_amd.pencils import illcond_pencil, load_reduced, save_pencil
    n, r, p, nq = 200, 16, 2, 8
    A0, B0 = illcond_pencil(n, seed=8)
    f, g = tmp_path / "in.pht", tmp_path / "out.pht"
    save_pencil(f, A0, B0)
    out = subprocess.run([dropin.DRIVER, "0", "1", str(r), str(p), str(nq), f"pencil={f}",
                          f"out={g}"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    m = LINE.search(out.stdout)
    assert m and int(m.group(1)) == n, out.stdout
    H, T, Q, Z = load_reduced(g)
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    u = 2.0 ** -53
    assert ea <= 10 * n * u and eb <= 10 * n * u, (ea, eb)
    assert O.below_band(H, 1) == 0.0 and O.below_band(T, 0) == 0.0


def test_driver_paraht_general_b_from_a_file(tmp_path):
    """Algorithm 1 on a pencil file with a general B: the QR of B on the GPU
    (paraht_qr_b through the module's binding) and both stages, checked
    against the pencil in the file."""
    import sys
    import os
    import numpy as np
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
    import oracle as O
    from paraht_amd.pencils import load_reduced, save_pencil
    n, r, p, nq = 300, 16, 4, 8
    print(f"fortran driver general B n={n}", flush=True)
    rng = np.random.default_rng(21)
    A0, B0 = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    f, g = tmp_path / "gen.pht", tmp_path / "gen_out.pht"
    save_pencil(f, A0, B0)
    out = subprocess.run([dropin.DRIVER, "0", "1", str(r), str(p), str(nq), f"pencil={f}",
                          f"out={g}"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    m = LINE.search(out.stdout)
    assert m and int(m.group(1)) == n, out.stdout
    H, T, Q, Z = load_reduced(g)
    ea, eb = O.backward_errors(A0, B0, H, T, Q, Z)
    u = 2.0 ** -53
    assert ea <= 10 * n * u and eb <= 10 * n * u, (ea, eb)
    assert O.below_band(H, 1) == 0.0 and O.below_band(T, 0) == 0.0
    assert O.chordal_eig_distance(A0, B0, H, T) <= 1e-6


def test_driver_qz_after_the_reduction():
    """qz=1 after algorithm 1: the GPU QZ iteration (paraht_hgeqz through the
    Fortran binding) on the reduced pencil; every eigenvalue lies within 1e-11
    (chordal) of LAPACK DHGEQZ's on the same pencil."""
    print("fortran driver qz=1 n=200", flush=True)
    out = subprocess.run([dropin.DRIVER, "200", "1", "16", "4", "8", "qz=1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert LINE.search(out.stdout), out.stdout
    m = re.search(r"qz: time:\s*([-\d.]+) ;info:\s*(\d+) ;infinite:\s*(\d+) ;chordal vs "
                  r"dhgeqz:\s*(\S+)", out.stdout)
    assert m and int(m.group(2)) == 0, out.stdout
    assert float(m.group(4)) <= 1e-11, out.stdout
