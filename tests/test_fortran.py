"""The Fortran drop-in on the CPU: the modules define the reference's
procedure names (src/ab2rht.f90:8, src/rht2ht.f90:696,1346,2053 as mangled by
Flang), the driver keeps the reference's command line and result line
(src/profilereduction.f90:25-69,176-185), and the GPU path fails loudly when
no device is present."""
import os
import re
import subprocess

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
