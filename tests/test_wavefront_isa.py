"""The split-form far update hands window entries between the lanes of a row
through LDS: at a step's end one lane stores an entry that the next lane loads
(paraht_amd/csrc/s2_far.hpp, wf_steps).  Seen from one thread the load and the
store touch different addresses, so only a compiler barrier keeps the load
after the store.  This test compiles the four-lane instantiation that went
wrong without it (r = 16, nq = 4: the compiler merged a step's loads into
the previous step's, ahead of its stores) and checks in the gfx950 assembly
that, within every wavefront step, the loads that follow the step's stores
come after a compiler barrier (the unfixed kernels fail this in most steps).
(Compiles on the CPU; skipped without hipcc.)"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r'''
#include "s2_far.hpp"
namespace pht {
template __global__ void k_s2_rows_w2<16, 4, 4, true>(double*, int, double*, int, double*, int,
    double*, int, int, int, int, const double*, const double*, const double*, const double*,
    S2Range);
template __global__ void k_s2_rows_w2<64, 16, 2, true>(double*, int, double*, int, double*, int,
    double*, int, int, int, int, const double*, const double*, const double*, const double*,
    S2Range);
}
'''


def _steps_violating(asm):
    """Wavefront steps (the text between consecutive step-boundary markers,
    prologue and epilogue left out) in which a 64-bit LDS load (a window entry; coefficient pairs are
    128-bit) follows the step's last store before any compiler barrier, i.e.
    with nothing to keep it behind the store."""
    bad = []
    segs = asm.split("; sched_barrier")
    # the first piece holds the block prologue (tile fills, then the initial
    # loads of the first step), the last one the epilogue
    for k in range(1, len(segs) - 1):
        seg = segs[k]
        ops = re.findall(r"^\s*(ds_write\w*|ds_read2?_b64\b|;;#ASMSTART)", seg, re.M)
        writes = [i for i, o in enumerate(ops) if o.startswith("ds_write")]
        if not writes:
            continue
        after = [o for o in ops[writes[-1] + 1:] if not o.startswith("ds_write")]
        if after and after[0] != ";;#ASMSTART":
            bad.append(k)
    return bad


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_split_wavefront_loads_follow_stores(tmp_path):
    src = tmp_path / "wf.hip"
    src.write_text(SRC)
    out = tmp_path / "wf.s"
    cmd = [HIPCC, "-O3", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I", os.path.join(ROOT, "paraht_amd", "csrc"),
           "--cuda-device-only", "-S", str(src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    asm = out.read_text()
    kernels = re.split(r"^(_ZN3pht12k_s2_rows_w2\w*):", asm, flags=re.M)
    found = 0
    for name, body in zip(kernels[1::2], kernels[2::2]):
        body = body.split("s_endpgm")[0]
        assert "sched_barrier" in body, name
        assert _steps_violating(body) == [], name
        found += 1
    assert found == 2
