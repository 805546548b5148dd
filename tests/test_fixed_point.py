"""Host-side checks of the engine's fixed-point helpers (hacc_gpm_amd/csrc/common.h,
compiled for the host with hipcc): the 32-bit drift the gather uses on power-of-two
grids, fx_drift_p2, gives the bits of the 64-bit form fx_drift (the one the oracle's
orc_drift_split restates, pinned on the GPU by the kick/drift parity tests) on random
and boundary inputs: moves of up to four boxes, exactly ng/2, ng and 2 ng, past 512
cells, the box seam and tiny moves."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_drift_p2_matches_drift(tmp_path):
    exe = tmp_path / "fixed_point_check"
    r = subprocess.run([HIPCC, "-O2", "-x", "hip", "--offload-arch=gfx950", "--cuda-host-only",
                        "-I", os.path.join(ROOT, "hacc_gpm_amd", "csrc"),
                        os.path.join(ROOT, "tests", "host", "fixed_point_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, check=True).stdout
    rows = [tuple(int(v) for v in line.split()) for line in out.splitlines() if line.strip()]
    assert [ng for ng, _, _ in rows] == [16, 32, 64, 128, 256, 512, 1024]
    for ng, cases, bad in rows:
        assert cases > 2_000_000 and bad == 0, (ng, cases, bad)
