"""The chunk cull's margin (master-of-puppets_amd/csrc/cull.h), checked on
the host with the vertex transform's own arithmetic: every point of a culled
box fails the plane the cull names, and no box with a corner in view is
culled.  Frames rendered with culling are compared with the reference CPU
backend by the full-size parity tests (the 4K terrain seen from inside it)."""
import os
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_culled_boxes_hold_only_rejected_points():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "cull_check")
        subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O2",
                        "-ffp-contract=off",
                        "-I", os.path.join(REPO, "include"),
                        "-I", os.path.join(REPO, "master-of-puppets_amd", "csrc"),
                        os.path.join(REPO, "tests", "cull_check.cpp"), "-o", exe],
                       check=True, capture_output=True)
        r = subprocess.run([exe, "1000000"], capture_output=True, text=True)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert " wrong 0 " in r.stdout and "visible_culled 0" in r.stdout
