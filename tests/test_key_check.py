"""The shading pass's key checks (shade_key_fetch, master-of-puppets_amd/
csrc/shade_key.h -- the code k_shade runs before it shades a pixel) compiled
for the host with AddressSanitizer and fed random and stale visibility keys
against tables of garbage (tests/key_check.cpp): no read leaves its table,
and every accepted key hands in-range draw and attribute slots to the
shading.  This is the check of the validation added after a timing variant
that removed the visibility reset faulted on stale keys; the variant itself
is never run."""
import os
import shutil
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_random_and_stale_keys_stay_in_bounds():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "key_check")
        subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g",
                        "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=address",
                        "-I", os.path.join(REPO, "include"),
                        "-I", os.path.join(REPO, "master-of-puppets_amd", "csrc"),
                        os.path.join(REPO, "tests", "key_check.cpp"), "-o", exe],
                       check=True, capture_output=True)
        r = subprocess.run([exe, "20000000"], capture_output=True, text=True)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert "AddressSanitizer" not in r.stderr
    assert "inconsistent 0" in r.stdout
