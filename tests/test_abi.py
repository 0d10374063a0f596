"""The C-ABI boundary: mirror-type layouts equal the reference's, the library
exports every declared entry point, and it refuses to run without a GPU
(no CPU fallback)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import mop_amd

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "rhi")),
                    reason="reference headers not present")
def test_layout_matches_reference_headers():
    # static asserts over sizeof/offsetof of every mirrored type and table slot
    r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", f"-I{REF}/include", f"-I{REF}/src",
                        f"-I{REPO}/include", os.path.join(REPO, "tests", "abi_check.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def declared_symbols():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(REPO, "include")) if f.endswith(".h")):
        src = open(os.path.join(REPO, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(mop_(?:hip|rhi)_\w+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = mop_amd.lib()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) > 60


def test_rhi_table_is_filled():
    from mop_amd import abi
    lib = mop_amd.lib()
    tbl = lib.mop_rhi_backend_hip()
    assert tbl
    name = C.cast(C.c_void_p.from_address(tbl).value, C.c_char_p).value
    assert name == b"hip"
    # mandatory slots of rhi.h:115-121 (device_create .. buffer_read) non-NULL
    mandatory = [1, 2, 3, 4, 5, 6, 7, 9, 10, 13, 14, 15, 16, 17, 18, 19, 22, 24, 25]
    ptrs = (C.c_void_p * 46).from_address(tbl)
    assert all(ptrs[i] for i in mandatory)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_gpu_means_no_device():
    lib = mop_amd.lib()
    assert not lib.mop_hip_device_create()
    assert b"device" in lib.mop_hip_last_error()
