"""Python host side of the MI355X HIP backend for Master of Puppets' frame path.

The product is the C-ABI library ``libmop_hip.so`` built in-tree from
``csrc/`` (gfx950 kernels, HIP runtime, C11 RHI table).  This package only
loads it with ctypes and offers the scene generators and the frame driver
used by the tests and the benchmark.  There is no CPU fallback: ``lib()``
raises if the library is missing or cannot be loaded.
"""
import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libmop_hip.so")

_lib = None


def build(verbose=False):
    """Compile libmop_hip.so for gfx950 (always recompiles)."""
    csrc = os.path.join(PKG_DIR, "csrc")
    subprocess.run(["make", "-C", csrc, "clean"], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)
    subprocess.run(["make", "-C", csrc, "-j8"], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)


def lib():
    """The loaded C-ABI library; raises RuntimeError when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        from . import abi
        abi.declare(_lib)
    return _lib
