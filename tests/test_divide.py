"""mop_div_recip (master-of-puppets_amd/csrc/libmf.h) -- the shading's
divisions by a constant or shared divisor, done as a multiply by the
precomputed reciprocal and one fused multiply-add correction -- returns
IEEE division's bits: the same header compiled for the host is compared on
a stride of every float numerator for each constant divisor and on random
pairs; `python tests/test_divide.py 1` runs the exhaustive sweep."""
import os
import subprocess
import sys
import tempfile

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_check(stride):
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "divide_check")
        subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fno-fast-math", "-mfma",
                        "-I", os.path.join(REPO, "master-of-puppets_amd", "csrc"), "-o", exe,
                        os.path.join(REPO, "tests", "divide_check.c"), "-lm"], check=True)
        r = subprocess.run([exe, str(stride)], capture_output=True, text=True)
    return r.returncode, r.stdout.strip()


def test_div_recip_equals_ieee_division():
    rc, out = run_check(101)
    print(out)
    assert rc == 0, out


if __name__ == "__main__":
    rc, out = run_check(int(sys.argv[1]) if len(sys.argv) > 1 else 1)
    print(out)
    sys.exit(rc)
