"""The op-counting build of the CPU restatement (oracle/opcount.py,
oracle/opcount_rt.c, tests/opcount_run.py).

* The counter is pinned on a program whose operation count is known in
  closed form: a loop with a data-dependent call, run on seeded data, must
  count exactly n multiplies, n + 1 adds, n compares, one divide and one
  conversion in its own phase and exactly k multiply-adds in the callee's
  phase (k = the products above the threshold, computed in float32 by numpy).
* The counting build of the restatement renders the same frame as the
  reference (depth, ID, HDR and RGBA8 bit-identical), so what it counts is
  the frame the parity tests compare.
* Its counts are deterministic and equal the committed fixtures
  (tests/golden/opcount_*.json, written by tests/golden/gen_opcount.py).
"""
import ctypes as C
import importlib.util
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import refcore

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang"
GOLDEN = os.path.join(REPO, "tests", "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG), reason="clang not present")


def _opcount_module():
    spec = importlib.util.spec_from_file_location(
        "opcount", os.path.join(REPO, "oracle", "opcount.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


KNOWN = r"""
static float g_acc;
__attribute__((noinline)) void hit(float v) { g_acc = g_acc * 0.5f + v; }
float run(const float *a, const float *b, int n) {
  float s = 0.0f;
  for (int i = 0; i < n; i++) {
    float t = a[i] * b[i];
    s = s + t;
    if (t > 0.25f)
      hit(t);
  }
  return s / (float)n + g_acc;
}
"""


def test_counter_on_a_program_of_known_count(tmp_path):
    opc = _opcount_module()
    src = tmp_path / "known.c"
    src.write_text(KNOWN)
    ll = tmp_path / "known.ll"
    subprocess.run([CLANG, "-std=c11", "-O2", "-fno-inline", "-fno-vectorize",
                    "-fno-slp-vectorize", "-ffp-contract=off", "-fPIC", "-S", "-emit-llvm",
                    str(src), "-o", str(ll)], check=True)
    opc.PHASE_OF = {"run": "shade", "hit": "walk"}
    res, blocks = opc.instrument(ll.read_text().split("\n"))
    (tmp_path / "known_opc.ll").write_text("\n".join(res) + "\n")
    subprocess.run([CLANG, "-O0", "-fPIC", "-c", str(tmp_path / "known_opc.ll"), "-o",
                    str(tmp_path / "known.o")], check=True)
    subprocess.run(["gcc", "-O2", "-fPIC", "-c", os.path.join(REPO, "oracle", "opcount_rt.c"),
                    "-o", str(tmp_path / "rt.o")], check=True)
    so = tmp_path / "libknown.so"
    subprocess.run(["gcc", "-shared", "-o", str(so), str(tmp_path / "known.o"),
                    str(tmp_path / "rt.o")], check=True)
    lib = C.CDLL(str(so))
    lib.run.restype = C.c_float
    lib.run.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    nb, nph = lib.mop_opcount_blocks(), lib.mop_opcount_phases()
    assert nb == len(blocks) and nph == len(opc.PHASES)
    kinds = sorted({k for b in blocks for k in b["ops"]})
    static = np.array([[b["ops"].get(k, 0) for k in kinds] for b in blocks], np.int64)
    rng = np.random.default_rng(5)
    for n in (0, 1, 7, 1000):
        a = rng.random(n, dtype=np.float32)
        b = rng.random(n, dtype=np.float32)
        lib.mop_opcount_reset()
        lib.run(a.ctypes.data, b.ctypes.data, n)
        runs = np.zeros(nph * nb, np.uint64)
        lib.mop_opcount_read(runs.ctypes.data_as(C.c_void_p))
        per = runs.reshape(nph, nb).astype(np.int64) @ static
        got = {ph: {k: int(per[p, j]) for j, k in enumerate(kinds) if per[p, j]}
               for p, ph in enumerate(opc.PHASES)}
        k = int(np.count_nonzero(a * b > np.float32(0.25)))
        want = {"mul": n, "add": n + 1, "cmp": n, "div": 1, "cvt": 1}
        assert got["shade"] == {k_: v for k_, v in want.items() if v}, (n, got)
        assert got["walk"] == ({"mul": k, "add": k} if k else {}), (n, k, got)
        assert all(not got[p] for p in opc.PHASES if p not in ("shade", "walk"))


def _run(scene, shadow=0, check=False):
    cmd = [sys.executable, os.path.join(REPO, "tests", "opcount_run.py"), "--scene", scene]
    if shadow:
        cmd += ["--shadow", str(shadow)]
    if check:
        cmd += ["--check"]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    return json.loads(p.stdout)


@pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")
def test_counting_build_renders_the_reference_frame_and_matches_fixture():
    got = _run("small", check=True)
    with open(os.path.join(GOLDEN, "opcount_small.json")) as f:
        want = json.load(f)
    assert got == want
    # the shading pass does work on covered pixels only in proportion to
    # them: every covered pixel runs two IBL direction look-ups
    sh = got["phases"]["shade"]
    assert sh["libm_atan2f"] == sh["libm_asinf"] == 2 * got["covered_px"]


@pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")
def test_1080p_pbr_count_matches_fixture():
    got = _run("config2", shadow=2048)
    with open(os.path.join(GOLDEN, "opcount_1080p_pbr.json")) as f:
        want = json.load(f)
    assert got == want
