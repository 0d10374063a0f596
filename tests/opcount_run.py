"""Floating-point operations per frame of the CPU restatement, by phase and
kind, from its op-counting build (oracle/_build/liboracle_opcount.so; how it
counts is in oracle/opcount.py).  The reference viewport core drives the
counting build exactly as it drives the plain restatement, so the count is of
the frame the parity tests compare.

Runs in a process of its own (the counting build takes the restatement's
place in the harness):

    python3 tests/opcount_run.py --scene config2 --shadow 2048 --out FILE

Output: {"scene", "width", "height", "shadow", "triangles", "covered_px",
"phases": {phase: {kind: ops}}, "total": {kind: ops}, "ops": {phase: ops}}.
"""
import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

LIB = os.path.join(REPO, "oracle", "_build", "liboracle_opcount.so")
TABLE = os.path.join(REPO, "oracle", "_build", "liboracle_opcount.json")


def scene_by_name(name):
    import conftest
    conftest.load_pkg()
    from mop_amd import scenes
    if name == "config2":  # the 1080p PBR frame
        return scenes.config2()
    if name == "small":  # 160x120, shadowed sun + point + spot, IBL
        return scenes.small_scene(seed=7, with_env=True)
    raise SystemExit(f"unknown scene {name}")


def count(scene, shadow_size=None):
    """Runs one frame of scene through the counting build; returns the
    per-phase operation counts and the frame's buffers."""
    import refcore
    refcore.ORACLE_LIB = LIB
    lib = C.CDLL(LIB, mode=C.RTLD_GLOBAL)
    lib.mop_opcount_blocks.restype = C.c_int
    lib.mop_opcount_phases.restype = C.c_int
    with open(TABLE) as f:
        table = json.load(f)
    nb, nph = lib.mop_opcount_blocks(), lib.mop_opcount_phases()
    assert nb == len(table["blocks"]) and nph == len(table["phases"])
    vp = refcore.RefViewport(scene, 3, shadow_size=shadow_size)
    try:
        lib.mop_opcount_reset()
        vp.render()
        runs = np.zeros(nph * nb, np.uint64)
        lib.mop_opcount_read(runs.ctypes.data_as(C.c_void_p))
        out = vp.read()
        tris = vp.triangle_count()
    finally:
        vp.close()
    runs = runs.reshape(nph, nb).astype(np.int64)
    kinds = sorted({k for b in table["blocks"] for k in b["ops"]})
    static = np.zeros((nb, len(kinds)), np.int64)
    for i, b in enumerate(table["blocks"]):
        for k, v in b["ops"].items():
            static[i, kinds.index(k)] = v
    per = runs @ static  # phases x kinds
    phases = {}
    for p, name in enumerate(table["phases"]):
        row = {k: int(per[p, j]) for j, k in enumerate(kinds) if per[p, j]}
        if row:
            phases[name] = row
    total = {k: int(per[:, j].sum()) for j, k in enumerate(kinds) if per[:, j].sum()}
    res = {"phases": phases, "total": total,
           "ops": {p: sum(v.values()) for p, v in phases.items()},
           "triangles": int(tris),
           "covered_px": int(np.count_nonzero(out["ids"]))}
    return res, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scene", default="small")
    ap.add_argument("--shadow", type=int, default=0)
    ap.add_argument("--out", default="-")
    ap.add_argument("--check", action="store_true",
                    help="also render the frame with the reference (single-threaded) and "
                         "require every buffer bit-identical")
    a = ap.parse_args()
    sc = scene_by_name(a.scene)
    res, out = count(sc, a.shadow or None)
    if a.check:
        import refcore
        ref = refcore.render(sc, 1, shadow_size=a.shadow or None)
        cmp = refcore.compare(out, ref)
        print(f"counting build vs reference: {cmp}", file=sys.stderr)
        refcore.assert_parity(cmp, exact=True)
    res = {"scene": a.scene, "width": sc.width, "height": sc.height,
           "shadow": a.shadow or 1024, **res}
    text = json.dumps(res, indent=1, sort_keys=True)
    if a.out == "-":
        print(text)
    else:
        with open(a.out, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
