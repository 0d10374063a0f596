"""Writes the golden fixtures: buffers rendered by the reference CPU backend
(compiled from the reference sources into oracle/_ref, forced
single-threaded so draws run in submission order) through the unmodified
reference viewport core.

Small scenes are stored whole (RGBA8, depth, object-ID, HDR); full-size
configuration 1 is stored as SHA-256 digests of each buffer.  Run from the
repository root:  python tests/golden/gen_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import conftest  # noqa: E402,F401  (registers the package as mop_amd)
import refcore  # noqa: E402
from mop_amd import scenes  # noqa: E402


def digest(buf):
    return hashlib.sha256(np.ascontiguousarray(buf).tobytes()).hexdigest()


def small_cases():
    return {
        "small_shadow_96x72": scenes.small_scene(seed=21, width=96, height=72, n=8),
        "small_ibl_96x72": scenes.small_scene(seed=22, width=96, height=72, n=8, with_env=True),
        "cfg1_96x72": scenes.config1(96, 72),
        "cfg3_near_96x54": scenes.config3(96, 54, n_quads=48, n_spheres=4, shadow_size=1024),
    }


def main():
    out = {}
    for name, sc in small_cases().items():
        r = refcore.render(sc, 1)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), color=r["color"],
                            depth=r["depth"], ids=r["ids"], hdr=r["hdr"])
        out[name] = {"triangles": int(r["triangles"]), "exposure": float(r["exposure"])}
        print("wrote", name)
    sc = scenes.config1()
    r = refcore.render(sc, 1)
    out["cfg1_800x600_sha256"] = {k: digest(r[k]) for k in ("color", "depth", "ids", "hdr")}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
