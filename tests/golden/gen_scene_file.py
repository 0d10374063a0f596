"""Writes the .mop v2 fixture: a small scene saved by the reference's own
mop_scene_save with quantized vertices (scene_quantized.mop), and the
reference loader's decode of each mesh (mop_scene_get_mesh) beside it
(scene_quantized_decoded.npz).  Needs oracle/_ref (the reference built from
its sources by oracle/Makefile).  Run from the repository root:
    python tests/golden/gen_scene_file.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import conftest  # noqa: E402,F401  (imports the package as mop_amd)
import refcore  # noqa: E402
from mop_amd import scenes  # noqa: E402


def main():
    sc = scenes.file_scene(seed=33, big=False)
    sc.meshes = sc.meshes[:4]
    path = os.path.join(HERE, "scene_quantized.mop")
    refcore.save_scene_file(sc, path, quantize=True)
    nm = refcore.scene_file_info(path)[0]
    arrays = {"n_meshes": np.int64(nm)}
    for k in range(nm):
        v, i, box = refcore.scene_file_mesh(path, k)
        arrays[f"v{k}"], arrays[f"i{k}"], arrays[f"box{k}"] = v, i, box
    np.savez_compressed(os.path.join(HERE, "scene_quantized_decoded.npz"), **arrays)
    print(f"{path}: {os.path.getsize(path)} bytes, {nm} meshes")


if __name__ == "__main__":
    main()
