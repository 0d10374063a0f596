"""Writes the op-count fixtures: floating-point operations per frame of the
CPU restatement's counting build (tests/opcount_run.py) for the small mixed
scene and for the 1080p PBR frame with its 2048^2 shadow map.

    make -C oracle opcount && python3 tests/golden/gen_opcount.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
RUN = os.path.join(os.path.dirname(HERE), "opcount_run.py")

for scene, shadow, name in (("small", 0, "opcount_small.json"),
                            ("config2", 2048, "opcount_1080p_pbr.json")):
    cmd = [sys.executable, RUN, "--scene", scene, "--out", os.path.join(HERE, name), "--check"]
    if shadow:
        cmd += ["--shadow", str(shadow)]
    subprocess.run(cmd, check=True)
    print("wrote", name)
