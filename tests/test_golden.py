"""Golden fixtures (tests/golden/, written by gen_golden.py from the reference
CPU backend): the CPU restatement must reproduce them bit for bit; the HIP
frame driver (standalone path used by the benchmark, batching and serving)
must match depth/ID bit-exactly and colour within 1/255."""
import hashlib
import json
import os

import numpy as np
import pytest

import refcore
from golden.gen_golden import small_cases

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLD = json.load(open(os.path.join(HERE, "golden.json")))
NAMES = sorted(small_cases().keys())


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"))
    return {k: z[k] for k in ("color", "depth", "ids", "hdr")}


@pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    sc = small_cases()[name]
    r = refcore.render(sc, 3)
    cmp = refcore.compare(r, load(name))
    assert cmp["depth_diff_px"] == 0 and cmp["id_diff_px"] == 0
    assert cmp["color_diff_px"] == 0 and cmp["hdr_bitdiff_px"] == 0


@pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")
def test_oracle_config1_full_size_digests():
    from mop_amd import scenes
    r = refcore.render(scenes.config1(), 3)
    got = {k: hashlib.sha256(np.ascontiguousarray(r[k]).tobytes()).hexdigest()
           for k in ("color", "depth", "ids", "hdr")}
    assert got == GOLD["cfg1_800x600_sha256"]


def render_driver(sc):
    """Render a Scene through the HIP frame driver; returns the buffers.  The
    shadow map is the reference's fixed 1024^2 (viewport.c:2286), which the
    goldens were rendered with whatever size the scene asks for."""
    from mop_amd import driver
    dev = driver.Device()
    try:
        drv = driver.SceneDriver(dev, sc, shadow_size=1024)
        fb = dev.framebuffer(sc.width, sc.height)
        tris = drv.render(fb)
        out = fb.read(hdr=True)
        out["triangles"] = tris
        out["exposure"] = drv.exposure()
        fb.close()
        drv.close()
        return out
    finally:
        dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_frame_driver_matches_golden(name):
    sc = small_cases()[name]
    print(f"[case] driver {name}", flush=True)
    r = render_driver(sc)
    g = load(name)
    cmp = refcore.compare(r, g)
    print(f"[parity] driver {name}: {cmp} exposure={r['exposure']} ref={GOLD[name]['exposure']}",
          flush=True)
    assert abs(r["exposure"] - GOLD[name]["exposure"]) <= 1e-6 * GOLD[name]["exposure"]
    refcore.assert_parity(cmp)
