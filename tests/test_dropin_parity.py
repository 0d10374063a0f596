"""End-to-end parity: the unmodified reference viewport core driving the HIP
backend (drop-in through the RHI table) against the same core driving the
reference CPU backend forced single-threaded (submission order,
cpu_backend.c:588-620).

Bar (north star): depth and object-ID bit-exact; RGBA8 within 1/255 per
channel; linear HDR within 1e-4 relative (|a-b| <= 1e-4 * max(|ref|, 1e-6)).
This backend meets the stronger bar the tests assert: every buffer
bit-identical (refcore.assert_parity).  The number of differing pixels is
printed for every case.
"""
import pytest

import refcore
from mop_amd import scenes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")]

def _check(scene, frames=1):
    print(f"[case] {scene.name} tris={scene.triangle_count}", flush=True)
    ref = refcore.render(scene, 1, frames=frames)
    hip = refcore.render(scene, 2, frames=frames)
    cmp = refcore.compare(hip, ref)
    print(f"[parity] {scene.name}: {cmp}", flush=True)
    refcore.assert_parity(cmp)
    return cmp


def test_config1_plain_800x600():
    _check(scenes.config1())


def test_config1_two_frames():
    _check(scenes.config1(320, 240), frames=2)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_small_scene_shadow(seed):
    _check(scenes.small_scene(seed=seed, with_shadow=True))


def test_small_scene_ibl():
    _check(scenes.small_scene(seed=5, with_shadow=True, with_env=True))


def test_config2_scaled():
    sc = scenes.config2(320, 180, n_instances=24, subdiv=2, shadow_size=1024,
                        env_size=(256, 128))
    _check(sc)


def test_config3_scaled_near_plane():
    sc = scenes.config3(320, 180, n_quads=96, n_spheres=8, shadow_size=1024)
    _check(sc)


@pytest.mark.parametrize("seed", [9, 10])
def test_ordered_draws_transparent_and_instanced(seed):
    _check(scenes.ordered_scene(seed, with_env=(seed == 10)))


def test_ordered_draws_larger_view():
    _check(scenes.ordered_scene(12, 480, 320))


@pytest.mark.parametrize("lit", [True, False])
def test_flat_and_unlit_shading_models(lit):
    # flat shading (incremental z, centroid lighting), unlit Lambert and the
    # legacy smooth path's silhouette AA, opaque and blended
    _check(scenes.shading_scene(21, lit=lit))
    _check(scenes.shading_scene(22, 320, 200, lit=lit))


@pytest.mark.parametrize("lit", [True, False])
def test_wireframe_render_mode(lit):
    # Xiaolin Wu lines, blended in submission order, over the core's grid
    _check(scenes.wireframe_scene(31, lit=lit))
    _check(scenes.wireframe_scene(32, 320, 200, lit=lit))


@pytest.mark.parametrize("rotation", [0.0, 0.7])
def test_environment_background_skybox(rotation):
    # the core's CPU skybox, drawn by the backend's background hook before the
    # frame's draws (opaque, blended and instanced over it)
    sc = scenes.ordered_scene(10, 200, 150, with_env=True)
    sc.env_background = True
    sc.env_rotation = rotation
    sc.name += f"_sky{rotation}"
    _check(sc)


@pytest.mark.parametrize("alpha_unselected", [0.0, 0.35])
def test_object_outline(alpha_unselected):
    # mop_overlay_builtin_outline on the device, fused into the frame tail
    # between the resolve and FXAA + gamma: thick (selected) and thin
    # (unselected) outlines with order-dependent overlaps, a chrome-range id
    sc = scenes.outline_scene(alpha_unselected=alpha_unselected)
    sc.name += f"_un{alpha_unselected}"
    _check(sc)


def test_object_outline_no_post_two_frames_ssaa():
    # the outline-only tail (no post effects), re-recorded every frame, on the
    # internal SSAA image before the downsample
    sc = scenes.outline_scene(seed=22, width=120, height=90, selected=(2, 5))
    sc.post_effects = 0
    sc.ssaa = 2
    sc.name += "_nopost_ssaa2"
    _check(sc, frames=2)


POST_CASES = {
    "fog_gamma_fxaa": (1 | 8 | 16, ((0.55, 0.6, 0.7, 1.0), 6.0, 20.0)),
    "tonemap_gamma_vignette": (1 | 2 | 4, None),
    "tonemap_fxaa": (2 | 16, None),
    "all": (1 | 2 | 4 | 8 | 16, ((0.2, 0.25, 0.3, 1.0), 2.0, 30.0)),
}


@pytest.mark.parametrize("case", sorted(POST_CASES))
def test_post_effects(case):
    # fog, tonemap, gamma, vignette: without fog and vignette the colour
    # effects are per-channel tables the host fills with its own powf; with
    # them the device evaluates the sRGB curve per pixel with its own powf,
    # which the required bar (1/255) covers -- the differing pixels are
    # reported
    effects, fog = POST_CASES[case]
    sc = scenes.small_scene(seed=41, width=180, height=130, with_env=True)
    sc.post_effects, sc.fog = effects, fog
    sc.name += "_" + case
    print(f"[case] {sc.name}", flush=True)
    ref = refcore.render(sc, 1)
    hip = refcore.render(sc, 2)
    cmp = refcore.compare(hip, ref)
    print(f"[parity] {sc.name}: {cmp}", flush=True)
    refcore.assert_parity(cmp, exact=not (effects & (4 | 8)))


def test_ssaa_3_device_downsample():
    # mop_viewport_read_color's box filter on the device (factor 3)
    sc = scenes.small_scene(seed=26, width=150, height=100, with_env=True)
    sc.ssaa = 3
    _check(sc)


def test_ssaa_2_downsample_and_pick():
    sc = scenes.small_scene(seed=25, width=128, height=96, with_env=True)
    sc.ssaa = 2
    _check(sc)
    ref = refcore.RefViewport(sc, 1)
    hip = refcore.RefViewport(sc, 2)
    ref.render()
    hip.render()
    for x, y in [(10, 10), (64, 48), (100, 80), (5, 90)]:
        assert ref.pick(x, y) == hip.pick(x, y), (x, y)
    ref.close()
    hip.close()
