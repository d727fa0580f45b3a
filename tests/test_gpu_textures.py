"""Textured materials on the device against the oracle: base-colour, metallic-roughness,
normal and emission maps, stochastic transparency and alpha shadows through a textured veil."""
import numpy as np
import pytest

import pt_amd
import oracle_bind as ob

pytestmark = pytest.mark.gpu

W, H = 160, 160


@pytest.fixture(scope="module")
def tex(scene_dir):
    from pt_amd.scenes import generate
    scene = pt_amd.Scene(generate.cached_scene("textured", root=scene_dir))
    cam = scene.camera(W, H)
    t = pt_amd.Tracer(scene, W, H, max_bounces=4)
    t.set_camera(cam)
    return scene, cam, t, ob.Oracle(scene)


def test_textured_render_matches_oracle_per_pixel(tex):
    scene, cam, t, orc = tex
    t.set_camera(cam)
    t.render(4)
    g = t.read()
    o, counts = orc.render(cam, W, H, 4, 4)
    close = np.abs(g[..., :3] - o[..., :3]) <= 2e-3 * (1.0 + np.abs(o[..., :3]))
    frac = close.all(axis=-1).mean()
    assert frac >= 0.98, frac
    st = t.stats()
    assert abs(st["extension_rays"] - counts["extension_rays"]) <= 0.01 * counts["extension_rays"]
    assert abs(st["shadow_rays"] - counts["shadow_rays"]) <= 0.01 * counts["shadow_rays"]


def test_textured_render_converges_to_oracle(tex):
    scene, cam, t, orc = tex
    t.set_camera(cam)
    t.render(64)
    g = t.read()
    o, _ = orc.render(cam, W, H, 64, 4)
    assert pt_amd.rel_mse(g, o) <= 1e-4


def test_textures_change_the_image(tex):
    """Sanity: the maps are sampled (an image with every map bit cleared differs)."""
    scene, cam, t, orc = tex
    t.set_camera(cam)
    t.render(4)
    g = t.read()
    # the checker floor: neighbouring 8-texel squares differ strongly in base colour
    floor = g[int(H * 0.85):, W // 4:3 * W // 4, :3]
    assert floor.std() > 0.02
