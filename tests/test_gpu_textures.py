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
    d = np.abs(g[..., :3] - o[..., :3])
    close = d <= 2e-3 * (1.0 + np.abs(o[..., :3]))
    frac = close.all(axis=-1).mean()
    # measured: 99.996 % of pixels agree, mean difference 2e-6 of the mean radiance
    assert frac >= 0.999, frac
    assert d.mean() <= 1e-4 * np.abs(o[..., :3]).mean(), d.mean()
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


def test_specular_and_sheen_maps_match_oracle(tex):
    """Specular and sheen maps (get_specular / get_sheen, material.hlsli:70-88) on the back wall
    and the sphere. Their effect is small (f0 = 0.08 x specular, sheen x Fresnel), so the test is
    differential: with the same seeds the paths are the same, and on the pixels where device and
    oracle agree in both scenes, the change the maps make must be the oracle's."""
    import json
    import os
    scene, cam, t, orc = tex
    d = os.path.dirname(scene.path)
    doc = json.load(open(os.path.join(d, "textured.gltf")))
    m = doc["materials"]
    m[1]["extensions"] = {"KHR_materials_specular": {"specularFactor": 1.0, "specularTexture": {"index": 1}},
                          "KHR_materials_sheen": {"sheenColorFactor": [1, 1, 1], "sheenColorTexture": {"index": 3}}}
    m[6]["extensions"] = {"KHR_materials_specular": {"specularFactor": 1.0, "specularTexture": {"index": 0}},
                          "KHR_materials_sheen": {"sheenColorFactor": [1, 1, 1], "sheenColorTexture": {"index": 0}}}
    json.dump(doc, open(os.path.join(d, "maps.gltf"), "w"))
    with open(os.path.join(d, "maps.scn"), "w") as fh:
        fh.write(open(scene.path).read().replace("textured.gltf", "maps.gltf"))
    s2 = pt_amd.Scene(os.path.join(d, "maps.scn"))
    u = s2.arrays()["materials"].view(np.uint32).reshape(-1, 24)
    assert u[1, 11] & (1 << 5) and u[1, 11] & (1 << 7) and u[6, 11] & (1 << 5) and u[6, 11] & (1 << 7)
    cam2 = s2.camera(W, H)
    t2 = pt_amd.Tracer(s2, W, H, max_bounces=4)
    t2.set_camera(cam2)
    t2.render(4)
    g1 = t2.read()[..., :3]
    t2.close()
    t.set_camera(cam)
    t.render(4)
    g0 = t.read()[..., :3]
    o0 = orc.render(cam, W, H, 4, 4)[0][..., :3]
    o1 = ob.Oracle(s2).render(cam2, W, H, 4, 4)[0][..., :3]
    close = (np.abs(g0 - o0) <= 2e-3 * (1.0 + np.abs(o0))).all(-1) & \
            (np.abs(g1 - o1) <= 2e-3 * (1.0 + np.abs(o1))).all(-1)
    assert close.mean() >= 0.97, close.mean()
    d_g, d_o = (g1 - g0)[close], (o1 - o0)[close]
    assert np.abs(d_o).sum() > 0.0005 * np.abs(o0[close]).sum()  # the maps change the image
    assert np.abs(d_g - d_o).sum() <= 0.05 * np.abs(d_o).sum()
