"""Textured materials in the host loader: map bits and indices as process_materials packs them
(src/scene/scene.cpp:74-395), texture data as TextureData holds it, index offsets across .scn
objects (PbrMaterial::offset_maps_by, src/scene/material.cpp:16-42) and the Monte Carlo
intensity of a textured emitter (calculate_object_intensity, src/scene/lightsample.cpp:41-87)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import pt_amd
from pt_amd.scenes import generate
from pt_amd.scenes.rng import Splitmix32, Xshiro128
from test_png import decode

BASE, EMISSION, RM, NORMAL, TRANSPARENT = 1 << 0, 1 << 1, 1 << 4, 1 << 12, 1 << 13


@pytest.fixture(scope="module")
def tex_dir(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("textured"))
    generate.textured_scene(d)
    return d


def _materials(scene):
    m = scene.arrays()["materials"].view(np.float32).reshape(-1, 24)
    return m, m.view(np.uint32)


def _textures(scene):
    b = scene.blob.contents
    descs = [(b.textures[i].offset, b.textures[i].width, b.textures[i].height, b.textures[i].channels,
              b.textures[i].srgb) for i in range(b.texture_count)]
    texels = bytes(np.ctypeslib.as_array(b.texels, (b.texel_bytes,))) if b.texel_bytes else b""
    return descs, texels


def test_material_maps(tex_dir):
    s = pt_amd.Scene(os.path.join(tex_dir, "textured.scn"))
    f, u = _materials(s)
    bits = u[:, 11]
    # materials in glTF order: floor, back, ceil, left, right, veil, sphere, light
    assert bits[0] == BASE | RM | NORMAL
    assert (u[0, 0], u[0, 8], u[0, 10]) == (0, 1, 2)  # base colour / metallic-roughness / normal map
    assert bits[1] == BASE and u[1, 0] == 3
    assert bits[5] == BASE | TRANSPARENT and u[5, 0] == 4  # alpha stripes -> transparent
    assert bits[7] == EMISSION and u[7, 4] == 5 and f[7, 7] == 15.0
    assert bits[2] == 0 and bits[6] == 0
    descs, texels = _textures(s)
    assert [d[4] for d in descs] == [1, 0, 0, 1, 1, 1]  # sRGB: base-colour and emission maps
    names = ["floor_base", "floor_mr", "floor_normal", "wall_base", "veil_base", "light_emission"]
    for (off, w, h, c, _), name in zip(descs, names):
        data = open(os.path.join(tex_dir, name + ".png"), "rb").read()
        dw, dh, dc, px = decode(data)
        assert (w, h, c) == (dw, dh, dc) == (64, 64, 4)
        assert texels[off:off + w * h * c] == px


def test_texture_indices_offset_per_object(tex_dir):
    p = os.path.join(tex_dir, "two.scn")
    with open(p, "w") as fh:
        fh.write("camera:\nobjects:\n a textured.gltf\n b textured.gltf\n\n a\n b\n  t 3 0 0\n")
    s = pt_amd.Scene(p)
    f, u = _materials(s)
    assert s.blob.contents.texture_count == 12
    assert (u[8, 0], u[8, 8], u[8, 10]) == (6, 7, 8)  # the second object's floor
    assert u[15, 4] == 11  # its light's emission map


def test_shared_image_is_one_texture(tex_dir):
    import json
    doc = json.load(open(os.path.join(tex_dir, "textured.gltf")))
    doc["textures"].append({"source": 0})  # a second texture object over the floor image
    doc["materials"][1]["pbrMetallicRoughness"]["baseColorTexture"] = {"index": len(doc["textures"]) - 1}
    path = os.path.join(tex_dir, "shared.gltf")
    json.dump(doc, open(path, "w"))
    with open(os.path.join(tex_dir, "shared.scn"), "w") as fh:
        fh.write("camera:\nobjects:\n sh shared.gltf\n\n sh\n")
    s = pt_amd.Scene(os.path.join(tex_dir, "shared.scn"))
    f, u = _materials(s)
    assert u[1, 0] == 0  # the back wall now shares texture 0 with the floor
    assert s.blob.contents.texture_count == 5


def _intensity(positions_uv, tris, tex, strength):
    """calculate_object_intensity restated in float32/float64 as the reference computes it, with
    the generator state of its g++ build (arguments evaluated right to left)."""
    sm = Splitmix32(0xBD7A0D48)
    d = [sm.next() for _ in range(4)]
    rng = Xshiro128(d[3], d[2], d[1], d[0])
    h, w, c = tex.shape
    f32 = np.float32
    total = 0.0
    for ia, ib, ic in tris:
        ua, ub, uc = positions_uv[ia], positions_uv[ib], positions_uv[ic]
        tri_sum = 0.0
        for _ in range(20):
            beta = f32(1.0) - np.sqrt(f32(rng.next_float()))
            gamma = (f32(1.0) - beta) * f32(rng.next_float())
            alpha = f32(1.0) - beta - gamma
            uv = [f32(alpha * f32(ua[k])) + f32(beta * f32(ub[k])) + f32(gamma * f32(uc[k])) for k in range(2)]
            xs = []
            for k, size in ((0, w), (1, h)):
                frac = f32(math.modf(float(uv[k]))[0])
                v = f32(frac * f32(size))
                r = math.floor(float(v) + 0.5) if v >= 0 else -math.floor(-float(v) + 0.5)
                xs.append(size - 1 if r < 0 else min(int(r), size - 1))
            px = tex[xs[1], xs[0]]
            chans = [f32(px[0]) / f32(255), f32(px[1]) / f32(255) if c >= 2 else f32(0),
                     f32(px[2]) / f32(255) if c == 4 else f32(0)]
            tri_sum += float(f32(f32(chans[0] + chans[1]) + chans[2]))
        total += tri_sum
    return float(f32(total * strength / (len(tris) * 20)))


def test_textured_emitter_intensity(tex_dir):
    # a point light beside the textured emitter: the light table's p values fix the ratio of
    # the two weights
    p = os.path.join(tex_dir, "lit.scn")
    with open(p, "w") as fh:
        fh.write("camera:\nobjects:\n plight 0 -1 0 , 1 1 1\n tex textured.gltf\n\n tex\n")
    s = pt_amd.Scene(p)
    words = s.arrays()["light_words"].view(np.uint32)
    assert s.arrays()["light_count"] == 2
    bins = words[:20].reshape(2, 10)
    assert bins[0, 3] == 3 and bins[1, 3] == 1  # emitter first, then the point light
    pvals = bins[:, 0].view(np.float32)
    # the light quad: uv (0,0) (1,0) (1,1) (0,1), triangles (0,1,2) (0,2,3)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    tex = decode(open(os.path.join(tex_dir, "light_emission.png"), "rb").read())
    img = np.frombuffer(tex[3], np.uint8).reshape(tex[1], tex[0], tex[2])
    w_e = _intensity(uv, [(0, 1, 2), (0, 2, 3)], img, 15.0)
    w_p = 3.0 * pt_amd.from_srgb(1.0)
    inv = 1.0 / (np.float64(np.float32(w_e)) + np.float64(np.float32(w_p)))
    assert pvals[0] == np.float32(np.float64(np.float32(w_e)) * inv)
    assert pvals[1] == np.float32(np.float64(np.float32(w_p)) * inv)


def test_jpeg_texture_loads_like_png(tex_dir):
    """A glTF image may be a JPEG (the reference's stb_image reads both): the wall's base-colour
    map replaced by a JPEG file becomes a texture of the JPEG's decoded pixels."""
    import json
    import shutil
    jpg = os.path.join(os.path.dirname(__file__), "golden", "jpeg", "rgb420_q75.jpg")
    shutil.copy(jpg, os.path.join(tex_dir, "wall.jpg"))
    doc = json.load(open(os.path.join(tex_dir, "textured.gltf")))
    src = doc["textures"][doc["materials"][1]["pbrMetallicRoughness"]["baseColorTexture"]["index"]]["source"]
    doc["images"][src] = {"uri": "wall.jpg", "mimeType": "image/jpeg"}
    json.dump(doc, open(os.path.join(tex_dir, "jpeg.gltf"), "w"))
    with open(os.path.join(tex_dir, "jpeg.scn"), "w") as fh:
        fh.write("camera:\nobjects:\n j jpeg.gltf\n\n j\n")
    s = pt_amd.Scene(os.path.join(tex_dir, "jpeg.scn"))
    f, u = _materials(s)
    descs, texels = _textures(s)
    off, w, h, c, srgb = descs[u[1, 0]]
    dw, dh, dc, px = decode(open(jpg, "rb").read())
    assert (w, h, c, srgb) == (dw, dh, dc, 1) == (33, 21, 4, 1)
    assert texels[off:off + w * h * c] == px


SPECULAR, SHEEN, CLEARCOAT = 1 << 5, 1 << 7, 1 << 9


def _edited(tex_dir, name, edit):
    import json
    doc = json.load(open(os.path.join(tex_dir, "textured.gltf")))
    edit(doc)
    json.dump(doc, open(os.path.join(tex_dir, name + ".gltf"), "w"))
    with open(os.path.join(tex_dir, name + ".scn"), "w") as fh:
        fh.write("camera:\nobjects:\n e %s.gltf\n\n e\n" % name)
    return pt_amd.Scene(os.path.join(tex_dir, name + ".scn"))


def test_specular_sheen_clearcoat_slots(tex_dir):
    """KHR_materials_specular / _sheen / _clearcoat as Assimp 5.4.3 passes them to
    process_specular / process_sheen / process_clearcoat_anisotropy (src/scene/scene.cpp:208-290):
    slot-0 maps after the normal map, linear; factors otherwise; all-zero extensions dropped with
    their textures (Assimp's defaults are zero); the specular factor read into a colour sets red."""
    base_bits = _materials(pt_amd.Scene(os.path.join(tex_dir, "textured.scn")))[1][:, 11].copy()

    def edit(doc):
        m = doc["materials"]
        m[2]["extensions"] = {  # ceiling
            "KHR_materials_specular": {"specularFactor": 1.0, "specularTexture": {"index": 1}},
            "KHR_materials_sheen": {"sheenColorFactor": [0.2, 0.3, 0.4], "sheenColorTexture": {"index": 3}},
            "KHR_materials_clearcoat": {"clearcoatFactor": 0.7, "clearcoatTexture": {"index": 4},
                                        "clearcoatRoughnessTexture": {"index": 0}}}
        m[3]["extensions"] = {  # left wall: every factor zero -> nothing kept, no texture loaded
            "KHR_materials_specular": {"specularTexture": {"index": 5}},
            "KHR_materials_sheen": {"sheenColorTexture": {"index": 5}},
            "KHR_materials_clearcoat": {"clearcoatTexture": {"index": 5}, "clearcoatRoughnessFactor": 0.9}}
        m[6]["extensions"] = {  # sphere: factors only
            "KHR_materials_specular": {"specularFactor": 0.6, "specularColorFactor": [1, 1, 1]},
            "KHR_materials_sheen": {"sheenColorFactor": [0.1, 0.2, 0.3]},
            "KHR_materials_clearcoat": {"clearcoatFactor": 0.25}}

    s = _edited(tex_dir, "slots", edit)
    f, u = _materials(s)
    assert u[2, 11] == base_bits[2] | SPECULAR | SHEEN | CLEARCOAT
    # the ceiling's maps: the floor's metallic-roughness image (texture 1), the back wall's
    # base-colour image (3), and the veil image, first used here, so texture 4 and linear
    assert (u[2, 12], u[2, 16], u[2, 20]) == (1, 3, 4)
    assert f[2, 21] == 0.0  # clearcoatRoughnessFactor's default in the importer
    assert u[3, 11] == base_bits[3]
    np.testing.assert_array_equal(f[3, 12:15], [0.5, 0.5, 0.5])
    np.testing.assert_array_equal(f[3, 16:19], [0, 0, 0])
    np.testing.assert_array_equal(f[3, 20:22], [0.0, 0.5])
    assert u[6, 11] == base_bits[6]
    np.testing.assert_allclose(f[6, 12:15], [0.6, 0, 0])
    np.testing.assert_allclose(f[6, 16:19], [0.1, 0.2, 0.3])
    np.testing.assert_allclose(f[6, 20:22], [0.25, 0.0])
    # the veil's base colour and the light's emission reuse / follow those textures
    assert u[5, 0] == 4 and u[7, 4] == 5
    descs, _ = _textures(s)
    assert [d[4] for d in descs] == [1, 0, 0, 1, 0, 1]


def test_texture_without_slot0_texture_is_refused(tex_dir):
    import re

    def edit(doc):
        doc["materials"][2]["extensions"] = {
            "KHR_materials_specular": {"specularFactor": 1.0, "specularColorTexture": {"index": 1}}}

    with pytest.raises(pt_amd.PtError, match=re.escape("specularColorTexture but no specularTexture")):
        _edited(tex_dir, "noslot0", edit)
