"""Host surface and generators against the reference's own host code (golden vectors).

tests/golden/ref_host.json was produced by tests/golden/make_golden.py from the reference's
rng.cpp, aliastable.h, colour.cpp, lalgebra.cpp, scene/camera.cpp, io/imagewrite.cpp
(stb_image_write 1.16) and util/meansq, compiled where they lie in the reference checkout.
"""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import pt_amd
from pt_amd.scenes import rng as prng

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_host.json")))


def f32(bits):
    return np.array(bits, dtype=np.uint32).view(np.float32)


# --- src/rng.cpp ------------------------------------------------------------------------

@pytest.mark.parametrize("case", GOLDEN["rng"], ids=lambda c: "seed%d" % c["seed"])
def test_rng_streams_match_reference(case):
    sm = prng.Splitmix32(case["seed"])
    assert [sm.next() for _ in range(8)] == case["splitmix"]
    assert list(prng.splitmix32_stream(case["seed"], 8)) == case["splitmix"]
    x = prng.Xshiro128.seeded(case["seed"])
    assert [x.next() for _ in range(16)] == case["xshiro"]
    y = prng.Xshiro128.seeded(case["seed"])
    got = np.array([y.next_float() for _ in range(16)], np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), np.array(case["xshiro_float_bits"], np.uint32))


# --- src/aliastable.h ---------------------------------------------------------------------

@pytest.mark.parametrize("k", range(len(GOLDEN["alias"])))
def test_alias_table_matches_reference(k):
    case = GOLDEN["alias"][k]
    w = f32(case["weights_bits"])
    values = np.arange(len(w), dtype=np.uint32) + 1000
    words = pt_amd.alias_table(w, values, 1)
    np.testing.assert_array_equal(words, np.array(case["words"], np.uint32))


def test_alias_table_marginals_equal_weights():
    # Closed form: P(i) = (q_i + sum_{j: alias_j = i} (1 - q_j)) / n = w_i / sum(w).
    r = np.random.default_rng(5)
    for n in (1, 2, 7, 64, 333):
        w = r.random(n).astype(np.float32) * 4
        w[r.random(n) < 0.1] = 0.0
        if w.sum() == 0:
            w[0] = 1.0
        words = pt_amd.alias_table(w, np.zeros(n, np.uint32), 1).reshape(n, 4)
        q = words[:, 1].view(np.float32).astype(np.float64)
        alias = words[:, 2]
        p = q.copy()
        np.add.at(p, alias, 1.0 - q)
        np.testing.assert_allclose(p / n, w / w.sum(), atol=2e-6)
        np.testing.assert_allclose(words[:, 0].view(np.float32), (w / w.astype(np.float64).sum()), rtol=1e-6)


# --- src/colour.cpp ---------------------------------------------------------------------

def test_srgb_matches_reference():
    g = GOLDEN["srgb"]
    xs = f32(g["x_bits"])
    frm = np.array([pt_amd.from_srgb(float(x)) for x in xs], np.float32)
    to = np.array([pt_amd.to_srgb(float(x)) for x in xs], np.float32)
    np.testing.assert_array_equal(frm.view(np.uint32), np.array(g["from_bits"], np.uint32))
    np.testing.assert_array_equal(to.view(np.uint32), np.array(g["to_bits"], np.uint32))


# --- camera: scenebuild.cpp:184-194, camera.h:16, camera.cpp:21-25, pathtrace.cpp:381-396 -

@pytest.mark.parametrize("k", range(len(GOLDEN["cameras"])))
def test_camera_block_matches_reference(k):
    c = GOLDEN["cameras"][k]
    aspect = float(f32(c["aspect_bits"])[0])
    cam = pt_amd.camera_block(c["position"], c["lookdir"], c["up"], c["fov"], aspect, c["near"], c["far"])
    iv = np.array(cam.inv_view[:], np.float32)
    ip = np.array(cam.inv_proj[:], np.float32)
    np.testing.assert_allclose(iv, f32(c["inv_view_rowmajor"]), rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(ip, f32(c["inv_proj_rowmajor"]), rtol=2e-6, atol=2e-6)


def test_camera_is_point_reflected():
    # The view matrix translates by -position (camera.h:16): .scn position (0,1,-3) puts the eye
    # at (0,-1,3), and the camera looks along -lookdir.
    cam = pt_amd.camera_block((0, 1, -3), (0, 0, 1), fov_deg=60)
    iv = np.array(cam.inv_view[:]).reshape(4, 4)
    np.testing.assert_allclose(iv[:3, 3], [0, -1, 3], atol=1e-6)
    np.testing.assert_allclose(iv[:3, :3] @ [0, 0, -1], [0, 0, -1], atol=1e-6)


# --- InstanceData packing of .scn transforms (scenebuild.cpp:213-221, scenebuffer.cpp:126-132)

def _write_one_tri_gltf(path):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.array([[0, 0, 1]] * 3, np.float32)
    uv = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    idx = np.array([0, 1, 2], np.uint32)
    from pt_amd.scenes import generate
    layout = generate.write_bin(os.path.join(os.path.dirname(path), "tri.bin"), (pos, nrm, uv, idx.reshape(1, 3)))
    generate.write_gltf(path, [(layout, generate._material("m", (0.5, 0.5, 0.5), 0.5, 0.0))])


def test_instance_records_match_reference(tmp_path):
    _write_one_tri_gltf(str(tmp_path / "tri.gltf"))
    lines = ["camera:", " position 0 0 0", "", "objects:", " tri tri.gltf"]
    for c in GOLDEN["transforms"]:
        lines += [" tri", "  s %.9g %.9g %.9g" % tuple(c["s"]),
                  "  r %.9g %.9g %.9g %.9g" % (c["axis"][0], c["axis"][1], c["axis"][2], c["angle"]),
                  "  t %.9g %.9g %.9g" % tuple(c["t"])]
    (tmp_path / "s.scn").write_text("\n".join(lines) + "\n")
    scene = pt_amd.Scene(str(tmp_path / "s.scn"))
    inst = scene.arrays()["instances"].view(np.uint32).reshape(-1, 32)
    assert len(inst) == len(GOLDEN["transforms"])
    for k, c in enumerate(GOLDEN["transforms"]):
        rec = inst[k]
        assert list(rec[:4].view(np.uint32)) == [0, 0, 0xFFFFFFFF, 0]
        # transform = transpose(M): bit-exact (same float products as the reference)
        np.testing.assert_array_equal(rec[4:20].view(np.uint32), np.array(c["instance_transform"], np.uint32))
        # normal = columns 0-2 of M^-1 (a 4x4 inverse; a few ulps apart between implementations)
        np.testing.assert_allclose(rec[20:32].view(np.float32), f32(c["instance_normal"]), rtol=1e-5, atol=1e-6)


# --- imagewrite.cpp (stbi_write_hdr) and util/meansq -------------------------------------

@pytest.mark.parametrize("k", range(len(GOLDEN["hdr"])))
def test_hdr_bytes_match_reference(tmp_path, k):
    c = GOLDEN["hdr"][k]
    img = f32(c["rgba_bits"]).reshape(c["h"], c["w"], 4)
    p = str(tmp_path / "o.hdr")
    pt_amd.write_hdr(p, img)
    assert open(p, "rb").read().hex() == c["file_hex"]


@pytest.mark.parametrize("k", range(len(GOLDEN["hdr"])))
def test_meansq_matches_reference(tmp_path, k):
    c = GOLDEN["hdr"][k]
    a, b = tmp_path / "a.hdr", tmp_path / "b.hdr"
    a.write_bytes(bytes.fromhex(c["file_hex"]))
    b.write_bytes(bytes.fromhex(c["twin_file_hex"]))
    ia, ib = pt_amd.read_hdr(str(a)), pt_amd.read_hdr(str(b))
    assert ia.shape == (c["h"], c["w"], 4)
    got = pt_amd.meansq(ib, ia)
    ref = float(c["meansq_twin_vs_file"])
    assert abs(got - ref) <= 5e-6 * abs(ref) + 1e-12  # the tool prints 6 significant digits
    assert pt_amd.meansq(ia, ia) == 0.0


def test_hdr_round_trip_bounds(tmp_path):
    # RGBE keeps 8 mantissa bits of the largest channel: relative error < 2^-7 of the max.
    r = np.random.default_rng(3)
    img = (r.random((5, 37, 4)) * np.exp(r.normal(0, 3, (5, 37, 1)))).astype(np.float32)
    p = str(tmp_path / "x.hdr")
    pt_amd.write_hdr(p, img)
    back = pt_amd.read_hdr(p)
    m = img[..., :3].max(axis=-1, keepdims=True)
    assert np.all(np.abs(back[..., :3] - img[..., :3]) <= m * 2.0 ** -7 + 1e-30)
    assert np.all(back[..., 3] == 1.0)
