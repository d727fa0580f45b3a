"""Host-side surface: .scn parsing semantics and errors, glTF loading, the packed scene and
light tables, the CLI's argument handling, and the C ABI's exported symbols and argument
checks (no GPU needed; nothing here launches a kernel)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pt_amd
from pt_amd.scenes import generate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _scn(tmp_path, text, gltf=True):
    if gltf:
        from test_golden_host import _write_one_tri_gltf
        _write_one_tri_gltf(str(tmp_path / "tri.gltf"))
    p = tmp_path / "s.scn"
    p.write_text(text)
    return str(p)


# --- .scn semantics (src/scene/scenebuild.cpp) ----------------------------------------------

def test_scn_defaults_and_envcol_srgb(tmp_path):
    s = pt_amd.Scene(_scn(tmp_path, "camera:\n envcol 0.5 0.25 1\nobjects:\n tri tri.gltf\n tri\n"))
    p = s.camera_params(400, 200)
    assert list(p.lookdir) == [0, 0, -1] and list(p.up) == [0, 1, 0]
    assert p.fov_deg == 45 and abs(p.near_clip - 0.1) < 1e-7 and p.far_clip == 1000
    assert p.focus_dist == 0 and p.lens_radius == 0 and p.exposure == 1 and p.aspect == 2.0
    env = s.arrays()["envcol"]
    np.testing.assert_allclose(env[:3], [pt_amd.from_srgb(0.5), pt_amd.from_srgb(0.25), 1.0], rtol=1e-6)


def test_scn_tabs_and_carriage_returns_are_removed(tmp_path):
    # "\t" and "\r" are deleted, not replaced (scenebuild.cpp:284-286)
    s = pt_amd.Scene(_scn(tmp_path, "camera:\r\n\tfov\t 30\r\nobjects:\r\n tri tri.gltf\r\n tri\r\n"))
    assert s.camera_params(1, 1).fov_deg == 30


def test_scn_transform_order_and_material_keywords(tmp_path):
    s = pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n tri tri.gltf\n tri\n  basecol 1 0 0\n  s 2 2 2\n"
                                    "  t 1 0 0\n tri\n  t 1 0 0\n  s 2 2 2\n"))
    inst = s.arrays()["instances"].view(np.float32).reshape(-1, 32)
    m0 = inst[0, 4:20].reshape(4, 4)  # row-major M
    m1 = inst[1, 4:20].reshape(4, 4)
    assert m0[0, 3] == 1.0 and m0[0, 0] == 2.0   # T * S: translation unscaled
    assert m1[0, 3] == 2.0 and m1[0, 0] == 2.0   # S * T: translation scaled


def test_scn_rotation_below_threshold_is_identity(tmp_path):
    s = pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n tri tri.gltf\n tri\n  r 0 0 0 0.000001\n"))
    inst = s.arrays()["instances"].view(np.float32).reshape(-1, 32)
    np.testing.assert_array_equal(inst[0, 4:20].reshape(4, 4), np.eye(4, dtype=np.float32))


@pytest.mark.parametrize("text,msg", [
    ("objects:\n tri tri.gltf\n tri\n", "No camera section"),
    ("camera:\ncamera:\n", "Camera already set"),
    ("camera:\n bogus 1\n", "Unrecognised token in camera section"),
    ("stuff:\n", "Unrecognised token"),
    ("camera:\nobjects:\n tri tri.gltf\n other\n", "Object other not recognised"),
    ("camera:\nobjects:\n t 1 2 3\n", "No instance to translate"),
    ("camera:\nobjects:\n tri tri.gltf\n tri\n r 1 0 0\n", "Rotation did not have three coordinates and angle"),
    ("camera:\nobjects:\n tri tri.gltf\n tri\n s 1 1\n", "Scaling did not have three coordinates"),
    ("camera:\nobjects:\n tri tri.gltf\n tri\n r 0 0 0 1\n", "Rotation axis was (close to) zero"),
    ("camera:\nobjects:\n plight 1 2 3 4 5 6\n", "Invalid light syntax"),
    ("camera:\n position 1 2\n", "Vectors need three components"),
    ("camera:\nobjects:\n tri missing.gltf\n tri\n", "Could not open file"),
    ("camera:\nobjects:\n basecol 1 0 0\n", "No instance to apply base colour parameter to."),
    ("camera:\nobjects:\n metalness 1\n", "No instance to apply metalness parameter to."),
    ("camera:\nobjects:\n r 1 0 0 1\n", "No instance to rotate."),
    ("camera:\nobjects:\n s 1 1 1\n", "No instance to scale."),
    ("camera:\nobjects:\n dlight 1 2 3 x 4 5 6\n", "Invalid light syntax"),
    ("camera:\nobjects:\n plight 1 2 3 x 4 5 6\n", "Invalid light syntax."),
    ("camera:\n fov\n", "Token stream has no next token"),
    ("camera:\n fov abc\n", "stof"),
    ("camera:\nobjects:\n tri tri.gltf\n tri\n t 1 2 3 4\n", "Translation did not have three coordinates"),
    ("camera: junk\nobjects:\n tri tri.gltf\n tri\n objects\n", "Object objects not recognised."),
    ("camera:\nobjects:\n tri tri.gltf\n tri\nfoo:\n", "Unrecognised token foo:"),
])
def test_scn_errors(tmp_path, text, msg):
    with pytest.raises(pt_amd.PtError, match=re.escape(msg)):
        pt_amd.Scene(_scn(tmp_path, text))


def test_scn_light_comma_messages_differ_in_their_period(tmp_path):
    # scenebuild.cpp:70 vs :87: the directional light's comma check has no final period
    for kw, want in (("plight", "Invalid light syntax."), ("dlight", "Invalid light syntax")):
        with pytest.raises(pt_amd.PtError) as e:
            pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n %s 1 2 3 x 4 5 6\n" % kw))
        assert str(e.value).endswith(want), str(e.value)


def test_scn_declarations_and_instance_flushing(tmp_path):
    """A name followed by any word declares an object (so 'position 3 0 0' in the objects section
    declares 'position' with file '3', unused); a name declared again is a new object whose
    instances use the new file; the instance being described when an objects section ends is
    added then, and again at the end of a later objects section, as the reference adds it
    (scenebuild.cpp:98-125, 245-257)."""
    s = pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n tri tri.gltf\n position 3 0 0\n tri\n  t 1 0 0\n"))
    assert s.info()["instances"] == 1
    # a second objects section adds the pending instance once more
    s2 = pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n tri tri.gltf\n tri\n  t 2 0 0\nobjects:\n"))
    inst = s2.arrays()["instances"].view(np.float32).reshape(-1, 32)
    assert len(inst) == 2 and inst[0, 7] == 2.0 and inst[1, 7] == 2.0
    # re-declaring tri: instances before it keep the old file (a missing one here would fail the load)
    import shutil
    shutil.copy(str(tmp_path / "tri.gltf"), str(tmp_path / "tri2.gltf"))
    s3 = pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n tri tri.gltf\n tri\n tri tri2.gltf\n tri\n  t 0 1 0\n"))
    assert s3.info()["instances"] == 2


def test_point_and_directional_lights_in_light_table(tmp_path):
    s = pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n plight 0 2 0 , 1 1 1\n dlight 0 -1 0 , 0.5 0.5 0.5\n"
                                    " tri tri.gltf\n tri\n"))
    a = s.arrays()
    assert a["light_count"] == 2
    words = a["light_words"].reshape(2, 10)
    assert words[0, 3] == 1 and words[1, 3] == 2  # LIGHT_TYPE_POINT, LIGHT_TYPE_DIRECTIONAL
    np.testing.assert_allclose(words[0, 4:7].view(np.float32), [0, 2, 0])
    np.testing.assert_allclose(words[1, 7:10].view(np.float32), [pt_amd.from_srgb(0.5)] * 3, rtol=1e-6)
    p = words[:, 0].view(np.float32)
    f = 3 * pt_amd.from_srgb(0.5)  # light colours are sRGB-decoded; weight = r+g+b power
    np.testing.assert_allclose(p, [3 / (3 + f), f / (3 + f)], rtol=1e-6)


def test_gltf_punctual_lights(tmp_path):
    """KHR_lights_punctual as process_lights (src/scene/scene.cpp:449-474) receives it from
    Assimp: colour = color x intensity, not sRGB-decoded; position (0, 0, 0) / direction
    (0, 0, -1) in the referencing node's space, without the node transform; spot lights
    dropped; then one copy per .scn instance in the instance's frame (scenebuild.cpp:323-335).
    Assimp itself is absent here: parity unpinned beyond these rules."""
    import json
    _scn(tmp_path, "camera:\nobjects:\n tri tri.gltf\n tri\n")  # writes tri.gltf
    doc = json.load(open(os.path.join(tmp_path, "tri.gltf")))
    doc.setdefault("extensions", {})["KHR_lights_punctual"] = {"lights": [
        {"type": "point", "color": [1.0, 0.5, 0.25], "intensity": 4.0},
        {"type": "spot", "intensity": 9.0, "spot": {}},
        {"type": "directional", "intensity": 2.0},
    ]}
    doc["nodes"].append({"extensions": {"KHR_lights_punctual": {"light": 0}}, "translation": [5, 5, 5]})
    json.dump(doc, open(os.path.join(tmp_path, "lit.gltf"), "w"))
    s = pt_amd.Scene(_scn(tmp_path, "camera:\nobjects:\n lit lit.gltf\n\n lit\n  t 1 2 3\n lit\n"
                                    "  r 0 1 0 1.5707964\n  t -1 0 0\n", gltf=False))
    info = s.info()
    assert info["point_lights"] == 2 and info["directional_lights"] == 2
    a = s.arrays()
    assert a["light_count"] == 4
    words = a["light_words"].reshape(-1, 10)[:4]
    assert list(words[:, 3]) == [1, 1, 2, 2]
    pos = words[:2, 4:7].view(np.float32)
    np.testing.assert_array_equal(pos, [[1, 2, 3], [-1, 0, 0]])  # the instances' origins
    np.testing.assert_array_equal(words[0, 7:10].view(np.float32), [4.0, 2.0, 1.0])
    np.testing.assert_array_equal(words[2, 7:10].view(np.float32), [2.0, 2.0, 2.0])
    d0, d1 = words[2, 4:7].view(np.float32), words[3, 4:7].view(np.float32)
    np.testing.assert_array_equal(d0, [0, 0, -1])
    np.testing.assert_allclose(d1, [-1, 0, 0], atol=1e-6)  # (0, 0, -1) turned pi/2 about y


# --- the generated correctness scene ------------------------------------------------------

def test_box_scene_packing(box_scn):
    s = pt_amd.Scene(box_scn)
    info = s.info()
    assert info["unique_triangles"] == 1292 and info["instances"] == 5 and info["emissive_instances"] == 1
    a = s.arrays()
    assert a["light_count"] == 1
    words = a["light_words"]
    # light bin: p = 1, q = 1, alias 0, emitter type, table at word 10, 2 triangles, instance 4
    assert list(words[:10]) == [0x3F800000, 0x3F800000, 0, 3, 10, 2, 4, 1, 0, 0]
    tri_bins = words[10:].reshape(2, 6)
    np.testing.assert_allclose(tri_bins[:, 0].view(np.float32), [0.5, 0.5])  # equal-area halves
    inst = a["instances"].view(np.uint32).reshape(-1, 32)
    assert inst[4, 2] == 0 and all(inst[k, 2] == 0xFFFFFFFF for k in range(4))
    mats = a["materials"]
    np.testing.assert_allclose(mats[4, 4:8], [1, 1, 1, 15])  # emission rgb, strength
    np.testing.assert_allclose(mats[0, 0:3], [0.7, 0.7, 0.7], rtol=1e-5)  # linear albedo 0.7


def test_identical_emitters_share_the_first_table(tmp_path):
    # lightsample.cpp:126: every emitter's table offset points at the first emitter table.
    scn = generate.instanced_scene(str(tmp_path), seed=9, counts={"ico2": 2}, emitters=3, grid_res=4,
                                   soup_tris=10)
    s = pt_amd.Scene(scn)
    a = s.arrays()
    assert a["light_count"] == 3
    bins = a["light_words"][:30].reshape(3, 10)
    assert set(bins[:, 4]) == {30}
    assert list(bins[:, 6]) == [2, 3, 4] and list(bins[:, 7]) == [1, 1, 1]


# --- CLI (src/main.cpp, src/argparse.cpp) -----------------------------------------------------

PT = os.path.join(ROOT, "pt_amd", "bin", "pt")


@pytest.mark.parametrize("args,code,msg", [
    ([], 1, "Expected path to scene"),
    (["a.scn", "b.scn"], 1, "Unexpected number of arguments"),
    (["-s0", "a.scn"], 1, "Invalid sample count: 0"),
    (["-b", "-1", "a.scn"], 1, "Invalid number of ray bounces: -1"),
    (["--headless", "a.scn"], 1, "No output file was given"),
    (["-w0", "a.scn"], 1, "Invalid dimensions: (0, 400)"),
    (["-f", "exr", "a.scn"], 1, "Output format is invalid: exr"),
    (["--nope", "a.scn"], 1, "Unknown option --nope."),
    (["-s", "x", "a.scn"], 1, "was not an integer: x"),
    (["--help"], 0, "Usage:"),
    (["a.scn", "--width"], 1, "Option --width has no value."),
    (["a.scn", "-o"], 1, "Option -o has no value."),
    (["-x", "a.scn"], 1, "Unknown option -x."),
    (["-qfoo", "a.scn"], 1, "Unknown option -q."),
    (["-t", "1.5x", "a.scn"], 1, "was not a decimal number: 1.5x"),
    (["-s", "+", "a.scn"], 1, "was not an integer: +"),
    (["-w64", "-h0", "a.scn"], 1, "Invalid dimensions: (64, 0)"),
    (["--", "-w", "a.scn"], 1, "Unexpected number of arguments"),
    (["-", "a.scn"], 1, "Unexpected number of arguments"),
])
def test_cli_argument_errors(args, code, msg):
    if not os.path.exists(PT):
        pytest.skip("pt CLI not built")
    r = subprocess.run([PT] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == code
    assert msg in r.stdout + r.stderr


def test_cli_help_lists_the_options_in_order():
    if not os.path.exists(PT):
        pytest.skip("pt CLI not built")
    r = subprocess.run([PT, "--help"], capture_output=True, text=True, timeout=60)
    lines = r.stdout.splitlines()
    assert lines[0].endswith(" [options] path-to-scene") and lines[1] == "Options:"
    assert [ln.split("\t")[0].strip() for ln in lines[2:]] == [
        "--help", "--width, -w", "--height, -h", "--headless", "--samples, -s", "--bounces, -b", "--out, -o",
        "--format, -f", "--time, -t"]


# --- C ABI ---------------------------------------------------------------------------------

def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(pth?_[a-z_0-9]+)\s*\(", text, re.M)))


@pytest.mark.parametrize("header,lib", [("pthip.h", "libpthip.so"), ("pthost.h", "libpthost.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(ROOT, "pt_amd", "lib", lib)
    names = _declared(header)
    assert len(names) >= 8
    dl = C.CDLL(path)
    for n in names:
        assert hasattr(dl, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(r"\bT %s$" % n, out, re.M), n


def test_pt_create_argument_checks():
    lib = pt_amd.hip_lib()
    h = C.c_void_p()
    assert lib.pt_create(None, 16, 16, C.byref(h)) == -1
    assert b"scene is null" in lib.pt_last_error(None)
    blob = pt_amd.SceneBlob()
    assert lib.pt_create(C.byref(blob), 0, 16, C.byref(h)) == -1
    # a mesh whose indices run past its vertices is rejected before any device work
    verts = np.zeros((3, 8), np.float32)
    idx = np.array([0, 1, 5], np.uint32)
    mesh = pt_amd.MeshDesc(0, 0, 1, 3)
    blob.vertices = verts.ctypes.data
    blob.vertex_count = 3
    blob.indices = idx.ctypes.data_as(C.POINTER(C.c_uint32))
    blob.index_count = 3
    blob.mesh_count = 1
    blob.meshes = C.pointer(mesh)
    assert lib.pt_create(C.byref(blob), 16, 16, C.byref(h)) == -1
    assert b"index out of range" in lib.pt_last_error(None)
    assert lib.pt_render(None, 1) == -1


def test_gltf_without_normals_gets_flat_normals(tmp_path):
    """A primitive without NORMAL loads with Assimp's generated face normals (the reference's
    import requests aiProcess_GenNormals, src/scene/scene.cpp:537-550; glTF 2.0 prescribes flat
    normals): every corner of a face carries (v1 - v0) x (v2 - v0) / |...|, the triangles'
    corners stay where they were, and JoinIdenticalVertices keeps one vertex per distinct
    (position, normal, uv): a corner is shared only by faces of one plane."""
    import json
    from pt_amd.scenes import generate
    generate.box_scene(str(tmp_path))
    doc = json.load(open(tmp_path / "box.gltf"))
    for m in doc["meshes"]:
        for p in m["primitives"]:
            p["attributes"].pop("NORMAL", None)
    json.dump(doc, open(tmp_path / "flat.gltf", "w"))
    (tmp_path / "flat.scn").write_text((tmp_path / "box.scn").read_text().replace("box.gltf", "flat.gltf"))
    ref = pt_amd.Scene(str(tmp_path / "box.scn")).arrays()
    got = pt_amd.Scene(str(tmp_path / "flat.scn")).arrays()
    assert len(ref["meshes"]) == len(got["meshes"])
    for (rv0, ri0, rn, _), (gv0, gi0, gn, gvn) in zip(ref["meshes"], got["meshes"]):
        assert rn == gn
        corners = {tuple(r) for r in got["vertices"][gv0 + got["indices"][gi0:gi0 + 3 * gn]].tolist()}
        assert gvn == len(corners)
        rv, gv = ref["vertices"], got["vertices"]
        rt = ref["indices"][ri0:ri0 + 3 * rn].reshape(-1, 3)
        gt = got["indices"][gi0:gi0 + 3 * gn].reshape(-1, 3)
        np.testing.assert_array_equal(gv[gv0 + gt, :3], rv[rv0 + rt, :3])  # same corners
        p = gv[gv0 + gt, :3]
        n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]).astype(np.float32)
        n /= np.linalg.norm(n, axis=1, keepdims=True).astype(np.float32)
        for k in range(3):
            np.testing.assert_allclose(gv[gv0 + gt[:, k], 3:6], n, atol=1e-6)
        if gn <= 2:  # the box's quads: their given normals are the face normals (up to side)
            np.testing.assert_allclose(np.abs((gv[gv0 + gt[:, 0], 3:6] * rv[rv0 + rt[:, 0], 3:6]).sum(1)), 1,
                                       atol=1e-5)


def _gltf_doc(buffers_bytes, accessors, views, prims):
    import base64
    return {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
        "meshes": [{"primitives": prims}],
        "materials": [{"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.5, 0.5, 1]}}],
        "buffers": [{"byteLength": len(buffers_bytes),
                     "uri": "data:application/octet-stream;base64," + base64.b64encode(buffers_bytes).decode()}],
        "bufferViews": views, "accessors": accessors,
    }


def test_gltf_modes_quantised_and_sparse_accessors(tmp_path):
    """glTF primitive modes as Assimp's glTF importer builds faces (strips (i, i+1, i+2) for even i
    and (i+1, i, i+2) for odd i, fans (0, i+1, i+2); points and lines are faces of fewer than three
    corners, which process_meshes leaves as all-zero triangles, scene.cpp:433-437), normalised
    integer texture coordinates and a sparse accessor. The four primitives share one material and
    vertex layout in one node, so OptimizeMeshes merges them: the later ones in its scan order
    (the fan, then the last primitive moved into the scanned slot, then the lines), the scanning
    one (the strip) last; JoinIdenticalVertices then numbers the distinct vertices by first use."""
    import json
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    uv8 = np.array([[0, 0], [255, 0], [0, 255], [255, 128]], np.uint8)  # normalised UNSIGNED_BYTE
    strip = np.array([0, 1, 2, 3], np.uint16)
    fan = np.array([0, 1, 3, 2], np.uint16)
    sp_idx = np.array([3], np.uint8)
    sp_val = np.array([[2, 2, 0]], np.float32)
    parts = [pos.tobytes(), nrm.tobytes(), uv8.tobytes(), strip.tobytes(), fan.tobytes(), sp_idx.tobytes() + b"\0\0\0",
             sp_val.tobytes()]
    offs = np.cumsum([0] + [len(p) for p in parts])
    views = [{"buffer": 0, "byteOffset": int(o), "byteLength": len(p)} for o, p in zip(offs, parts)]
    acc = [
        {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 2, "componentType": 5121, "normalized": True, "count": 4, "type": "VEC2"},
        {"bufferView": 3, "componentType": 5123, "count": 4, "type": "SCALAR"},
        {"bufferView": 4, "componentType": 5123, "count": 4, "type": "SCALAR"},
        {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3",
         "sparse": {"count": 1, "indices": {"bufferView": 5, "componentType": 5121},
                    "values": {"bufferView": 6}}},
    ]
    attrs = {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}
    prims = [dict(attributes=attrs, indices=3, mode=5, material=0),
             dict(attributes=attrs, indices=4, mode=6, material=0),
             dict(attributes=attrs, mode=1, material=0),
             dict(attributes=dict(attrs, POSITION=5), indices=3, mode=5, material=0)]
    json.dump(_gltf_doc(b"".join(parts), acc, views, prims), open(tmp_path / "m.gltf", "w"))
    (tmp_path / "m.scn").write_text("camera:\nobjects:\n m m.gltf\n\n m\n")
    a = pt_amd.Scene(str(tmp_path / "m.scn")).arrays()
    assert len(a["meshes"]) == 1
    v0, i0, n, nv = a["meshes"][0]
    tris = a["indices"][i0:i0 + 3 * n].reshape(-1, 3).tolist()
    assert tris == [[0, 1, 2], [0, 2, 3],    # fan over vertices 0 1 3 2: (0, 1, 3), (0, 3, 2)
                    [0, 1, 3], [3, 1, 4],    # sparse strip: (0, 1, 2), (2, 1, 3'), 3' moved to (2, 2, 0)
                    [0, 0, 0], [0, 0, 0],    # two lines
                    [0, 1, 3], [3, 1, 2]]    # strip: (0, 1, 2), (2, 1, 3)
    v = a["vertices"][v0:v0 + nv]
    order = [0, 1, 3, 2, 3]
    np.testing.assert_array_equal(v[:, :3], np.vstack([pos[[0, 1, 3, 2]], sp_val]))
    one = np.float32(1)
    uv = uv8[order].astype(np.float32) / np.float32(255)
    uv[:, 1] = one - (one - uv[:, 1])  # the importer's v flip and FlipUVs' flip back
    np.testing.assert_array_equal(v[:, 6:8], uv)


@pytest.mark.parametrize("edit, msg", [
    (lambda d: d["bufferViews"][0].update(byteStride=1e30), "glTF: bad byteStride"),
    (lambda d: d["bufferViews"][0].update(byteStride=2 ** 39), "glTF: accessor exceeds its buffer"),
    (lambda d: d["accessors"][0].update(count=2 ** 33), "glTF: accessor exceeds its buffer"),
    (lambda d: d["accessors"][0].update(byteOffset=-4), "glTF: bad byteOffset"),
    (lambda d: d["accessors"][0].update(count=2.5), "glTF: bad accessor count"),
    (lambda d: d["meshes"][0]["primitives"][0]["attributes"].update(POSITION=1e20), "glTF: missing accessor"),
    (lambda d: d["meshes"][0]["primitives"][0].update(material=-1), "glTF: bad material index"),
])
def test_gltf_sizes_and_indices_out_of_range_are_errors(tmp_path, edit, msg):
    """Counts, strides, offsets and indices that do not fit (huge, negative, fractional) are
    refused before any read: a stride near 2^63 once wrapped the bounds check of an accessor
    (found by a sanitizer fuzz run of the loader)."""
    import json
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": pos.nbytes}]
    acc = [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}]
    prims = [dict(attributes={"POSITION": 0}, material=0)]
    doc = _gltf_doc(pos.tobytes(), acc, views, prims)
    edit(doc)
    json.dump(doc, open(tmp_path / "m.gltf", "w"))
    (tmp_path / "m.scn").write_text("camera:\nobjects:\n m m.gltf\n\n m\n")
    with pytest.raises(pt_amd.PtError, match=re.escape(msg)):
        pt_amd.Scene(str(tmp_path / "m.scn"))
