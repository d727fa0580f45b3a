"""The glTF import's post-processing (the reference's Assimp flags, src/scene/scene.cpp:537-550).

Pinned three ways:
* arithmetic: baked vertices, normals, instance transforms and generated face normals equal, bit
  for bit, what Assimp 5.4.3's own inline math (vendored by the reference) computes for the same
  inputs (tests/golden/assimp_math.json, from oracle/assimp_harness.cpp);
* invariants: the world-space triangles of every material are the same set before and after the
  steps for random node hierarchies (nothing lost, doubled or misplaced by collapsing the graph,
  joining nodes, merging meshes or replacing instances);
* the documented behaviour of each step (include/assimp/postprocess.h), case by case.
Which meshes merge and in which order is Assimp's (its step sources are not in the reference):
parity of that numbering is unpinned beyond these cases.
"""
import base64
import json
import os

import numpy as np
import pytest

import pt_amd

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "assimp_math.json")))


def bits(v):
    return np.array(v, np.uint32).view(np.float32)


class Doc:
    """A glTF document with one embedded buffer."""

    def __init__(self):
        self.blob = bytearray()
        self.d = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": []}], "nodes": [], "meshes": [],
                  "materials": [], "bufferViews": [], "accessors": []}

    def accessor(self, a, typ):
        a = np.ascontiguousarray(a)
        while len(self.blob) % 4:
            self.blob.append(0)
        self.d["bufferViews"].append({"buffer": 0, "byteOffset": len(self.blob), "byteLength": a.nbytes})
        self.blob += a.tobytes()
        ct = {np.dtype(np.float32): 5126, np.dtype(np.uint32): 5125}[a.dtype]
        self.d["accessors"].append({"bufferView": len(self.d["bufferViews"]) - 1, "componentType": ct,
                                    "count": int(a.shape[0]), "type": typ})
        return len(self.d["accessors"]) - 1

    def material(self, **m):
        self.d["materials"].append(m)
        return len(self.d["materials"]) - 1

    def prim(self, pos, nrm=None, uv=None, material=None, idx=None):
        attrs = {"POSITION": self.accessor(np.asarray(pos, np.float32).reshape(-1, 3), "VEC3")}
        if nrm is not None:
            attrs["NORMAL"] = self.accessor(np.asarray(nrm, np.float32).reshape(-1, 3), "VEC3")
        if uv is not None:
            attrs["TEXCOORD_0"] = self.accessor(np.asarray(uv, np.float32).reshape(-1, 2), "VEC2")
        p = {"attributes": attrs}
        if idx is not None:
            p["indices"] = self.accessor(np.asarray(idx, np.uint32).reshape(-1), "SCALAR")
        if material is not None:
            p["material"] = material
        return p

    def mesh(self, prims):
        self.d["meshes"].append({"primitives": prims})
        return len(self.d["meshes"]) - 1

    def node(self, **n):
        self.d["nodes"].append(n)
        return len(self.d["nodes"]) - 1

    def load(self, tmp_path, name="m"):
        self.d["buffers"] = [{"byteLength": len(self.blob), "uri": "data:application/octet-stream;base64," +
                              base64.b64encode(bytes(self.blob)).decode()}]
        (tmp_path / (name + ".gltf")).write_text(json.dumps(self.d))
        (tmp_path / (name + ".scn")).write_text("camera:\nobjects:\n %s %s.gltf\n\n %s\n" % (name, name, name))
        return pt_amd.Scene(str(tmp_path / (name + ".scn"))).arrays()


def meshes_of(a):
    """Per mesh: vertices (n x 8) and triangles (m x 3) of the packed scene."""
    out = []
    for v0, i0, n, nv in a["meshes"]:
        out.append((a["vertices"][v0:v0 + nv], a["indices"][i0:i0 + 3 * n].reshape(-1, 3)))
    return out


def instances_of(a):
    """(mesh, material, 4x4 transform) per instance record (InstanceData, scenebuffer.h:13-20: the
    transform's transpose as glm stores it, so its words read row by row are the transform's rows)."""
    rec = a["instances"].view(np.uint32).reshape(-1, 32)
    return [(int(a["instance_mesh"][k]), int(rec[k, 3]), rec[k, 4:20].view(np.float32).reshape(4, 4).copy())
            for k in range(len(rec))]


def trs_of(c):
    return {"translation": bits(c["t"]).tolist(), "rotation": bits(c["q"]).tolist(), "scale": bits(c["s"]).tolist()}


# --- arithmetic against Assimp's own inline math ----------------------------------------------

def test_collapsed_hierarchy_matches_assimp_arithmetic(tmp_path):
    """root R -> A (mesh 0), B -> C (mesh 1), D1 and D2 (both mesh 2). OptimizeGraph pulls A, C,
    D1, D2 under the root (R x A, R x (B x C), ...) and joins C into A (C's triangle baked by
    inverse(R x A) x R x (B x C), its normals by the inverse transpose); meshes 0 and 1 have
    materials equal but for their names, so RemoveRedundantMaterials makes them one and
    OptimizeMeshes merges them, C's triangle first; mesh 2 is instanced and stays a mesh of its
    own, listed first."""
    g = GOLDEN
    d = Doc()
    tris = g["tris"]
    uv = [[0.1, 0.2], [0.7, 0.3], [0.4, 0.9]]
    m0 = d.material(name="first", pbrMetallicRoughness={"baseColorFactor": [0.5, 0.4, 0.3, 1.0]})
    m1 = d.material(name="second", pbrMetallicRoughness={"baseColorFactor": [0.5, 0.4, 0.3, 1.0]})
    m2 = d.material(name="third", pbrMetallicRoughness={"baseColorFactor": [0.2, 0.4, 0.3, 1.0]})
    mesh = [d.mesh([d.prim(bits(tris[t]["pos"]), bits(tris[t]["nrm"]), uv, m)])
            for t, m in (("t0", m0), ("t1", m1), ("t2", m2))]
    n = g["nodes"]
    a = d.node(name="A", mesh=mesh[0], **trs_of(n["A"]))
    c = d.node(name="C", mesh=mesh[1], **trs_of(n["C"]))
    b = d.node(name="B", children=[c], **trs_of(n["B"]))
    d1 = d.node(name="D1", mesh=mesh[2], **trs_of(n["D1"]))
    d2 = d.node(name="D2", mesh=mesh[2], **trs_of(n["D2"]))
    r = d.node(name="R", children=[a, b, d1, d2], **trs_of(n["R"]))
    d.d["scenes"][0]["nodes"] = [r]
    arr = d.load(tmp_path)

    ms = meshes_of(arr)
    assert len(ms) == 2 and len(arr["materials"]) == 2
    v, t = ms[1]
    assert t.tolist() == [[0, 1, 2], [3, 4, 5]]
    np.testing.assert_array_equal(v[:, :3].reshape(-1).view(np.uint32), np.array(g["merged_pos"], np.uint32))
    np.testing.assert_array_equal(v[:, 3:6].reshape(-1).view(np.uint32), np.array(g["merged_nrm"], np.uint32))
    v2, t2 = ms[0]
    np.testing.assert_array_equal(v2[:, :3].reshape(-1), bits(tris["t2"]["pos"]))
    inst = instances_of(arr)
    assert [(m, mat) for m, mat, _ in inst] == [(0, 1), (0, 1), (1, 0)]
    for (_, _, tr), key in zip(inst, ("instance_D1", "instance_D2", "instance_A")):
        np.testing.assert_array_equal(tr, bits(g[key]).reshape(4, 4))


def test_generated_face_normals_match_assimp_arithmetic(tmp_path):
    """GenNormals for triangles without NORMAL: ((v1 - v0) ^ (v2 - v0)).NormalizeSafe() on every
    corner (a multiplication by the reciprocal length, not a division)."""
    d = Doc()
    cases = [c for c in GOLDEN["face_normals"] if np.any(bits(c["normal"]) != 0)]
    prims = [d.prim(bits(c["pos"]), uv=[[0, 0], [1, 0], [0, 1]],
                    material=d.material(pbrMetallicRoughness={"baseColorFactor": [0.01 * k, 0.5, 0.5, 1]}))
             for k, c in enumerate(cases)]
    d.node(mesh=d.mesh(prims))
    d.d["scenes"][0]["nodes"] = [0]
    ms = meshes_of(d.load(tmp_path))
    assert len(ms) == len(cases)
    for (v, t), c in zip(ms, cases):
        assert t.tolist() == [[0, 1, 2]]
        np.testing.assert_array_equal(v[:, :3].reshape(-1), bits(c["pos"]))
        for k in range(3):
            np.testing.assert_array_equal(v[k, 3:6].view(np.uint32), np.array(c["normal"], np.uint32))


# --- invariant: the same world-space triangles per material ------------------------------------

def _quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(np.float32)


def _matrix(node):
    t = np.eye(4)
    if "translation" in node:
        t[:3, 3] = node["translation"]
    r = np.eye(4)
    if "rotation" in node:
        x, y, z, w = node["rotation"]
        r[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    s = np.diag(list(node.get("scale", [1, 1, 1])) + [1.0])
    return t @ r @ s


def _world_triangles(doc, material_names):
    """{material name: sorted (n, 3, 3) world-space triangles} straight from the glTF."""
    out = {}

    def walk(i, parent):
        node = doc["nodes"][i]
        m = parent @ _matrix(node)
        if "mesh" in node:
            for p in doc["meshes"][node["mesh"]]["primitives"]:
                pos = _accessor(doc, p["attributes"]["POSITION"]).reshape(-1, 3).astype(np.float64)
                idx = _accessor(doc, p["indices"]).reshape(-1, 3)
                w = pos @ m[:3, :3].T + m[:3, 3]
                out.setdefault(material_names[p["material"]], []).extend(w[idx])
        for c in node.get("children", []):
            walk(c, m)

    for r in doc["scenes"][0]["nodes"]:
        walk(r, np.eye(4))
    return {k: np.array(v) for k, v in out.items()}


_BLOB = {}


def _accessor(doc, i):
    a = doc["accessors"][i]
    v = doc["bufferViews"][a["bufferView"]]
    raw = _BLOB["b"][v["byteOffset"]:v["byteOffset"] + v["byteLength"]]
    return np.frombuffer(raw, np.float32 if a["componentType"] == 5126 else np.uint32)


def _same_triangles(got, want, tol):
    """Largest corner distance between the two sets of triangles, paired by their centroids
    (an optimal assignment) with corners paired in the best of the six orders (winding left out)."""
    from itertools import permutations

    from scipy.optimize import linear_sum_assignment
    assert got.shape == want.shape
    cg, cw = got.mean(axis=1), want.mean(axis=1)
    rows, cols = linear_sum_assignment(np.linalg.norm(cg[:, None] - cw[None], axis=2))
    worst = 0.0
    for r, c in zip(rows, cols):
        worst = max(worst, min(np.abs(got[r][list(p)] - want[c]).max() for p in permutations(range(3))))
    assert worst <= tol, worst


@pytest.mark.parametrize("seed", range(6))
def test_world_geometry_is_unchanged_by_the_steps(tmp_path, seed):
    """Random hierarchies (nested nodes, nodes without meshes, a shared mesh, mirroring scales,
    three materials of which two are equal but for their names): the world-space triangles of
    every material are the same set after the steps as in the file, within float rounding."""
    rng = np.random.default_rng(seed)
    d = Doc()
    mats = [d.material(name="red", pbrMetallicRoughness={"baseColorFactor": [1, 0, 0, 1]}),
            d.material(name="red too", pbrMetallicRoughness={"baseColorFactor": [1, 0, 0, 1]}),
            d.material(name="blue", pbrMetallicRoughness={"baseColorFactor": [0, 0, 1, 1]})]
    names = {mats[0]: "red", mats[1]: "red", mats[2]: "blue"}
    meshes = []
    for k in range(5):
        prims = []
        for _ in range(int(rng.integers(1, 3))):
            nt = int(rng.integers(1, 4))
            pos = rng.uniform(-1, 1, (nt * 3, 3)).astype(np.float32)
            nrm = rng.normal(size=(nt * 3, 3)).astype(np.float32)
            uv = rng.uniform(0, 1, (nt * 3, 2)).astype(np.float32)
            prims.append(d.prim(pos, nrm, uv, int(rng.integers(0, 3)), np.arange(nt * 3)))
        meshes.append(d.mesh(prims))

    def trs():
        s = rng.uniform(0.5, 2.0, 3)
        if rng.random() < 0.3:
            s[int(rng.integers(0, 3))] *= -1
        return {"translation": rng.uniform(-3, 3, 3).astype(np.float32).tolist(),
                "rotation": _quat(rng).tolist(), "scale": s.astype(np.float32).tolist()}

    def make(depth):
        kids = [make(depth + 1) for _ in range(int(rng.integers(0, 3)) if depth < 3 else 0)]
        n = trs()
        if kids:
            n["children"] = kids
        if rng.random() < 0.7:
            n["mesh"] = meshes[int(rng.integers(0, len(meshes)))]
        return d.node(**n)

    roots = [make(0) for _ in range(3)]
    d.node(mesh=meshes[0], **trs())  # mesh 0 used at least twice
    roots.append(len(d.d["nodes"]) - 1)
    d.d["scenes"][0]["nodes"] = roots
    arr = d.load(tmp_path)
    _BLOB["b"] = bytes(d.blob)
    want = _world_triangles(d.d, names)

    ms = meshes_of(arr)
    got = {}
    base = arr["materials"][:, 0]
    for mesh, mat, tr in instances_of(arr):
        v, t = ms[mesh]
        w = v[:, :3].astype(np.float64) @ tr[:3, :3].T.astype(np.float64) + tr[:3, 3]
        key = "red" if base[mat] > 0.5 else "blue"
        got.setdefault(key, []).extend(w[t])
    assert sorted(got) == sorted(want)
    for k in want:
        _same_triangles(np.array(got[k]), want[k], 1e-3)


# --- each step's documented behaviour ----------------------------------------------------------

def _tri(rng):
    return rng.uniform(-1, 1, (3, 3)).astype(np.float32)


def test_materials_equal_but_for_their_names_are_one_and_unused_ones_go(tmp_path):
    """RemoveRedundantMaterials: a material nothing uses is dropped; materials whose properties
    are equal (explicit glTF defaults equal to absent values, the name left out) become the first
    of them; a present extension is a property of its own."""
    rng = np.random.default_rng(1)
    d = Doc()
    d.material(name="unused", pbrMetallicRoughness={"baseColorFactor": [0.9, 0.1, 0.1, 1]})
    a = d.material(name="a")
    b = d.material(name="b", pbrMetallicRoughness={"metallicFactor": 1.0, "roughnessFactor": 1,
                                                  "baseColorFactor": [1, 1, 1, 1]}, alphaMode="OPAQUE")
    c = d.material(name="c", extensions={"KHR_materials_emissive_strength": {"emissiveStrength": 1.0}})
    uv = [[0, 0], [1, 0], [0, 1]]
    prims = [d.prim(_tri(rng), None, uv, m) for m in (a, b, c)]
    d.node(mesh=d.mesh(prims))
    d.node(mesh=d.mesh([d.prim(_tri(rng), None, uv)]))  # no material: the importer's default one
    d.d["scenes"][0]["nodes"] = [0, 1]
    arr = d.load(tmp_path)
    assert len(arr["materials"]) == 2  # {a, b, default}, {c}
    ms = meshes_of(arr)
    # a and b merged into one mesh in node 0; c its own; the default-material mesh joined into node 0
    # by OptimizeGraph and then merged with a/b (equal material)
    tri_counts = sorted(len(t) for _, t in ms)
    assert tri_counts == [1, 3]


def test_identical_meshes_become_instances_of_the_first(tmp_path):
    """FindInstances: two glTF meshes holding the same triangles with the same material are one
    mesh referenced by both nodes, so OptimizeGraph keeps both nodes and their transforms."""
    rng = np.random.default_rng(2)
    d = Doc()
    m = d.material()
    tri, nrm, uv = _tri(rng), rng.normal(size=(3, 3)), [[0, 0], [1, 0], [0, 1]]
    m1 = d.mesh([d.prim(tri, nrm, uv, m)])
    m2 = d.mesh([d.prim(tri, nrm, uv, m)])
    d.node(mesh=m1, translation=[1, 0, 0])
    d.node(mesh=m2, translation=[0, 2, 0])
    d.d["scenes"][0]["nodes"] = [0, 1]
    arr = d.load(tmp_path)
    assert len(arr["meshes"]) == 1
    inst = instances_of(arr)
    assert [i[0] for i in inst] == [0, 0]
    np.testing.assert_array_equal([i[2][:3, 3] for i in inst], [[1, 0, 0], [0, 2, 0]])


def test_light_and_camera_nodes_are_kept(tmp_path):
    """OptimizeGraph keeps nodes named by a camera or a light: their meshes stay in their own frame
    instead of being joined into the first node's."""
    rng = np.random.default_rng(3)
    d = Doc()
    m = d.material()
    uv = [[0, 0], [1, 0], [0, 1]]
    meshes = [d.mesh([d.prim(_tri(rng), None, uv, m)]) for _ in range(3)]
    d.node(name="plain", mesh=meshes[0], translation=[1, 0, 0])
    d.node(name="lamp", mesh=meshes[1], translation=[0, 1, 0], extensions={"KHR_lights_punctual": {"light": 0}})
    d.node(name="eye", mesh=meshes[2], translation=[0, 0, 1], camera=0)
    d.d["cameras"] = [{"type": "perspective", "perspective": {"yfov": 1.0, "znear": 0.1}}]
    d.d["extensions"] = {"KHR_lights_punctual": {"lights": [{"type": "point"}]}}
    d.d["scenes"][0]["nodes"] = [0, 1, 2]
    arr = d.load(tmp_path)
    inst = instances_of(arr)
    assert len(inst) == 3
    np.testing.assert_array_equal(sorted(i[2][:3, 3].tolist() for i in inst), [[0, 0, 1], [0, 1, 0], [1, 0, 0]])

    # without the light and the camera all three nodes are leaves of the root: one merged mesh
    for n in d.d["nodes"]:
        n.pop("extensions", None)
        n.pop("camera", None)
    d.d.pop("cameras")
    d.d.pop("extensions")
    arr = d.load(tmp_path, "n")
    assert len(arr["meshes"]) == 1 and len(instances_of(arr)) == 1


def test_invalid_texture_coordinates_and_normals(tmp_path):
    """FindInvalidData: a texture-coordinate set whose used values are all equal is dropped (the
    reference then reads zero); normals holding a zero vector are dropped and GenNormals makes face
    normals."""
    d = Doc()
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    mats = [d.material(pbrMetallicRoughness={"baseColorFactor": [0.1 * k, 0, 0, 1]}) for k in range(3)]
    p0 = d.prim(tri, [[0, 0, 1]] * 3, [[0.5, 0.25]] * 3, mats[0])
    p1 = d.prim(tri, [[0, 0, 1], [0, 0, 0], [0, 0, 1]], [[0, 0], [1, 0], [0, 1]], mats[1])
    p2 = d.prim(tri, [[0, 0, 1], [0, 1, 0], [1, 0, 0]], [[0, 0], [1, 0], [0, 1]], mats[2])
    d.node(mesh=d.mesh([p0, p1, p2]))
    d.d["scenes"][0]["nodes"] = [0]
    ms = meshes_of(d.load(tmp_path))
    np.testing.assert_array_equal(ms[0][0][:, 6:8], 0)          # constant uvs dropped
    np.testing.assert_array_equal(ms[1][0][:, 3:6], [[0, 0, 1]] * 3)  # face normal (0, 0, 1)
    np.testing.assert_array_equal(ms[2][0][:, 3:6], [[0, 0, 1], [0, 1, 0], [1, 0, 0]])  # kept


def test_coincident_corners_make_zero_triangles(tmp_path):
    """FindDegenerates without removal: corners at equal positions (== on floats, so -0 equals 0)
    reduce a face to a line or a point, which process_meshes leaves as an all-zero triangle in
    its slot."""
    d = Doc()
    pos = np.array([[0, 0, 0], [1, 0, 0], [-0.0, 0, 0], [0, 1, 0], [2, 0, 0], [2, 1, 0]], np.float32)
    d.node(mesh=d.mesh([d.prim(pos, None, [[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0], [0.5, 1]], d.material(),
                               idx=[0, 1, 3, 0, 2, 3, 1, 4, 5])]))
    d.d["scenes"][0]["nodes"] = [0]
    (v, t), = meshes_of(d.load(tmp_path))
    assert t[1].tolist() == [0, 0, 0]
    assert t[0].tolist() != [0, 0, 0] and t[2].tolist() != [0, 0, 0]
