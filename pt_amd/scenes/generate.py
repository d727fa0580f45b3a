"""Procedural test/benchmark scenes, written as glTF 2.0 + a .scn scene description.

The .scn files use the reference's scene-description syntax (README.md:25-52, parser
src/scene/scenebuild.cpp:132-277); the glTF files are the subset the host loader reads.
Every glTF keeps the reference's import stable: one primitive per material, distinct
materials, one node with an identity transform, no duplicate vertices, and every emissive
object is its own single-instance object with identical geometry (the reference's light
tables assume that, src/scene/lightsample.cpp:126 and src/scene/scenebuffer.cpp:129).

All randomness comes from the reference's generators (rng.py): scalars from
``Xshiro128.seeded(seed)``, bulk arrays from the Splitmix32 stream.

Scenes:
  box       256x256 correctness scene: 5 axis-aligned 2x2 quads, a level-3 icosphere of radius
            0.5 (roughness 0.5, metallic 0) and a 0.5x0.5 emissive ceiling quad of radiance 15.
  instanced 300 instances of 5 meshes (icospheres L2/L4/L6, a 512x512 noise terrain, a 200k
            triangle soup), per-instance materials, 20 emissive spheres (~3M triangles).
"""
import json
import os

import numpy as np

from .rng import Xshiro128, splitmix32_floats


def to_srgb(x):
    """src/colour.cpp:3-5 (the encode the reference pairs with from_srgb)."""
    x = float(x)
    return 12.92 * x if x <= 0.00031308 else 1.055 * x ** (1.0 / 2.4) - 0.055


# ------------------------------------------------------------------------------------------
# meshes


def icosphere(level, radius=1.0):
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    n = np.array(verts, dtype=np.float64)
    pos = (n * radius).astype(np.float32)
    uv = np.stack([0.5 + np.arctan2(n[:, 2], n[:, 0]) / (2 * np.pi),
                   np.arccos(np.clip(n[:, 1], -1, 1)) / np.pi], axis=1).astype(np.float32)
    return pos, n.astype(np.float32), uv, np.array(faces, dtype=np.uint32)


def quad(corner, edge_u, edge_v, normal):
    c = np.asarray(corner, dtype=np.float32)
    u = np.asarray(edge_u, dtype=np.float32)
    v = np.asarray(edge_v, dtype=np.float32)
    pos = np.stack([c, c + u, c + u + v, c + v]).astype(np.float32)
    nrm = np.tile(np.asarray(normal, dtype=np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    return pos, nrm, uv, idx


def merge(parts):
    pos, nrm, uv, idx = [], [], [], []
    base = 0
    for p, n, t, i in parts:
        pos.append(p)
        nrm.append(n)
        uv.append(t)
        idx.append(i + base)
        base += len(p)
    return (np.concatenate(pos), np.concatenate(nrm), np.concatenate(uv), np.concatenate(idx))


def _value_noise(x, z, lattice):
    n = lattice.shape[0]
    xi = np.floor(x).astype(np.int64)
    zi = np.floor(z).astype(np.int64)
    fx = x - xi
    fz = z - zi
    sx = fx * fx * (3 - 2 * fx)
    sz = fz * fz * (3 - 2 * fz)
    a = lattice[xi % n, zi % n]
    b = lattice[(xi + 1) % n, zi % n]
    c = lattice[xi % n, (zi + 1) % n]
    d = lattice[(xi + 1) % n, (zi + 1) % n]
    return (a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz


def terrain(res, seed, amplitude=0.25, octaves=4, half_size=1.0):
    """res x res quads over [-s,s]^2, height = 4-octave value noise (seeded lattices)."""
    g = np.linspace(-half_size, half_size, res + 1)
    x, z = np.meshgrid(g, g, indexing="ij")
    h = np.zeros_like(x)
    freq, amp = 4.0, 1.0
    for o in range(octaves):
        lattice = splitmix32_floats(seed + 7919 * o, 64 * 64).reshape(64, 64).astype(np.float64)
        h += amp * (_value_noise((x / half_size + 1) * freq, (z / half_size + 1) * freq, lattice) - 0.5)
        freq *= 2.0
        amp *= 0.5
    y = amplitude * h
    dx = np.gradient(y, g, axis=0)
    dz = np.gradient(y, g, axis=1)
    nrm = np.stack([-dx, np.ones_like(y), -dz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    pos = np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float32)
    uv = np.stack([(x / half_size + 1) / 2, (z / half_size + 1) / 2], axis=-1).reshape(-1, 2).astype(np.float32)
    r = res + 1
    i = np.arange(res)[:, None]
    j = np.arange(res)[None, :]
    v00 = (i * r + j).ravel()
    v10 = ((i + 1) * r + j).ravel()
    v01 = (i * r + j + 1).ravel()
    v11 = ((i + 1) * r + j + 1).ravel()
    idx = np.concatenate([np.stack([v00, v01, v11], 1), np.stack([v00, v11, v10], 1)]).astype(np.uint32)
    return pos, nrm.reshape(-1, 3).astype(np.float32), uv, idx


def triangle_soup(count, seed, lo=0.01, hi=0.5):
    """count independent triangles in [-1,1]^3, edge lengths log-uniform in [lo, hi]."""
    f = splitmix32_floats(seed, count * 10).astype(np.float64).reshape(count, 10)
    centre = f[:, 0:3] * 2.0 - 1.0
    size = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * f[:, 3])
    # two edge directions from spherical angles
    th1, ph1 = np.arccos(1 - 2 * f[:, 4]), 2 * np.pi * f[:, 5]
    th2, ph2 = np.arccos(1 - 2 * f[:, 6]), 2 * np.pi * f[:, 7]
    d1 = np.stack([np.sin(th1) * np.cos(ph1), np.sin(th1) * np.sin(ph1), np.cos(th1)], 1)
    d2 = np.stack([np.sin(th2) * np.cos(ph2), np.sin(th2) * np.sin(ph2), np.cos(th2)], 1)
    s2 = size * (0.5 + f[:, 8])
    a = centre
    b = centre + d1 * size[:, None]
    c = centre + d2 * s2[:, None]
    n = np.cross(b - a, c - a)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(ln > 1e-12, n / np.maximum(ln, 1e-30), np.array([0.0, 1.0, 0.0]))
    pos = np.stack([a, b, c], 1).reshape(-1, 3).astype(np.float32)
    nrm = np.repeat(n, 3, axis=0).astype(np.float32)
    uv = np.tile(np.array([[0, 0], [1, 0], [0, 1]], dtype=np.float32), (count, 1))
    idx = np.arange(count * 3, dtype=np.uint32).reshape(count, 3)
    return pos, nrm, uv, idx


# ------------------------------------------------------------------------------------------
# glTF writing


def _material(name, base_linear, roughness, metallic, emission=None, strength=None):
    m = {
        "name": name,
        "pbrMetallicRoughness": {
            "baseColorFactor": [to_srgb(base_linear[0]), to_srgb(base_linear[1]), to_srgb(base_linear[2]), 1.0],
            "roughnessFactor": float(roughness),
            "metallicFactor": float(metallic),
        },
    }
    if emission is not None:
        m["emissiveFactor"] = [float(e) for e in emission]
        m["extensions"] = {"KHR_materials_emissive_strength": {"emissiveStrength": float(strength)}}
    return m


def write_bin(path, mesh):
    """Writes positions | normals | uvs | indices; returns the layout for write_gltf."""
    pos, nrm, uv, idx = mesh
    parts = [np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(nrm, np.float32),
             np.ascontiguousarray(uv, np.float32), np.ascontiguousarray(idx, np.uint32)]
    offsets, off = [], 0
    with open(path, "wb") as f:
        for p in parts:
            b = p.tobytes()
            offsets.append((off, len(b)))
            f.write(b)
            off += len(b)
    return {"file": os.path.basename(path), "size": off, "offsets": offsets, "nv": len(pos), "ni": idx.size,
            "min": pos.min(0).tolist(), "max": pos.max(0).tolist()}


def write_png(path, px):
    """8-bit RGB or RGBA PNG (H x W x 3|4 uint8), no filtering, zlib-compressed."""
    import struct
    import zlib
    px = np.ascontiguousarray(px, np.uint8)
    h, w, n = px.shape
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    data = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2 if n == 3 else 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(data)


def write_gltf(path, prims, images=()):
    """prims: list of (bin_layout, material dict). One node, one mesh, one primitive each.
    images: PNG file names; texture i of the materials is image i."""
    doc = {"asset": {"version": "2.0", "generator": "pt_amd.scenes"}, "scene": 0,
           "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
           "meshes": [{"name": os.path.splitext(os.path.basename(path))[0], "primitives": []}],
           "materials": [], "buffers": [], "bufferViews": [], "accessors": []}
    if images:
        doc["images"] = [{"uri": im} for im in images]
        doc["textures"] = [{"source": i} for i in range(len(images))]
    buffers = {}
    for layout, mat in prims:
        if layout["file"] not in buffers:
            buffers[layout["file"]] = len(doc["buffers"])
            doc["buffers"].append({"uri": layout["file"], "byteLength": layout["size"]})
        b = buffers[layout["file"]]
        acc = []
        specs = [("VEC3", 5126, layout["nv"]), ("VEC3", 5126, layout["nv"]), ("VEC2", 5126, layout["nv"]),
                 ("SCALAR", 5125, layout["ni"])]
        for k, (typ, ct, count) in enumerate(specs):
            off, ln = layout["offsets"][k]
            doc["bufferViews"].append({"buffer": b, "byteOffset": off, "byteLength": ln})
            a = {"bufferView": len(doc["bufferViews"]) - 1, "componentType": ct, "count": int(count), "type": typ}
            if k == 0:
                a["min"], a["max"] = layout["min"], layout["max"]
            doc["accessors"].append(a)
            acc.append(len(doc["accessors"]) - 1)
        doc["materials"].append(mat)
        doc["meshes"][0]["primitives"].append({
            "attributes": {"POSITION": acc[0], "NORMAL": acc[1], "TEXCOORD_0": acc[2]},
            "indices": acc[3], "material": len(doc["materials"]) - 1, "mode": 4})
    doc["extensionsUsed"] = ["KHR_materials_emissive_strength"]
    with open(path, "w") as f:
        json.dump(doc, f)


def _fmt(v):
    return " ".join("%.9g" % float(x) for x in v)


# ------------------------------------------------------------------------------------------
# scenes


def box_scene(out_dir, seed=1):
    """The 256x256 correctness scene. Returns the .scn path.

    Albedos are linear: 0.7 is written as the glTF factor to_srgb(0.7), which the loader's
    sRGB decode turns back into 0.7 (src/scene/scene.cpp:94).
    """
    os.makedirs(out_dir, exist_ok=True)
    rng = Xshiro128.seeded(seed)
    col_a = [rng.uniform(0.2, 0.8) for _ in range(3)]
    col_b = [rng.uniform(0.2, 0.8) for _ in range(3)]
    white = merge([
        quad((-1, -2, 1), (2, 0, 0), (0, 0, -2), (0, 1, 0)),    # floor y=-2
        quad((-1, 0, -1), (2, 0, 0), (0, 0, 2), (0, -1, 0)),    # ceiling y=0
        quad((-1, -2, -1), (2, 0, 0), (0, 2, 0), (0, 0, 1)),    # back z=-1
    ])
    left = quad((-1, -2, 1), (0, 0, -2), (0, 2, 0), (1, 0, 0))     # x=-1
    right = quad((1, -2, -1), (0, 0, 2), (0, 2, 0), (-1, 0, 0))    # x=+1
    sp, sn, su, si = icosphere(3, 0.5)
    sphere = (sp + np.array([0, -1.5, 0], np.float32), sn, su, si)
    light = quad((-0.25, -0.001, -0.25), (0.5, 0, 0), (0, 0, 0.5), (0, -1, 0))
    prims = []
    for name, mesh, mat in [
        ("white", white, _material("white", (0.7, 0.7, 0.7), 1.0, 0.0)),
        ("left", left, _material("left", col_a, 1.0, 0.0)),
        ("right", right, _material("right", col_b, 1.0, 0.0)),
        ("sphere", sphere, _material("sphere", (0.7, 0.7, 0.7), 0.5, 0.0)),
        ("light", light, _material("light", (0.0, 0.0, 0.0), 1.0, 0.0, emission=(1, 1, 1), strength=15.0)),
    ]:
        prims.append((write_bin(os.path.join(out_dir, "box_%s.bin" % name), mesh), mat))
    write_gltf(os.path.join(out_dir, "box.gltf"), prims)
    scn = os.path.join(out_dir, "box.scn")
    with open(scn, "w") as f:
        f.write("camera:\n position 0 1 -3\n lookdir 0 0 1\n fov 60\n focusdist 3\n lensradius 0.03\n"
                " exposure 1\n envcol 0.375 0.45 0.9\n\nobjects:\n box box.gltf\n\n box\n")
    return scn


# instance counts per mesh kind for the instanced scene (~3M triangles over 300 instances)
INSTANCED_COUNTS = {"grid": 2, "soup": 4, "ico6": 10, "ico4": 40, "ico2": 224}
INSTANCED_EMITTERS = 20


def instanced_scene(out_dir, seed=2, counts=None, emitters=INSTANCED_EMITTERS, grid_res=512,
                    soup_tris=200000, camera_z=0.0):
    """300 instances of 5 seeded meshes with per-instance materials. Returns the .scn path.

    Each instance is its own .scn object (its own glTF file, one material) because a .scn
    instance cannot override materials (scenebuild.cpp:222-237); the glTF files of one mesh
    kind share that kind's .bin.
    """
    counts = dict(INSTANCED_COUNTS if counts is None else counts)
    os.makedirs(out_dir, exist_ok=True)
    rng = Xshiro128.seeded(seed)
    meshes = {
        "ico2": icosphere(2),
        "ico4": icosphere(4),
        "ico6": icosphere(6),
        "grid": terrain(grid_res, seed * 1000 + 17, amplitude=2.0, half_size=8.0),
        "soup": triangle_soup(soup_tris, seed * 1000 + 29),
    }
    layouts = {}
    for kind, mesh in meshes.items():
        if counts.get(kind, 0) > 0 or (kind == "ico2" and emitters > 0):
            layouts[kind] = write_bin(os.path.join(out_dir, "%s.bin" % kind), mesh)

    def placement():
        pos = [rng.uniform(-50.0, 50.0) for _ in range(3)]
        scale = rng.log_uniform(0.5, 4.0)
        # uniform axis on S^2, angle U(0, 2pi)
        z = rng.uniform(-1.0, 1.0)
        phi = rng.uniform(0.0, 2 * np.pi)
        r = (1 - z * z) ** 0.5
        axis = (r * np.cos(phi), r * np.sin(phi), z)
        angle = rng.uniform(0.0, 2 * np.pi)
        return pos, scale, axis, angle

    lines_decl, lines_inst = [], []
    k = 0
    order = [(kind, counts.get(kind, 0)) for kind in ("grid", "soup", "ico6", "ico4", "ico2")]
    for kind, n in order:
        for _ in range(n):
            base = [rng.uniform(0.1, 0.9) for _ in range(3)]
            rough = rng.uniform(0.05, 1.0)
            metal = 1.0 if rng.next_float() < 0.3 else 0.0
            name = "o%03d" % k
            write_gltf(os.path.join(out_dir, name + ".gltf"),
                       [(layouts[kind], _material(name, base, rough, metal))])
            pos, scale, axis, angle = placement()
            lines_decl.append(" %s %s.gltf" % (name, name))
            lines_inst.append(" %s\n  s %s\n  r %s %.9g\n  t %s" % (name, _fmt([scale] * 3), _fmt(axis), angle,
                                                                     _fmt(pos)))
            k += 1
    for _ in range(emitters):
        base = [rng.uniform(0.1, 0.9) for _ in range(3)]
        strength = rng.uniform(5.0, 50.0)
        name = "e%03d" % k
        write_gltf(os.path.join(out_dir, name + ".gltf"),
                   [(layouts["ico2"], _material(name, base, 1.0, 0.0, emission=(1, 1, 1), strength=strength))])
        pos, scale, axis, angle = placement()
        lines_decl.append(" %s %s.gltf" % (name, name))
        lines_inst.append(" %s\n  s %s\n  r %s %.9g\n  t %s" % (name, _fmt([scale] * 3), _fmt(axis), angle, _fmt(pos)))
        k += 1
    scn = os.path.join(out_dir, "instanced.scn")
    with open(scn, "w") as f:
        # The eye sits inside the instance cloud (the .scn position is negated by the
        # camera's view matrix, src/scene/camera.h:16) and looks along +z.
        f.write("camera:\n position 0 0 %.9g\n lookdir 0 0 -1\n fov 60\n focusdist 30\n lensradius 0.03\n"
                " exposure 1\n envcol 0.5 0.55 0.6\n\nobjects:\n" % (-camera_z))
        f.write("\n".join(lines_decl) + "\n\n" + "\n".join(lines_inst) + "\n")
    return scn


MESH_TRIS = {"grid": 2 * 512 * 512, "soup": 200000, "ico6": 81920, "ico4": 5120, "ico2": 320}


def counts_for_triangles(total):
    """Instance counts of the five mesh kinds whose triangles (with the 20 emissive
    level-2 spheres) add up to at least `total` and less than 320 more: the default scene's mix
    of kinds, scaled, the remainder filled with the largest spheres that fit and one more
    level-2 sphere for what is left."""
    default_total = sum(INSTANCED_COUNTS[k] * MESH_TRIS[k] for k in INSTANCED_COUNTS)
    left = max(0, int(total) - INSTANCED_EMITTERS * MESH_TRIS["ico2"])
    counts = {}
    for kind in ("grid", "soup", "ico6", "ico4"):
        share = INSTANCED_COUNTS[kind] * MESH_TRIS[kind] / default_total
        counts[kind] = int(left * share // MESH_TRIS[kind])
    counts["ico2"] = 0
    rem = left - sum(counts[k] * MESH_TRIS[k] for k in counts)
    for kind in ("ico6", "ico4", "ico2"):  # the remainder, largest meshes first
        n = rem // MESH_TRIS[kind]
        counts[kind] += n
        rem -= n * MESH_TRIS[kind]
    if rem > 0:
        counts["ico2"] += 1
    counts["ico2"] = max(1, counts["ico2"])
    return counts


def textured_scene(out_dir, seed=3):
    """The box scene with textured materials: a checker floor (sRGB base-colour map repeated 4x,
    metallic-roughness map, normal map), a back wall with an RGBA base-colour map, a floating
    quad whose base-colour map has alpha stripes (stochastic transparency and alpha shadows),
    and a ceiling light with an emission map. Returns the .scn path."""
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    n = 64
    yy, xx = np.mgrid[0:n, 0:n]
    checker = (((xx // 8) + (yy // 8)) % 2).astype(np.float64)
    base = np.stack([40 + 180 * checker, 60 + 120 * (1 - checker), 90 + 40 * checker], -1)
    base = (base + rng.integers(0, 30, (n, n, 3))).clip(0, 255).astype(np.uint8)
    write_png(os.path.join(out_dir, "floor_base.png"), base)
    mr = np.stack([np.zeros((n, n)), 60 + 190 * (xx / (n - 1)), 255 * (yy > n // 2)], -1).astype(np.uint8)
    write_png(os.path.join(out_dir, "floor_mr.png"), mr)
    bump = np.stack([0.35 * np.sin(xx * 2 * np.pi / 16), 0.35 * np.cos(yy * 2 * np.pi / 16), np.ones((n, n))], -1)
    bump = bump / np.linalg.norm(bump, axis=-1, keepdims=True)
    write_png(os.path.join(out_dir, "floor_normal.png"), ((bump * 0.5 + 0.5) * 255).round().astype(np.uint8))
    wall = np.stack([xx * 4, yy * 4, 255 - xx * 2, np.full((n, n), 255)], -1).clip(0, 255).astype(np.uint8)
    write_png(os.path.join(out_dir, "wall_base.png"), wall)
    stripes = np.stack([np.full((n, n), 230), np.full((n, n), 120), np.full((n, n), 40),
                        np.where((xx // 6) % 2 == 0, 255, np.where((xx // 6) % 3 == 0, 0, 90))], -1).astype(np.uint8)
    write_png(os.path.join(out_dir, "veil_base.png"), stripes)
    glow = np.stack([200 + 55 * checker, 180 + 40 * (1 - checker), 120 + 100 * (xx / (n - 1))], -1).astype(np.uint8)
    write_png(os.path.join(out_dir, "light_emission.png"), glow)

    floor = quad((-1, -2, 1), (2, 0, 0), (0, 0, -2), (0, 1, 0))
    floor = (floor[0], floor[1], floor[2] * 4.0, floor[3])  # REPEAT addressing
    back = quad((-1, -2, -1), (2, 0, 0), (0, 2, 0), (0, 0, 1))
    ceil = quad((-1, 0, -1), (2, 0, 0), (0, 0, 2), (0, -1, 0))
    left = quad((-1, -2, 1), (0, 0, -2), (0, 2, 0), (1, 0, 0))
    right = quad((1, -2, -1), (0, 0, 2), (0, 2, 0), (-1, 0, 0))
    veil = quad((-0.6, -0.8, -0.5), (1.2, 0, 0), (0, 0, 1.0), (0, 1, 0))
    sp, sn, su, si = icosphere(3, 0.45)
    sphere = (sp + np.array([0.35, -1.55, -0.2], np.float32), sn, su, si)
    light = quad((-0.3, -0.001, -0.3), (0.6, 0, 0), (0, 0, 0.6), (0, -1, 0))
    images = ["floor_base.png", "floor_mr.png", "floor_normal.png", "wall_base.png", "veil_base.png",
              "light_emission.png"]

    def tex(i):
        return {"index": i}
    m_floor = {"name": "floor", "pbrMetallicRoughness": {"baseColorTexture": tex(0), "metallicRoughnessTexture": tex(1)},
               "normalTexture": tex(2)}
    m_back = {"name": "back", "pbrMetallicRoughness": {"baseColorTexture": tex(3), "roughnessFactor": 0.8,
                                                        "metallicFactor": 0.0}}
    m_veil = {"name": "veil", "pbrMetallicRoughness": {"baseColorTexture": tex(4), "roughnessFactor": 0.6,
                                                        "metallicFactor": 0.0}}
    m_light = {"name": "light", "pbrMetallicRoughness": {"baseColorFactor": [0, 0, 0, 1], "roughnessFactor": 1.0,
                                                         "metallicFactor": 0.0},
               "emissiveTexture": tex(5), "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 15.0}}}
    prims = []
    for name, mesh, mat in [
        ("floor", floor, m_floor), ("back", back, m_back),
        ("ceil", ceil, _material("ceil", (0.7, 0.7, 0.7), 1.0, 0.0)),
        ("left", left, _material("left", (0.6, 0.25, 0.2), 1.0, 0.0)),
        ("right", right, _material("right", (0.2, 0.4, 0.6), 1.0, 0.0)),
        ("veil", veil, m_veil),
        ("sphere", sphere, _material("sphere", (0.7, 0.7, 0.7), 0.4, 1.0)),
        ("light", light, m_light),
    ]:
        prims.append((write_bin(os.path.join(out_dir, "tex_%s.bin" % name), mesh), mat))
    write_gltf(os.path.join(out_dir, "textured.gltf"), prims, images)
    scn = os.path.join(out_dir, "textured.scn")
    with open(scn, "w") as f:
        f.write("camera:\n position 0 1 -3\n lookdir 0 0 1\n fov 60\n exposure 1\n envcol 0.375 0.45 0.9\n\n"
                "objects:\n tex textured.gltf\n\n tex\n")
    return scn


def cached_scene(kind, root=None, **kw):
    """Generates a scene once under root (default $PT_SCENE_DIR or /tmp/pt_scenes)."""
    root = root or os.environ.get("PT_SCENE_DIR", "/tmp/pt_scenes")
    def fmt(v):
        return "-".join("%s%s" % kv for kv in sorted(v.items())) if isinstance(v, dict) else str(v)
    tag = kind + "".join("_%s%s" % (k, fmt(v)) for k, v in sorted(kw.items()))
    out = os.path.join(root, tag)
    name = {"box": "box.scn", "instanced": "instanced.scn", "textured": "textured.scn"}[kind]
    path = os.path.join(out, name)
    done = os.path.join(out, ".done")
    if not os.path.exists(done):
        fn = {"box": box_scene, "instanced": instanced_scene, "textured": textured_scene}[kind]
        fn(out, **kw)
        open(done, "w").close()
    return path
