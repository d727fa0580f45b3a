"""Checks of the oracle's estimator (no GPU): what the mathematics fixes, independently of how
the oracle computes it.

The oracle restates the reference's estimator (pathtrace.rg.hlsl, pathtrace.ch.hlsl,
lightsample.hlsli, pathtrace.hlsli): next-event estimation with a balance-heuristic MIS
weight against the BSDF pdf (pathtrace.ch.hlsl:159-225, 243-271), light sampling through two
alias tables with an area-to-solid-angle conversion (lightsample.hlsli:36-108), and Russian
roulette (pathtrace.hlsli:118-129). Equal images of the device and the oracle only mean
something if the oracle's estimator is itself right, so it is pinned here by:

  * equivalence: the NEE + MIS estimator and the BSDF-only estimator the reference builds
    without SAMPLE_LIGHTS (pathtrace.ch.hlsl:8, 266-268) converge to the same image, and so do
    the estimators with and without Russian roulette. A wrong solid-angle conversion, MIS
    weight or roulette weight biases one side;
  * closed forms: the expected one-bounce radiance of a plane under a uniform environment is
    the environment times the BSDF's directional albedo, integrated by quadrature of
    evaluate_brdf over the hemisphere (not by its sampling routine); the light sampler's
    1/pdf averages to the solid angle the emitter subtends (Van Oosterom & Strackee's formula),
    its points are uniform over the emitter's area, and light_pdf at an emitter triangle is
    1/area of the emitter;
  * planted mistakes (oracle_render_ex's ORACLE_PLANT_* flags: a dropped d^2, swapped MIS
    weight, the roulette weight divided by q) each fail the corresponding check.

Sample batches are successive index ranges of the same Owen-scrambled Sobol sequence; their
spread estimates the error of the mean (conservatively: scrambled nets vary less than
independent samples).
"""
import os

import numpy as np
import pytest

import pt_amd
import oracle_bind as ob
from pt_amd.scenes import generate

W = H = 32
K = 16       # sample batches
SPP = 32     # samples per pixel per batch
BLOCK = 8    # z-test over BLOCK x BLOCK pixel blocks


def batch_means(orc, cam, w, h, spp, k, depth, flags):
    """k images, image j the mean of samples [j * spp, (j + 1) * spp) of every pixel, from the
    oracle's running mean (pathtrace.rg.hlsl:61-67) after each batch."""
    acc = np.zeros((h, w, 4), np.float32)
    out, prev = [], np.zeros((h, w, 3))
    for j in range(k):
        orc.render(cam, w, h, spp, depth, sample_index=j * spp, accum=acc, flags=flags)
        cur = acc[..., :3].astype(np.float64)
        out.append((j + 1) * cur - j * prev)
        prev = cur
    return np.stack(out)


def block_z(a, b, block=BLOCK):
    """z scores of the difference of two estimators per pixel block and for the whole image."""
    def blocks(x):
        k, h, w, c = x.shape
        return x.reshape(k, h // block, block, w // block, block, c).mean(axis=(2, 4)).reshape(k, -1)

    def z(xa, xb):
        ma, mb = xa.mean(0), xb.mean(0)
        se = np.sqrt(xa.var(0, ddof=1) / len(xa) + xb.var(0, ddof=1) / len(xb))
        return (ma - mb) / np.maximum(se, 1e-12 + 1e-6 * np.abs(mb))
    zb = z(blocks(a), blocks(b))
    zi = z(a.reshape(len(a), -1).mean(1, keepdims=True), b.reshape(len(b), -1).mean(1, keepdims=True))
    return np.abs(zb), float(abs(zi[0]))


@pytest.fixture(scope="module")
def box(box_scn):
    scene = pt_amd.Scene(box_scn)
    return scene, scene.camera(W, H), ob.Oracle(scene)


@pytest.fixture(scope="module")
def nee_mis(box):
    scene, cam, orc = box
    return batch_means(orc, cam, W, H, SPP, K, 3, 0)


def test_nee_mis_equals_bsdf_only_estimator(box, nee_mis):
    scene, cam, orc = box
    bsdf = batch_means(orc, cam, W, H, SPP, K, 3, ob.NO_NEE)
    zb, zi = block_z(nee_mis, bsdf)
    assert zb.max() < 5.0 and zi < 4.0, (zb.max(), zi)
    # the two means agree closely over the image
    assert abs(nee_mis.mean() / bsdf.mean() - 1.0) < 0.02


@pytest.mark.parametrize("plant", [ob.PLANT_NO_DIST2, ob.PLANT_MIS_SWAP])
def test_equivalence_detects_planted_light_mistakes(box, plant):
    scene, cam, orc = box
    bsdf = batch_means(orc, cam, W, H, SPP, K, 3, ob.NO_NEE)
    bad = batch_means(orc, cam, W, H, SPP, K, 3, plant)
    zb, zi = block_z(bad, bsdf)
    assert zb.max() > 8.0 or zi > 8.0, (zb.max(), zi)


def test_russian_roulette_is_unbiased(box):
    scene, cam, orc = box
    rr = batch_means(orc, cam, W, H, SPP, K, 6, 0)
    no_rr = batch_means(orc, cam, W, H, SPP, K, 6, ob.NO_RR)
    zb, zi = block_z(rr, no_rr)
    assert zb.max() < 5.0 and zi < 4.0, (zb.max(), zi)
    bad = batch_means(orc, cam, W, H, SPP, K, 6, ob.PLANT_RR_Q)
    zb2, zi2 = block_z(bad, no_rr)
    assert zb2.max() > 8.0 or zi2 > 8.0, (zb2.max(), zi2)


# --- one bounce under a uniform environment ------------------------------------------------

def _plane_scene(out_dir, base_linear, roughness, metallic):
    """A 400 x 400 quad in the plane z = 0 facing +z; environment colour 1 (linear), no lights."""
    os.makedirs(out_dir, exist_ok=True)
    plane = generate.quad((-200, -200, 0), (400, 0, 0), (0, 400, 0), (0, 0, 1))
    prims = [(generate.write_bin(os.path.join(out_dir, "plane.bin"), plane),
              generate._material("plane", base_linear, roughness, metallic))]
    generate.write_gltf(os.path.join(out_dir, "plane.gltf"), prims)
    scn = os.path.join(out_dir, "plane.scn")
    with open(scn, "w") as f:
        f.write("camera:\n position 0 0 -5\n lookdir 0 0 -1\n fov 1\n exposure 1\n envcol 1 1 1\n\n"
                "objects:\n plane plane.gltf\n\n plane\n")
    return scn


def _brdf_params(mat):
    """The BSDF input the closest-hit shader builds from a constant PbrMaterial record
    (pathtrace.ch.hlsl:283-302; material.h:24-29 layout: base, emission, (rough, metal, normal
    map, bits), specular, sheen, (clearcoat, clearcoat roughness, anisotropy, -))."""
    base, rmnb, spec, sheen, cc = mat[0:4], mat[8:12], mat[12:16], mat[16:20], mat[20:24]
    r2 = max(1e-5, float(rmnb[0]) ** 2)
    aspect = np.sqrt(1.0 - 0.9 * abs(float(cc[2])))
    alpha = (r2 / aspect, r2 * aspect) if cc[2] >= 0 else (r2 * aspect, r2 / aspect)
    cc_alpha = 0.001 + (0.1 - 0.001) * float(cc[1]) ** 2
    return np.array([base[0], base[1], base[2], alpha[0], alpha[1], rmnb[1], spec[0], spec[1], spec[2],
                     sheen[0], sheen[1], sheen[2], cc[0], cc_alpha], np.float32)


def directional_albedo(params, o, n_theta=600, n_phi=1200):
    """integral over the upper hemisphere of f(i, o) cos(theta_i) d omega, midpoint rule in
    (theta, phi)."""
    th = (np.arange(n_theta) + 0.5) * (0.5 * np.pi / n_theta)
    ph = (np.arange(n_phi) + 0.5) * (2 * np.pi / n_phi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    i = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    f = ob.eval_brdf_many(params, i, np.broadcast_to(np.asarray(o, np.float32), i.shape)).astype(np.float64)
    wgt = (np.cos(T) * np.sin(T)).reshape(-1, 1) * (0.5 * np.pi / n_theta) * (2 * np.pi / n_phi)
    return (f * wgt).sum(0)


@pytest.mark.parametrize("base,rough,metal,theta_deg", [
    ((0.8, 0.8, 0.8), 1.0, 0.0, 0.0),
    ((0.8, 0.5, 0.3), 0.6, 0.0, 50.0),
    ((0.9, 0.6, 0.3), 0.5, 1.0, 35.0),
    ((0.5, 0.7, 0.9), 0.35, 0.3, 65.0),
])
def test_one_bounce_environment_furnace(tmp_path, base, rough, metal, theta_deg):
    scene = pt_amd.Scene(_plane_scene(str(tmp_path), base, rough, metal))
    arr = scene.arrays()
    assert arr["light_count"] == 0 and np.allclose(arr["envcol"][:3], 1.0)
    params = _brdf_params(arr["materials"][0])
    # a camera 10 units from the plane's centre at theta_deg from its normal, 0.5 degree field
    th = np.radians(theta_deg)
    eye = np.array([0.0, -10 * np.sin(th), 10 * np.cos(th)])
    fwd = -eye / np.linalg.norm(eye)
    w = h = 16
    cam = pt_amd.camera_block(list(-eye), list(-fwd), fov_deg=0.5, aspect=1.0)
    orc = ob.Oracle(scene)
    img = batch_means(orc, cam, w, h, 64, 8, 1, 0)  # depth 1: the hit, one BSDF sample, the sky
    o = np.array([0.0, np.sin(th), np.cos(th)])  # direction to the eye in the plane's frame
    # the shading frame's x/y axes do not matter for an isotropic BSDF: evaluate at (0, sin, cos)
    expect = directional_albedo(params, o)
    got = img.mean(axis=(0, 1, 2))
    se = img.reshape(len(img), -1, 3).mean(1).std(0, ddof=1) / np.sqrt(len(img))
    tol = np.maximum(5 * se, 2e-3 * expect)
    assert np.all(np.abs(got - expect) <= tol), (got, expect, se)


# --- the light sampler ----------------------------------------------------------------------

def _solid_angle_tri(p, a, b, c):
    """Solid angle of triangle abc seen from p (Van Oosterom & Strackee 1983)."""
    r1, r2, r3 = a - p, b - p, c - p
    l1, l2, l3 = np.linalg.norm(r1), np.linalg.norm(r2), np.linalg.norm(r3)
    num = abs(np.dot(r1, np.cross(r2, r3)))
    den = l1 * l2 * l3 + np.dot(r1, r2) * l3 + np.dot(r1, r3) * l2 + np.dot(r2, r3) * l1
    return 2.0 * np.arctan2(num, den)


def _emitter(scene):
    """The box scene's only emitter: its instance, emitter index and world corners."""
    arr = scene.arrays()
    inst = arr["instances"].reshape(-1, 128)
    head = inst[:, :16].view(np.uint32)
    k = int(np.flatnonzero(head[:, 2] != 0xFFFFFFFF)[0])
    vs, is_, nt, nv = arr["meshes"][arr["instance_mesh"][k]]
    m = inst[k, 16:80].view(np.float32).reshape(4, 4).astype(np.float64)
    pos = arr["vertices"][vs:vs + nv, :3].astype(np.float64)
    world = pos @ m[:3, :3].T + m[:3, 3]
    tris = arr["indices"][is_:is_ + 3 * nt].reshape(nt, 3)
    return int(head[k, 2]), world, tris


def test_light_sampling_pdf_closed_forms(box):
    scene, cam, orc = box
    e_index, world, tris = _emitter(scene)
    assert len(tris) == 2
    area = sum(0.5 * np.linalg.norm(np.cross(world[t[1]] - world[t[0]], world[t[2]] - world[t[0]])) for t in tris)
    # light_pdf at either emitter triangle: the probability of that triangle per unit of area
    for prim, t in enumerate(tris):
        a_t = 0.5 * np.linalg.norm(np.cross(world[t[1]] - world[t[0]], world[t[2]] - world[t[0]]))
        assert orc.light_pdf(e_index, prim) / a_t == pytest.approx(1.0 / area, rel=1e-5)
    n = np.cross(world[tris[0][1]] - world[tris[0][0]], world[tris[0][2]] - world[tris[0][0]])
    n /= np.linalg.norm(n)
    rng = np.random.default_rng(5)
    m = 64
    for p in [np.array([0.2, -1.3, 0.1]), np.array([-0.7, -1.9, 0.6]), np.array([0.05, -0.4, -0.2])]:
        omega = sum(_solid_angle_tri(p, world[t[0]], world[t[1]], world[t[2]]) for t in tris)
        inv, pts = [], []
        for a in range(m):
            for b in range(m):
                u = np.array([(a + rng.random()) / m, (b + rng.random()) / m, rng.random()], np.float32)
                # the first uniform picks light and triangle, the other two the point
                u = np.array([u[2], u[0], u[1]], np.float32)
                e, d, dist, pdf = orc.sample_light(p, u)
                q = p + d.astype(np.float64)
                pts.append(q)
                assert abs(np.dot(q - world[0], n)) < 1e-5  # on the emitter's plane
                inv.append(1.0 / pdf)
        pts = np.array(pts)
        assert np.mean(inv) == pytest.approx(omega, rel=3e-3), (np.mean(inv), omega)
        # uniform over the emitter: the sample centroid is the quad's centre (5 standard errors)
        se = pts.std(0) / np.sqrt(len(pts))
        assert np.all(np.abs(pts.mean(0) - world.mean(0)) <= 5 * se + 1e-7), (pts.mean(0), world.mean(0), se)
