"""Pins the CPU oracle with mathematics that a plausible slip would break.

The oracle restates the reference shaders; nothing executes the reference GPU path here, so
each part is checked against a closed form, an invariant or brute force:
  sampler    Owen-scrambled Sobol prefixes of 2^m samples form (0,m,2)-nets per 2-D pair
             (a wrong direction number, scramble or index shuffle breaks this);
  BSDF       reciprocity, the pdf returned by sampling equals the pdf of evaluation, the
             mixture pdf integrates to <= 1 and ~1 for smooth lobes, energy <= 1;
  traversal  closest and any hits equal fp64 brute force over every triangle;
  camera     all lens samples of one pixel meet at the focus distance (thin-lens invariant);
  estimator  depth 0 renders exactly emission / environment; misses render exactly envcol.
"""
import numpy as np
import pytest

import pt_amd
import oracle_bind as ob


# --- sampler ----------------------------------------------------------------------------

def _is_net(pts, m):
    n = 1 << m
    for a in range(m + 1):
        b = m - a
        ix = np.floor(pts[:, 0] * (1 << a)).astype(np.int64)
        iy = np.floor(pts[:, 1] * (1 << b)).astype(np.int64)
        cells = ix * (1 << b) + iy
        if len(np.unique(cells)) != n:
            return False
    return True


@pytest.mark.parametrize("px,py,seed", [(0, 0, 0), (17, 5, 1), (255, 131, 3)])
def test_sobol_prefixes_are_nets(px, py, seed):
    for m in (4, 6, 8):
        n = 1 << m
        d01 = np.array([ob.sampler_draw(px, py, s, seed, [2])[:2] for s in range(n)])
        d23 = np.array([ob.sampler_draw(px, py, s, seed, [2, 2])[2:4] for s in range(n)])
        assert _is_net(d01, m), (m, "dims 0,1")
        # dims 2,3 are not a (0,2) pair; each is still stratified in 1-D
        for k in (0, 1):
            cells = np.floor(d23[:, k] * n).astype(np.int64)
            assert len(np.unique(cells)) == n, (m, "dim", 2 + k)


def test_sampler_values_in_unit_interval_and_padding():
    v = ob.sampler_draw(3, 4, 77, 2, [1, 1, 2, 1, 2, 2, 1])
    assert np.all(v >= 0) and np.all(v < 1)
    # a 2-D request at dimension 3 pads: the 4th value is the first of a fresh pad
    a = ob.sampler_draw(3, 4, 77, 2, [1, 1, 1, 2])
    b = ob.sampler_draw(3, 4, 77, 2, [1, 1, 1, 1])
    assert a[3] != b[3] or a[4] != b[3]


# --- BSDF -------------------------------------------------------------------------------

def _params(base=(0.7, 0.5, 0.3), alpha=(0.25, 0.25), metal=0.0, spec=(0.5, 0.5, 0.5), sheen=(0, 0, 0)):
    return np.array(list(base) + list(alpha) + [metal] + list(spec) + list(sheen) + [0.0, 0.01], np.float32)


def _hemi(n_theta=256, n_phi=512):
    th = (np.arange(n_theta) + 0.5) / n_theta * (np.pi / 2)
    ph = (np.arange(n_phi) + 0.5) / n_phi * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / n_theta) * (2 * np.pi / n_phi)).reshape(-1)
    return d.astype(np.float32), dw


def _norm(v):
    return v / np.linalg.norm(v)


def test_brdf_reciprocity():
    r = np.random.default_rng(1)
    for _ in range(200):
        i = _norm(r.normal(size=3)); i[2] = abs(i[2]) + 0.05; i = _norm(i)
        o = _norm(r.normal(size=3)); o[2] = abs(o[2]) + 0.05; o = _norm(o)
        m = _norm(i + o)
        p = _params(alpha=tuple(r.uniform(0.02, 1.0, 2)), metal=float(r.integers(0, 2)))
        f1, _ = ob.eval_brdf(p, i, o, m)
        f2, _ = ob.eval_brdf(p, o, i, m)
        np.testing.assert_allclose(f1, f2, rtol=1e-5)


@pytest.mark.parametrize("alpha,metal", [(0.05, 0.0), (0.3, 0.0), (0.8, 1.0), (0.5, 0.5)])
def test_sampled_pdf_equals_evaluated_pdf(alpha, metal):
    r = np.random.default_rng(2)
    o = _norm(np.array([0.3, -0.2, 0.9]))
    p = _params(alpha=(alpha, alpha), metal=metal)
    for _ in range(300):
        u = r.random(3).astype(np.float32)
        i, m, w, pdf = ob.sample_direction(o, (alpha, alpha), metal, u)
        if i[2] <= 0:
            continue
        _, pdf_eval = ob.eval_brdf(p, i, o, m)
        # one-sample MIS weight: w * (mixture pdf) == pdf of the chosen lobe
        np.testing.assert_allclose(w * pdf_eval, pdf, rtol=2e-3)


@pytest.mark.parametrize("alpha,metal", [(0.1, 0.0), (0.4, 0.0), (0.4, 1.0), (1.0, 0.3)])
def test_pdf_integrates_to_at_most_one(alpha, metal):
    o = _norm(np.array([0.2, 0.1, 0.95]))
    d, dw = _hemi(128, 256)
    p = _params(alpha=(alpha, alpha), metal=metal)
    tot_pdf = tot_e = 0.0
    for k in range(len(d)):
        i = d[k]
        m = _norm(i + o)
        f, pdf = ob.eval_brdf(p, i, o, m)
        tot_pdf += pdf * dw[k]
        tot_e += f.max() * i[2] * dw[k]
    assert tot_pdf <= 1.0 + 1e-2
    # the diffuse share integrates exactly; VNDF reflections of rough lobes partly fall below
    # the horizon, smooth lobes keep nearly all their mass
    assert tot_pdf >= 0.5 * (1.0 - metal) - 1e-2
    if alpha <= 0.1:
        assert tot_pdf >= 0.98
    assert tot_e <= 1.0 + 2e-2  # energy conservation of the lobes
    # The evaluated pdf over the upper hemisphere is the probability that sampling lands there.
    r = np.random.default_rng(11)
    n = 4000
    up = sum(ob.sample_direction(o, (alpha, alpha), metal, r.random(3))[0][2] > 0 for _ in range(n)) / n
    assert abs(up - tot_pdf) <= 4 * np.sqrt(tot_pdf * (1 - tot_pdf) / n) + 5e-3, (up, tot_pdf)


# --- traversal vs brute force ---------------------------------------------------------------

def _brute(scene, rays, any_hit):
    a = scene.arrays()
    verts = a["vertices"][:, :3].astype(np.float64)
    inst = a["instances"].view(np.float32).reshape(-1, 32)
    out = []
    for r in rays:
        o = r["origin"].astype(np.float64); d = r["dir"].astype(np.float64)
        best = (r["tmax"], None)
        for k, mi in enumerate(a["instance_mesh"]):
            vs, ist, nt, _ = a["meshes"][mi]
            M = inst[k, 4:20].astype(np.float64).reshape(4, 4)
            tri = a["indices"][ist:ist + 3 * nt].reshape(nt, 3) + vs
            P = verts[tri] @ M[:3, :3].T + M[:3, 3]   # world-space corners
            e1, e2 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
            pv = np.cross(d, e2)
            det = np.einsum("ij,ij->i", e1, pv)
            ok = np.abs(det) > 1e-14
            inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
            tv = o - P[:, 0]
            u = np.einsum("ij,ij->i", tv, pv) * inv
            q = np.cross(tv, e1)
            v = (q @ d) * inv
            t = np.einsum("ij,ij->i", e2, q) * inv
            hit = ok & (u >= -1e-9) & (v >= -1e-9) & (u + v <= 1 + 1e-9) & (t >= r["tmin"]) & (t <= best[0])
            if hit.any():
                j = np.argmin(np.where(hit, t, np.inf))
                best = (t[j], (k, j))
        out.append(best)
    return out


def test_oracle_traversal_equals_brute_force(box_scn):
    scene = pt_amd.Scene(box_scn)
    orc = ob.Oracle(scene)
    r = np.random.default_rng(4)
    n = 300
    rays = np.zeros(n, pt_amd.RAY_DTYPE)
    rays["origin"] = r.uniform([-1.5, -2.5, -1.5], [1.5, 0.5, 3.0], (n, 3))
    dirs = r.normal(size=(n, 3))
    rays["dir"] = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    rays["tmin"] = 0.0
    rays["tmax"] = 1e30
    got = orc.trace_rays(rays)
    ref = _brute(scene, rays, False)
    for k in range(n):
        t, hit = ref[k]
        if hit is None:
            assert got[k]["instance"] == pt_amd.PT_NO_HIT
        else:
            assert got[k]["instance"] != pt_amd.PT_NO_HIT
            np.testing.assert_allclose(got[k]["t"], t, rtol=1e-5)
            if (got[k]["instance"], got[k]["primitive"]) != hit:
                # a tie: another triangle at the same depth (shared edge)
                assert abs(got[k]["t"] - t) <= 1e-6 * max(1.0, t)
    any_hits = orc.trace_rays(rays, any_hit=True)
    assert np.array_equal(any_hits["instance"] != pt_amd.PT_NO_HIT, got["instance"] != pt_amd.PT_NO_HIT)


# --- camera ---------------------------------------------------------------------------------

def test_thin_lens_rays_meet_at_focus_distance():
    cam = pt_amd.camera_block((0, 1, -3), (0, 0, 1), fov_deg=60, lens_radius=0.2, focus_dist=3.0)
    pts = []
    for ul in [(0.1, 0.2), (0.9, 0.5), (0.4, 0.95), (0.6, 0.05)]:
        o, d = ob.camera_ray(cam, 100, 80, 256, 256, (0.5, 0.5), ul)
        # the eye is (0,-1,3) looking along -z: the focus plane at distance 3 is z = 0
        t = o[2] / -d[2]
        pts.append(o + d * t)
    pts = np.array(pts)
    np.testing.assert_allclose(pts, np.repeat(pts[:1], len(pts), 0), atol=2e-5)


# --- estimator closed forms -----------------------------------------------------------------

def test_depth_zero_renders_emission_and_environment(box_scn):
    scene = pt_amd.Scene(box_scn)
    orc = ob.Oracle(scene)
    w = h = 64
    cam = scene.camera(w, h)
    img, cnt = orc.render(cam, w, h, 1, 0)
    hits = orc.primary(cam, w, h, 0)
    env = np.array(scene.arrays()["envcol"][:3], np.float32)
    miss = hits["instance"] == pt_amd.PT_NO_HIT
    np.testing.assert_array_equal(img[miss][:, :3], np.broadcast_to(env, (miss.sum(), 3)))
    mats = scene.arrays()["materials"]
    inst = scene.arrays()["instances"].view(np.uint32).reshape(-1, 32)
    for (y, x) in np.argwhere(~miss):
        mat = mats[inst[hits[y, x]["instance"], 3]]
        e = mat[4:7] * mat[7]
        px = img[y, x, :3]
        assert np.allclose(px, e) or np.allclose(px, 0.0), (px, e)  # one-sided emitters
    assert cnt["shadow_rays"] == 0 and cnt["extension_rays"] == w * h


def test_running_mean_over_split_renders(box_scn):
    # Rendering samples [0,4) then [4,8) equals rendering [0,8) in one call (same recurrence).
    scene = pt_amd.Scene(box_scn)
    orc = ob.Oracle(scene)
    w = h = 32
    cam = scene.camera(w, h)
    a, _ = orc.render(cam, w, h, 8, 3)
    b, _ = orc.render(cam, w, h, 4, 3)
    b, _ = orc.render(cam, w, h, 4, 3, sample_index=4, accum=b)
    np.testing.assert_array_equal(a, b)
