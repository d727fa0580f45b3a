"""The batched workloads through the C ABI: 64 views of the instanced scene in one batch, and
16 scenes of 10k..20M triangles with 4 views each in one batch (pt_amd/scenes/workloads.py).

Each batch renders every view with all views' paths streaming through one path pool; a view
must equal the same view rendered alone (same seeds per pixel and sample index, same running
mean), and the scenes' device-built hierarchies pass the host validation pt_create runs
(every node index in range, every triangle and instance referenced once) and give the
oracle's primary hits, every mismatch an fp64-confirmed tie.
"""
import numpy as np
import pytest

import pt_amd
import oracle_bind as ob
from pt_amd.scenes import generate
from pt_amd.scenes import workloads as wl
from tie_check import Adjudicator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def inst(scene_dir):
    scene = pt_amd.Scene(generate.cached_scene("instanced", root=scene_dir))
    return scene, pt_amd.Tracer(scene, 64, 64)


@pytest.mark.timeout(600)
def test_batch_of_64_views_equals_single_renders(inst):
    scene, tr = inst
    views = wl.views_batch()
    assert len(views) == 64
    batch = pt_amd.Batch([tr], [(0,) + v for v in views])
    batch.render()
    st = batch.stats()
    assert st["paths"] == sum(w * h * spp for w, h, spp, _, _ in views)
    assert st["extension_rays"] >= st["paths"]
    imgs = []
    for i, (w, h, spp, depth, cam) in enumerate(views):
        img = batch.read(i)
        assert img.shape == (h, w, 4) and np.isfinite(img).all(), i
        assert np.all(img[..., 3] == 1.0), i  # every pixel's first sample was accumulated
        imgs.append(img)
    batch.close()
    # four views rendered alone: the smallest, the largest, the most and the fewest samples
    pick = {int(np.argmin([w * h for w, h, *_ in views])), int(np.argmax([w * h for w, h, *_ in views])),
            int(np.argmax([v[2] for v in views])), int(np.argmin([v[2] for v in views]))}
    for i in sorted(pick):
        w, h, spp, depth, cam = views[i]
        one = pt_amd.Tracer(scene, w, h, max_bounces=depth)
        one.set_camera(cam)
        one.render(spp)
        np.testing.assert_array_equal(imgs[i], one.read(), err_msg="view %d" % i)
        one.close()


def _view_of_largest_instance(scene):
    """An eye and view direction from which the scene's largest instance fills the view."""
    a = scene.arrays()
    inst = a["instances"].reshape(-1, 128)
    m = inst[:, 16:80].copy().view(np.float32).reshape(-1, 4, 4).astype(np.float64)
    tri_counts = [a["meshes"][int(a["instance_mesh"][i])][2] for i in range(len(inst))]
    i = int(np.argmax(tri_counts))
    vs, _, _, nv = a["meshes"][int(a["instance_mesh"][i])]
    p = a["vertices"][vs:vs + nv, :3].astype(np.float64) @ m[i, :3, :3].T + m[i, :3, 3]
    centre, extent = p.mean(0), np.ptp(p, axis=0).max()
    fwd = np.array([0.3, -0.4, 1.0])
    fwd /= np.linalg.norm(fwd)
    return centre - 1.2 * extent * fwd, fwd


@pytest.mark.timeout(900)
def test_batch_of_16_scenes_up_to_20m_triangles(scene_dir):
    sizes = wl.scene_sizes()
    tracers, views, tris = [], [], []
    for k, target in enumerate(sizes):
        scn = generate.cached_scene("instanced", root=scene_dir, seed=100 + k,
                                    counts=generate.counts_for_triangles(target))
        scene = pt_amd.Scene(scn)
        info = scene.info()
        tris.append(info["total_triangles"])
        t = pt_amd.Tracer(scene, 64, 64)  # builds and validates the BVH
        s = t.stats()
        assert s["tri_record_count"] == info["total_triangles"] and s["node_count"] > 0
        tracers.append((scene, t))
        # the batch's views at 1-2 spp (the timed workload runs 16-128)
        views += [(k,) + v for v in wl.scene_views(100 + k, spp_lo=1, spp_hi=2)]
    assert tris[0] >= 10000 and tris[-1] >= 20_000_000
    batch = pt_amd.Batch([t for _, t in tracers], views)
    batch.render()
    st = batch.stats()
    assert st["paths"] == sum(w * h * spp for _, w, h, spp, _, _ in views)
    for i in range(len(views)):
        assert np.isfinite(batch.read(i)).all(), i
    batch.close()
    # primary hits against the oracle on a small view of two scenes, looking at their largest
    # instance
    for k in (0, 8):
        scene, t = tracers[k]
        w, h = 96, 64
        cam = wl.camera(*_view_of_largest_instance(scene), w, h, lens_radius=0.0)
        t2 = pt_amd.Tracer(scene, w, h)
        t2.set_camera(cam)
        g = t2.trace_primary(0)
        orc = ob.Oracle(scene)
        o, rays = orc.primary(cam, w, h, 0, want_rays=True)
        assert (o["instance"] != pt_amd.PT_NO_HIT).mean() > 0.2  # the view sees geometry
        adj = Adjudicator(scene)
        same = (g["instance"] == o["instance"]) & (g["primitive"] == o["primitive"])
        for y, x in np.argwhere(~same):
            assert adj.ok(rays[y, x], g[y, x], o[y, x]), (k, y, x, g[y, x], o[y, x])
        assert (~same).mean() <= 1e-3
        t2.close()
    for _, t in tracers:
        t.close()


def _camera_path_run(scene, path, w, h, query, learning_rate):
    tr = pt_amd.Tracer(scene, w, h, max_bounces=8)
    tr.cache_configure(query=query, train_stride=4, train_steps=4, batch=16384, learning_rate=learning_rate,
                       frame_spp=4, seed=1)
    losses, records = [], []
    for spp, cam in path:
        tr.set_camera(cam)
        tr.render(spp)
        assert np.isfinite(tr.read()).all()
        cs = tr.cache_stats()
        losses.append(cs["last_loss"])
        records.append(cs["last_records"])
    steps = tr.cache_stats()["steps"]
    tr.close()
    return np.array(losses), np.array(records), steps


@pytest.mark.timeout(600)
def test_learned_cache_on_a_camera_path(inst):
    """The camera-path workload at reduced resolution: 64 frames at 1-4 spp through 8 keyframes,
    the cache trained online after every frame (Adam, learning rate 1e-3, 4 x 16384 record
    samples per frame) and queried at the second vertex: every frame's image is finite and
    records arrive every frame. That the network learns is checked on paired data: trained
    only (no queries, so the paths and records do not depend on the network), the per-frame
    loss of the last step falls well below that of the same run at a learning rate of 1e-7
    (the network as initialised) on the same records."""
    scene, _ = inst
    w, h = 320, 180
    path = wl.camera_path(w=w, h=h)
    assert len(path) == 64 and all(1 <= spp <= 4 for spp, _ in path)
    losses, records, steps = _camera_path_run(scene, path, w, h, True, 1e-3)
    assert records.min() > 0 and steps == 4 * 64
    print("cache losses (queried)", np.round(losses, 3).tolist())
    trained, rec_a, _ = _camera_path_run(scene, path, w, h, False, 1e-3)
    frozen, rec_b, _ = _camera_path_run(scene, path, w, h, False, 1e-7)
    np.testing.assert_array_equal(rec_a, rec_b)  # the same records in both runs
    ratio = trained / frozen
    print("trained / initial loss per frame", np.round(ratio, 3).tolist())
    assert np.median(ratio[-32:]) < 0.5, ratio
