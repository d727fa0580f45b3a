"""Batches of views and scenes in one wavefront run, and graph replay of the iterations.

Every view of a batch traces the same paths as a context rendering that view alone (same
sampler seeding per pixel and sample index), so the images must be bit-identical."""
import numpy as np
import pytest

import pt_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(box_scn, scene_dir):
    from pt_amd.scenes import generate
    box = pt_amd.Scene(box_scn)
    inst = pt_amd.Scene(generate.cached_scene("instanced", root=scene_dir))
    tb = pt_amd.Tracer(box, 64, 64)
    ti = pt_amd.Tracer(inst, 64, 64)
    return {"box": (box, tb), "inst": (inst, ti)}


def _cam(pos, look, w, h, fov=60.0, lens=0.0, focus=0.0, exposure=1.0):
    return pt_amd.camera_block(pos, look, fov_deg=fov, aspect=w / h, focus_dist=focus, lens_radius=lens,
                               exposure=exposure)


def _alone(scene, w, h, spp, bounces, cam, calls=1):
    t = pt_amd.Tracer(scene, w, h, max_bounces=bounces)
    t.set_camera(cam)
    for _ in range(calls):
        t.render(spp)
    img = t.read()
    t.close()
    return img


BOX_VIEWS = [
    (96, 64, 3, 4, ((0, 1, -3), (0, 0, 1))),
    (40, 72, 5, 8, ((0.3, 1.2, -2.5), (-0.1, 0.1, 1))),
    (128, 128, 2, 2, ((0, 1, -3.5), (0, 0.05, 1))),
]


def test_batch_views_equal_single_renders(scenes):
    box, tb = scenes["box"]
    views = [(0, w, h, spp, b, _cam(p, l, w, h, lens=0.03, focus=3.0)) for w, h, spp, b, (p, l) in BOX_VIEWS]
    batch = pt_amd.Batch([tb], views)
    batch.render()
    batch.render()  # accumulation continues per view
    for i, (_, w, h, spp, b, cam) in enumerate(views):
        np.testing.assert_array_equal(batch.read(i), _alone(box, w, h, spp, b, cam, calls=2), err_msg="view %d" % i)
    st = batch.stats()
    assert st["paths"] == 2 * sum(w * h * spp for _, w, h, spp, _, _ in views)
    assert st["extension_rays"] >= st["paths"]
    batch.close()


def test_batch_of_two_scenes_equals_single_renders(scenes):
    box, tb = scenes["box"]
    inst, ti = scenes["inst"]
    views = [
        (0, 80, 60, 4, 4, _cam((0, 1, -3), (0, 0, 1), 80, 60)),
        (1, 120, 90, 3, 8, _cam((0, 0, 0), (0, 0, -1), 120, 90, lens=0.03, focus=30.0)),
        (1, 50, 50, 6, 8, _cam((10, 5, 20), (-0.3, -0.1, -1), 50, 50)),
        (0, 33, 47, 2, 8, _cam((0.5, 1, -2), (-0.2, 0, 1), 33, 47)),
    ]
    batch = pt_amd.Batch([tb, ti], views)
    batch.render()
    for i, (sc, w, h, spp, b, cam) in enumerate(views):
        scene = box if sc == 0 else inst
        np.testing.assert_array_equal(batch.read(i), _alone(scene, w, h, spp, b, cam), err_msg="view %d" % i)
    batch.close()


def test_batch_set_view_restarts_accumulation(scenes):
    box, tb = scenes["box"]
    c1 = _cam((0, 1, -3), (0, 0, 1), 32, 32)
    c2 = _cam((0.2, 1, -3), (0, 0, 1), 32, 32)
    batch = pt_amd.Batch([tb], [(0, 32, 32, 2, 3, c1)])
    batch.render()
    batch.set_view(0, (0, 32, 32, 3, 5, c2))
    batch.render()
    np.testing.assert_array_equal(batch.read(0), _alone(box, 32, 32, 3, 5, c2))
    with pytest.raises(pt_amd.PtError):
        batch.set_view(0, (0, 16, 32, 3, 5, c2))  # the image size is fixed
    batch.close()


def test_batch_argument_checks(scenes):
    _, tb = scenes["box"]
    c = _cam((0, 1, -3), (0, 0, 1), 8, 8)
    with pytest.raises(pt_amd.PtError, match="scene index"):
        pt_amd.Batch([tb], [(1, 8, 8, 1, 2, c)])
    with pytest.raises(pt_amd.PtError, match="Sample count was zero"):
        pt_amd.Batch([tb], [(0, 8, 8, 0, 2, c)])
    with pytest.raises(pt_amd.PtError, match="size"):
        pt_amd.Batch([tb], [(0, 0, 8, 1, 2, c)])


@pytest.mark.parametrize("kind", ["box", "inst"])
def test_graph_replay_equals_direct_launches(scenes, kind):
    scene, _ = scenes[kind]
    w, h = 72, 40
    cam = scene.camera(w, h)
    imgs = []
    for graphs in (0, 1):
        t = pt_amd.Tracer(scene, w, h, max_bounces=8)
        t.set_option(pt_amd.PT_OPT_GRAPHS, graphs)
        t.set_camera(cam)
        for spp in (1, 3, 2):
            t.render(spp)
        imgs.append(t.read())
        t.close()
    np.testing.assert_array_equal(imgs[0], imgs[1])


@pytest.mark.parametrize("kind", ["box", "inst"])
def test_pipelined_shadow_rays_across_mode_changes(scenes, kind):
    """With graphs, an iteration's shadow rays are traced in the next iteration's launch, with
    its extension rays; iterations that time their stages or count fetches trace them in their
    own iteration and first trace any the previous iteration left. Renders that switch between
    the two (a small pool, so paths stream across calls) equal renders in either mode alone."""
    scene, _ = scenes[kind]
    w, h = 96, 64
    cam = scene.camera(w, h)

    def run(plan):
        t = pt_amd.Tracer(scene, w, h, max_bounces=8)
        t.set_option(pt_amd.PT_OPT_POOL_SLOTS, 4096)
        t.set_camera(cam)
        for option, value, spp in plan:
            if option is not None:
                t.set_option(option, value)
            t.render(spp)
        img = t.read()
        t.close()
        return img

    pipelined = run([(None, 0, 3), (None, 0, 5), (None, 0, 4)])
    per_stage = run([(pt_amd.PT_OPT_PROFILE_KERNELS, 1, 3), (None, 0, 5), (None, 0, 4)])
    mixed = run([(None, 0, 3), (pt_amd.PT_OPT_PROFILE_KERNELS, 1, 5), (pt_amd.PT_OPT_PROFILE_KERNELS, 0, 1),
                 (pt_amd.PT_OPT_COUNT_FETCHES, 1, 2), (pt_amd.PT_OPT_COUNT_FETCHES, 0, 1)])
    assert np.isfinite(pipelined).all() and pipelined[..., :3].mean() > 0
    np.testing.assert_array_equal(pipelined, per_stage)
    np.testing.assert_array_equal(pipelined, mixed)
