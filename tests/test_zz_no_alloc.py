"""pt_trace_rays_noalloc (through the C ABI) against pt_trace_rays: the allocation-free variant
stages rays and hits through the library's fixed device array and must return the same hits,
call after call, whatever the previous call left in that array."""
import numpy as np
import pytest

import pt_amd

pytestmark = pytest.mark.gpu


def _rays(n, seed, tmax=np.float32(np.inf)):
    r = np.random.default_rng(seed)
    rays = np.zeros(n, pt_amd.RAY_DTYPE)
    rays["origin"] = r.uniform(-60, 60, (n, 3))
    d = r.normal(size=(n, 3))
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmin"] = 0.0
    rays["tmax"] = tmax
    return rays


@pytest.fixture(scope="module")
def tracers(scene_dir, box_scn):
    from pt_amd.scenes import generate
    inst = pt_amd.Scene(generate.cached_scene("instanced", root=scene_dir))
    box = pt_amd.Scene(box_scn)
    a, b = pt_amd.Tracer(inst, 32, 32), pt_amd.Tracer(box, 32, 32)
    yield a, b
    a.close()
    b.close()


def test_noalloc_rays_equal_trace_rays_call_after_call(tracers):
    a, b = tracers
    big = _rays(200_000, 1)
    small = _rays(3_000, 2, tmax=np.float32(40.0))
    for any_hit in (False, True):
        ref_big = a.trace_rays(big, any_hit=any_hit)
        got_big = a.trace_rays_noalloc(big, any_hit=any_hit)
        np.testing.assert_array_equal(got_big, ref_big)
        assert (got_big["instance"] != pt_amd.PT_NO_HIT).mean() > 0.01
        # a second, smaller call: the array still holds the first call's rays and hits beyond it
        np.testing.assert_array_equal(a.trace_rays_noalloc(small, any_hit=any_hit),
                                      a.trace_rays(small, any_hit=any_hit))
    # another context on the same GPU shares the array
    box_rays = _rays(5_000, 3)
    box_rays["origin"] = np.random.default_rng(4).uniform(-0.9, 0.9, (5_000, 3)) + np.array([0.0, -1.0, 0.0])
    np.testing.assert_array_equal(b.trace_rays_noalloc(box_rays), b.trace_rays(box_rays))
    np.testing.assert_array_equal(a.trace_rays_noalloc(small), a.trace_rays(small))


def test_noalloc_rejects_more_rays_than_its_array_holds(tracers):
    a, _ = tracers
    limit = 262144
    with pytest.raises(pt_amd.PtError, match="at most 262144 rays") as e:
        a.trace_rays_noalloc(_rays(limit + 1, 5))
    assert e.value.code == -1
    out = a.trace_rays_noalloc(_rays(limit, 6))  # the maximum itself is accepted
    assert len(out) == limit
