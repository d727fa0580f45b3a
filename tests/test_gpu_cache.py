"""The learned radiance cache on the device: MLP kernels against the NumPy model, Adam, and
the cache inside the renderer (unchanged images when only training, gain when querying)."""
import numpy as np
import pytest

import mlp_ref
import pt_amd

pytestmark = pytest.mark.gpu

W, H = 96, 96


@pytest.fixture(scope="module")
def box(box_scn):
    scene = pt_amd.Scene(box_scn)
    cam = scene.camera(W, H)
    tr = pt_amd.Tracer(scene, W, H, max_bounces=8)
    tr.set_camera(cam)
    tr.cache_configure(query=False, seed=7)
    return scene, cam, tr


def _features(n, seed):
    r = np.random.default_rng(seed)
    return r.uniform(-1, 1, (n, 96)).astype(np.float32)


def _params(seed, scale=1.0):
    r = np.random.default_rng(seed)
    p = np.zeros(mlp_ref.layout()[1], np.float32)
    for w, b, ni, no in mlp_ref.layout()[0]:
        p[w:w + ni * no] = r.uniform(-1, 1, ni * no) * np.sqrt(6.0 / ni) * scale
        p[b:b + no] = r.uniform(-0.2, 0.2, no)
    return p


def test_param_count_and_initialisation(box):
    _, _, tr = box
    assert pt_amd.hip_lib().pt_cache_param_count() == mlp_ref.layout()[1]
    tr.cache_configure(query=False, seed=7)
    np.testing.assert_array_equal(tr.cache_params(), mlp_ref.he_uniform(7))
    st = tr.cache_stats()
    assert st["enabled"] == 1 and st["steps"] == 0 and st["param_count"] == mlp_ref.layout()[1]
    # world box of the scene: the box [-1,1] x [-2,0] x [-1,1] and the emitter inside it
    np.testing.assert_allclose(st["bounds_lo"], [-1, -2, -1], atol=1e-3)
    np.testing.assert_allclose(st["bounds_hi"], [1, 0, 1], atol=1e-3)


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000])
def test_forward_matches_numpy(box, n):
    _, _, tr = box
    p = _params(11)
    tr.cache_set_params(p)
    X = _features(n, n)
    y = tr.cache_forward(X)
    yb, _ = mlp_ref.forward(p, X, "bf16")
    ye, _ = mlp_ref.forward(p, X, "exact")
    # same rounding points as the kernel: only accumulation order and rare BF16 ties differ
    assert np.linalg.norm(y - yb) <= 5e-3 * np.linalg.norm(yb) + 1e-6, np.abs(y - yb).max()
    assert np.linalg.norm(y - ye) <= 5e-2 * np.linalg.norm(ye)


@pytest.mark.parametrize("n", [32, 77, 512])
def test_gradients_match_numpy(box, n):
    _, _, tr = box
    p = _params(12)
    tr.cache_set_params(p)
    X = _features(n, 100 + n)
    T = np.random.default_rng(n).uniform(0, 3, (n, 3)).astype(np.float32)
    loss, g = tr.cache_train_batch(X, T, apply=False)
    Lb, gb = mlp_ref.loss_and_grad(p, X, T, "bf16")
    Le, ge = mlp_ref.loss_and_grad(p, X, T, "exact")
    assert loss == pytest.approx(Lb, rel=5e-3)
    assert loss == pytest.approx(Le, rel=5e-2)
    for l, (w, b, ni, no) in enumerate(mlp_ref.layout()[0]):
        for lo, hi in ((w, w + ni * no), (b, b + no)):
            d = np.linalg.norm(g[lo:hi] - gb[lo:hi])
            assert d <= 2e-2 * np.linalg.norm(gb[lo:hi]) + 1e-7, (l, lo, d, np.linalg.norm(gb[lo:hi]))
            # no further from the fp64 gradient than BF16 arithmetic itself puts the model
            # (deep layers amplify the rounding of the back-propagated gradients)
            gap = np.linalg.norm(gb[lo:hi] - ge[lo:hi])
            assert np.linalg.norm(g[lo:hi] - ge[lo:hi]) <= 1.5 * gap + 2e-2 * np.linalg.norm(ge[lo:hi]) + 1e-7
    # parameters untouched without apply
    np.testing.assert_array_equal(tr.cache_params(), p)


def test_large_batch_gradients_are_deterministic(box):
    """A batch of more 32-sample tiles than the training launch has blocks (512): some blocks
    add a second tile into their row of partial gradients. The summed gradient matches the
    NumPy model and is the same, bit for bit, on a second call (the rows are added in a fixed
    order, no atomics)."""
    _, _, tr = box
    p = _params(21)
    tr.cache_set_params(p)
    n = 17000
    X = _features(n, 77)
    T = np.random.default_rng(77).uniform(0, 3, (n, 3)).astype(np.float32)
    loss1, g1 = tr.cache_train_batch(X, T, apply=False)
    loss2, g2 = tr.cache_train_batch(X, T, apply=False)
    np.testing.assert_array_equal(g1, g2)
    Lb, gb = mlp_ref.loss_and_grad(p, X, T, "bf16")
    assert loss1 == pytest.approx(Lb, rel=5e-3)
    for l, (w, b, ni, no) in enumerate(mlp_ref.layout()[0]):
        for lo, hi in ((w, w + ni * no), (b, b + no)):
            d = np.linalg.norm(g1[lo:hi] - gb[lo:hi])
            assert d <= 2e-2 * np.linalg.norm(gb[lo:hi]) + 1e-7, (l, lo, d, np.linalg.norm(gb[lo:hi]))


def test_adam_step_matches_numpy(box):
    _, _, tr = box
    tr.cache_configure(query=False, seed=3)  # fresh Adam state
    p = _params(13)
    tr.cache_set_params(p)
    X = _features(256, 5)
    T = np.random.default_rng(5).uniform(0, 1, (256, 3)).astype(np.float32)
    m = np.zeros_like(p, np.float64)
    v = np.zeros_like(p, np.float64)
    cur = p.astype(np.float64)
    for step in (1, 2, 3):
        _, g = tr.cache_train_batch(X, T, apply=True)
        cur, m, v = mlp_ref.adam(cur, g.astype(np.float64), m, v, step, lr=1e-3)
        got = tr.cache_params()
        np.testing.assert_allclose(got, cur, rtol=1e-5, atol=1e-6)
        cur = got.astype(np.float64)
    assert tr.cache_stats()["steps"] == 3


def test_query_weights_are_the_moving_average(box):
    """ema_decay d: after t Adam steps the queries read sum_k (1 - d) d^(t-k) w_k / (1 - d^t), the
    normalised weighted mean of the trained weights w_1..w_t (closed form, float64), restarted by
    pt_cache_set_params; with d = 0 they read the trained weights."""
    _, _, tr = box
    d = 0.75
    tr.cache_configure(query=False, seed=3, ema_decay=d)
    np.testing.assert_array_equal(tr.cache_query_params(), tr.cache_params())
    X = _features(256, 6)
    T = np.random.default_rng(6).uniform(0, 1, (256, 3)).astype(np.float32)
    for restart in range(2):
        ws = []
        for t in range(1, 6):
            tr.cache_train_batch(X, T, apply=True)
            ws.append(tr.cache_params().astype(np.float64))
            expect = sum((1 - d) * d ** (t - k) * ws[k - 1] for k in range(1, t + 1)) / (1 - d ** t)
            np.testing.assert_allclose(tr.cache_query_params(), expect, rtol=1e-5, atol=1e-6)
        assert np.abs(tr.cache_query_params() - ws[-1]).max() > 1e-4  # the mean lags the last step
        tr.cache_set_params(_params(14))
        np.testing.assert_array_equal(tr.cache_query_params(), _params(14))
    tr.cache_configure(query=False, seed=3)
    tr.cache_train_batch(X, T, apply=True)
    np.testing.assert_array_equal(tr.cache_query_params(), tr.cache_params())
    for bad in (1.0, -0.5):
        with pytest.raises(pt_amd.PtError, match="cache configuration"):
            tr.cache_configure(ema_decay=bad)


def test_encoding_matches_numpy(box):
    _, _, tr = box
    r = np.random.default_rng(8)
    rec = np.zeros((500, 13), np.float32)
    rec[:, :3] = r.uniform([-1.2, -2.2, -1.2], [1.2, 0.2, 1.2], (500, 3))  # some outside the box
    rec[:, 3:13] = r.uniform(-1, 1, (500, 10))
    st = tr.cache_stats()
    f = tr.cache_encode(rec)
    ref = mlp_ref.encode(rec, st["bounds_lo"], st["bounds_hi"])
    # BF16 rounding (2^-8 relative) plus fp32 argument reduction at the highest octave
    np.testing.assert_allclose(f, ref, rtol=2 ** -7, atol=2e-3)
    np.testing.assert_array_equal(f, mlp_ref.bf16(f))


def test_training_reduces_loss(box):
    _, _, tr = box
    tr.cache_configure(query=False, seed=4, learning_rate=3e-3)
    r = np.random.default_rng(9)
    X = r.uniform(-1, 1, (4096, 96)).astype(np.float32)
    T = (1.0 + 0.5 * np.sin(3 * X[:, :3]) + 0.3 * X[:, 3:6] ** 2).astype(np.float32)
    first = None
    for it in range(300):
        sl = slice((it % 8) * 512, (it % 8 + 1) * 512)
        loss, _ = tr.cache_train_batch(X[sl], T[sl], apply=True)
        first = loss if first is None else first
    final, _ = tr.cache_train_batch(X[:512], T[:512], apply=False)
    assert final < 0.1 * first, (first, final)


def test_training_only_leaves_the_image_unchanged(box):
    scene, cam, tr = box
    plain = pt_amd.Tracer(scene, W, H, max_bounces=8)
    plain.set_camera(cam)
    plain.render(8)
    tr.cache_configure(query=False, seed=1, train_stride=4, train_steps=2, batch=4096)
    tr.set_camera(cam)
    for _ in range(8):
        tr.render(1)
    np.testing.assert_array_equal(tr.read(), plain.read())
    st = tr.cache_stats()
    assert st["steps"] == 16
    assert st["records_total"] > 0.1 * 8 * W * H / 4  # one path in 4 trains when it reaches vertex 2
    assert np.isfinite(st["last_loss"]) and st["last_loss"] > 0
    plain.close()


def test_inflight_training_leaves_the_image_unchanged(box):
    scene, cam, tr = box
    plain = pt_amd.Tracer(scene, W, H, max_bounces=8)
    plain.set_camera(cam)
    plain.render(16)
    tr.cache_configure(query=False, seed=5, train_stride=4, train_steps=1, batch=2048, frame_spp=0)
    tr.set_camera(cam)
    tr.render(16)
    np.testing.assert_array_equal(tr.read(), plain.read())
    st = tr.cache_stats()
    assert st["steps"] >= 9  # one step after every wavefront iteration of the 16-spp job
    assert st["last_records"] > 0.1 * 16 * W * H / 4
    assert np.isfinite(st["last_loss"]) and st["last_loss"] > 0
    plain.close()


@pytest.mark.parametrize("vertex", [3, 8])
def test_training_at_a_later_vertex_leaves_the_image_unchanged(box, vertex):
    """Records taken at a later path vertex: the image still equals the plain tracer's, and
    fewer paths reach vertex 8 than vertex 3 (and some reach neither)."""
    scene, cam, tr = box
    plain = pt_amd.Tracer(scene, W, H, max_bounces=8)
    plain.set_camera(cam)
    plain.render(4)
    counts = {}
    for v in (2, vertex):
        tr.cache_configure(query=False, seed=1, train_stride=4, train_steps=1, batch=4096, query_vertex=v)
        tr.set_camera(cam)
        for _ in range(4):
            tr.render(1)
        np.testing.assert_array_equal(tr.read(), plain.read())
        counts[v] = tr.cache_stats()["records_total"]
    assert 0 < counts[vertex] < counts[2], counts
    for bad in (1, 9):
        with pytest.raises(pt_amd.PtError, match="cache configuration"):
            tr.cache_configure(query_vertex=bad)
    plain.close()


def test_control_variate_with_certain_continuation_is_the_plain_estimate(box):
    """continue_prob = 1: every queried path goes on with weight 1 and the prediction enters
    with weight 0, so the image is the plain one."""
    scene, cam, tr = box
    plain = pt_amd.Tracer(scene, W, H, max_bounces=8)
    plain.set_camera(cam)
    plain.render(8)
    for vertex in (2, 4):
        tr.cache_configure(query=True, seed=6, train_stride=8, train_steps=2, batch=4096, continue_prob=1.0,
                           query_vertex=vertex)
        tr.set_camera(cam)
        tr.render(2)  # trains; the first frame is not queried
        tr.set_camera(cam)
        for _ in range(8):
            tr.render(1)
        np.testing.assert_array_equal(tr.read(), plain.read(), err_msg="vertex %d" % vertex)
    plain.close()


def test_query_limit_below_every_pdf_queries_nothing(box):
    """query_max_pdf: only paths that arrived by a direction of pdf below the limit are ended
    with the prediction; with a limit below every sampled pdf no path is, and the image is the
    plain tracer's even though queries are on and the network is trained."""
    scene, cam, tr = box
    plain = pt_amd.Tracer(scene, W, H, max_bounces=8)
    plain.set_camera(cam)
    plain.render(6)
    tr.cache_configure(query=True, seed=6, train_stride=8, train_steps=2, batch=4096, query_max_pdf=1e-30)
    tr.set_camera(cam)
    tr.render(2)  # trains
    tr.set_camera(cam)
    for _ in range(6):
        tr.render(1)
    assert tr.cache_stats()["steps"] > 0
    np.testing.assert_array_equal(tr.read(), plain.read())
    with pytest.raises(pt_amd.PtError, match="cache configuration"):
        tr.cache_configure(query_max_pdf=-1.0)
    plain.close()


def test_cached_estimate_gains_at_equal_samples(box):
    """The cached estimator (paths end at their 2nd vertex with the prediction) has a lower
    error against the converged plain image than the plain estimator at the same sample
    count, and traces fewer rays."""
    scene, cam, tr = box
    ref_tr = pt_amd.Tracer(scene, W, H, max_bounces=8)
    ref_tr.set_camera(cam)
    ref_tr.render(4096)
    ref = ref_tr.read()
    ref_tr.close()

    spp = 16
    plain = pt_amd.Tracer(scene, W, H, max_bounces=8)
    plain.set_camera(cam)
    plain.render(spp)
    e_plain = pt_amd.rel_mse(plain.read(), ref)
    rays_plain = plain.stats()["extension_rays"]
    plain.close()

    tr.cache_configure(query=True, seed=2, train_stride=16, train_steps=8, batch=8192)
    tr.set_camera(cam)
    for _ in range(64):  # warm-up frames train the network
        tr.render(1)
    tr.set_camera(cam)  # restart the accumulation with the trained network
    tr.reset_stats()
    for _ in range(spp):
        tr.render(1)
    img = tr.read()
    e_cache = pt_amd.rel_mse(img, ref)
    rays_cache = tr.stats()["extension_rays"]
    print(f"relMSE plain {e_plain:.3e} cached {e_cache:.3e}; rays plain {rays_plain} cached {rays_cache}")
    assert np.all(np.isfinite(img))
    assert e_cache < e_plain
    assert rays_cache < rays_plain
