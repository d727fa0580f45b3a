"""Pins the NumPy MLP model (tests/mlp_ref.py) that checks the device cache kernels."""
import numpy as np
import pytest
import torch

import mlp_ref


def _net(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    _, count = mlp_ref.layout()
    p = np.zeros(count)
    for w, b, ni, no in mlp_ref.layout()[0]:
        p[w:w + ni * no] = rng.uniform(-1, 1, ni * no) * np.sqrt(6.0 / ni) * scale
        p[b:b + no] = rng.uniform(-0.1, 0.1, no)
    return p


def test_layout_count():
    lay, count = mlp_ref.layout()
    assert count == 96 * 64 + 64 + 4 * (64 * 64 + 64) + 64 * 3 + 3 == 22720 + 323
    assert lay[-1][1] + 3 == count


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0, np.inf, -np.inf])])
    ours = mlp_ref.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(ours.view(np.uint32), ref.view(np.uint32))
    # ties go to the even mantissa
    assert ours[-8] == 1.0 and ours[-7] == 1.0 and ours[-6] == np.float32(1 + 2 ** -6)


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = _net(3)
    X = rng.uniform(-1, 1, (6, 96))
    T = rng.uniform(0, 2, (6, 3))
    L, g = mlp_ref.loss_and_grad(p, X, T)
    y, _ = mlp_ref.forward(p, X)
    den = y * y + 0.01  # the loss holds its denominator constant
    lay, _ = mlp_ref.layout()
    idx = []
    for w, b, ni, no in lay:  # a few weights and biases of every layer
        idx += list(rng.integers(w, w + ni * no, 6)) + list(rng.integers(b, b + no, 2))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss(pp, X, T, den=den)[0] - mlp_ref.loss(pm, X, T, den=den)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
    assert L == pytest.approx(mlp_ref.loss(p, X, T)[0])


def test_gradient_is_zero_at_the_target():
    rng = np.random.default_rng(4)
    p = _net(5)
    X = rng.uniform(-1, 1, (9, 96))
    y, _ = mlp_ref.forward(p, X)
    L, g = mlp_ref.loss_and_grad(p, X, y)
    assert L == 0.0 and not np.any(g)


def test_bf16_mode_close_to_exact():
    rng = np.random.default_rng(6)
    p = _net(7)
    X = rng.uniform(-1, 1, (64, 96))
    T = rng.uniform(0, 2, (64, 3))
    ye, _ = mlp_ref.forward(p, X)
    yb, _ = mlp_ref.forward(p, X, "bf16")
    assert np.linalg.norm(yb - ye) <= 3e-2 * np.linalg.norm(ye)
    _, ge = mlp_ref.loss_and_grad(p, X, T)
    _, gb = mlp_ref.loss_and_grad(p, X, T, "bf16")
    assert np.linalg.norm(gb - ge) <= 5e-2 * np.linalg.norm(ge)


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([1e-3, -2.0, 5.0, -1e-4])
    p, m, v = mlp_ref.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)


def test_he_uniform_bounds_and_moments():
    p = mlp_ref.he_uniform(1)
    for w, b, ni, no in mlp_ref.layout()[0]:
        W = p[w:w + ni * no]
        lim = np.float32(np.sqrt(np.float32(6.0) / np.float32(ni)))
        assert np.all(np.abs(W) <= lim) and not np.any(p[b:b + no])
        assert abs(W.mean()) < 0.05 * lim
        assert W.var() == pytest.approx(lim * lim / 3, rel=0.15)
    assert not np.array_equal(p, mlp_ref.he_uniform(2))


def test_encode_octaves():
    rec = np.zeros((1, 13))
    rec[0, :3] = [0.25, 0.5, 0.75]  # inside the box [0, 1]^3
    rec[0, 3:13] = np.arange(10) + 1
    f = mlp_ref.encode(rec, [0, 0, 0], [1, 1, 1])[0]
    assert f[0] == pytest.approx(np.sin(np.pi * 0.25)) and f[1] == pytest.approx(np.cos(np.pi * 0.25))
    assert f[2] == pytest.approx(1.0) and f[3] == pytest.approx(np.cos(np.pi * 0.5), abs=1e-12)
    assert f[6] == pytest.approx(np.sin(2 * np.pi * 0.25))  # octave 1, x
    np.testing.assert_allclose(f[72:75], [0.25, 0.5, 0.75])
    np.testing.assert_allclose(f[75:85], np.arange(10) + 1)
    assert not np.any(f[85:])
