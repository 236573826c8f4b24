"""The warm-start network's oracle (numpy float64) pinned by finite
differences and closed forms, and the product's host-side layout (CPU)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.mlp as om
from hl2sdk_tools_sharp_amd import mlp


def small_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(om.N_PARAMS) * scale / np.sqrt(64)


def test_parameter_count_and_layout_agree():
    assert om.N_PARAMS == 36739 == mlp.N_PARAMS
    p = mlp.init_params(3, out_bias=[1.0, 2.0, 3.0])
    L = om.unpack(p)
    assert [W.shape for W, _ in L] == [(25, 128), (128, 128), (128, 128), (128, 3)]
    assert np.array_equal(L[3][1], [1.0, 2.0, 3.0])
    assert all(not b.any() for _, b in L[:3])
    np.testing.assert_allclose(L[1][0].std(), np.sqrt(2 / 128), rtol=0.05)
    assert np.array_equal(om.pack(L), p.astype(np.float64))
    assert np.array_equal(om.fourier_matrix(), mlp.fourier_matrix())


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    w = small_params(1)
    x = rng.standard_normal((6, om.N_IN))
    t = rng.standard_normal((6, 3))
    _loss, g = om.loss_and_grad(w, x, t)
    idx = np.concatenate([rng.choice(om.N_PARAMS, 300, replace=False),
                          [0, 25 * 128, om.N_PARAMS - 1, om.N_PARAMS - 4]])  # first W, first b, last b, last W
    h = 1e-6
    for i in idx:
        e = np.zeros(om.N_PARAMS)
        e[i] = h
        fd = (om.loss_and_grad(w + e, x, t)[0] - om.loss_and_grad(w - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_closed_form():
    w = np.zeros(om.N_PARAMS)
    L = om.unpack(w)
    L[3][1][:] = [1.0, 2.0, 3.0]
    w = om.pack(L)
    x = np.zeros((4, om.N_IN))
    t = np.zeros((4, 3))
    loss, g = om.loss_and_grad(w, x, t)
    assert loss == pytest.approx((1 + 4 + 9) / 3)
    gb = om.unpack(g)[3][1]
    np.testing.assert_allclose(gb, 2 * np.array([1.0, 2.0, 3.0]) / 3)


def test_bf16_rounding():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40], np.float32)
    r = om.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                 # tie, to even
    assert r[2] == 1 + 2 ** -6         # tie, to even (up)
    assert r[3] == 1 + 2 ** -7         # exact
    assert r[4] == -2.5
    assert np.isfinite(r[5]) and abs(r[5] - 3.0e38) <= 3.0e38 * 2 ** -8
    u = np.asarray(r, np.float32).view(np.uint32)
    assert np.all(u & 0xFFFF == 0)


def test_gelu_values_and_derivative():
    z = np.linspace(-6, 6, 101)
    assert om.gelu(0.0) == 0.0
    np.testing.assert_allclose(om.gelu(np.array([8.0])), [8.0], rtol=1e-12)
    np.testing.assert_allclose(om.gelu(np.array([-8.0])), [0.0], atol=1e-12)
    h = 1e-6
    np.testing.assert_allclose(om.gelu_grad(z), (om.gelu(z + h) - om.gelu(z - h)) / (2 * h), rtol=1e-6, atol=1e-9)


def test_adam_first_step_is_lr_sign():
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    w, m, v = om.adam(np.zeros(4), np.zeros(4), np.zeros(4), g, 1, lr=0.01)
    np.testing.assert_allclose(w, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-9)


def test_features_closed_forms():
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([100.0, 200.0, 50.0])
    o = np.array([[0.0, 0.0, 0.0], [100.0, 200.0, 50.0], [50.0, 100.0, 25.0]])
    n = np.array([[0, 0, 1.0]] * 3)
    d = np.array([[0.0, 0, 0], [np.e - 1, 0, -5], [1, 1, 1]])
    F = om.fourier_matrix()
    x = om.features(o, n, d, lo, hi, F, 1024.0)
    np.testing.assert_allclose(x[0, :3], -1.0)
    np.testing.assert_allclose(x[1, :3], 1.0)
    np.testing.assert_allclose(x[2, :3], 0.0)
    np.testing.assert_allclose(x[1, 6:9], [np.log(np.e - 1 + 1e-7), np.log(1e-7), np.log(1e-7)])
    np.testing.assert_allclose(x[:, 9:17] ** 2 + x[:, 17:25] ** 2, 1.0)
    np.testing.assert_allclose(x[0, 9:17], 0.0, atol=1e-15)
    np.testing.assert_allclose(x[0, 17:25], 1.0)
