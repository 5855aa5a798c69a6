"""The FP64 MLP oracle pinned without the device: analytic gradients against central finite
differences, the one-layer closed form, bf16 rounding against a direct bit-level statement,
and Adam's first step, whose size is lr in every coordinate with a nonzero gradient."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp


def sample(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, mlp.IN)), rng.normal(size=(n, mlp.OUT))


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    flat = mlp.init_params(seed) + 0.05 * np.random.default_rng(seed + 9).normal(size=mlp.N_PARAMS)
    x, t = sample(7, seed)
    _, g = mlp.loss_and_gradients(flat, x, t)
    rng = np.random.default_rng(seed + 3)
    probe = rng.choice(mlp.N_PARAMS, 60, replace=False)
    h = 1e-6
    for k in probe:
        e = np.zeros(mlp.N_PARAMS)
        e[k] = h
        lp, _ = mlp.loss_and_gradients(flat + e, x, t)
        lm, _ = mlp.loss_and_gradients(flat - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_output_layer_gradient_is_the_least_squares_closed_form():
    # With W1, W2 fixed the loss is quadratic in (W3, b3); its gradient is (2 / (B m)) (Y - T)^T [H2 1].
    flat = mlp.init_params(5)
    x, t = sample(11, 5)
    y, c = mlp.forward(flat, x)
    _, g = mlp.loss_and_gradients(flat, x, t)
    p = mlp.unpack(g)
    expect_w3 = 2.0 * (y - t).T @ c["h2"] / (11 * mlp.OUT)
    expect_b3 = 2.0 * (y - t).sum(0) / (11 * mlp.OUT)
    np.testing.assert_allclose(p["W3"], expect_w3, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(p["b3"], expect_b3, rtol=1e-12, atol=1e-15)


def test_bf16_rounding_is_round_to_nearest_even():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3 * 2^-8 -> 1 + 2^-6 (even).
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 3.0e-3])
    got = mlp.bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1 + 2 ** -6
    assert got[3] == 1 + 2 ** -7 and got[4] == -2.5
    assert abs(got[5] / 3.0e-3 - 1) < 2 ** -8


def test_adam_first_step_moves_every_coordinate_by_lr():
    flat = mlp.init_params(2)
    g = np.random.default_rng(4).normal(size=mlp.N_PARAMS)
    new, m, v = mlp.adam(flat, g, np.zeros_like(g), np.zeros_like(g), 1, 1e-3, eps=0.0)
    np.testing.assert_allclose(flat - new, 1e-3 * np.sign(g), rtol=1e-9)


def test_bf16_points_change_the_loss_only_at_bf16_precision():
    flat = mlp.init_params(3)
    x, t = sample(64, 3)
    l64, g64 = mlp.loss_and_gradients(flat, x, t)
    l16, g16 = mlp.loss_and_gradients(flat, x, t, bf16_points=True)
    assert abs(l16 / l64 - 1) < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2
