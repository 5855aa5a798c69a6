"""The closure network's HIP kernels (bf16 MFMA, f32 accumulation) against the FP64 oracle in
oracle/mlp.py: the forward pass and the full-batch gradients match the oracle that rounds to
bf16 at the kernels' rounding points, an Adam step matches the oracle's update of the same
gradient, and training on the device fits a smooth target, reproducibly."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as O
from pocketwind_amd.closure import N_PARAMS, Mlp, init_params

pytestmark = pytest.mark.gpu


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 12)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4]),
                  0.3 * x[:, 5] ** 2 - 0.2 * x[:, 6] + 0.1 * x[:, 11]], axis=1).astype(np.float32)
    return x, t


def params(seed):
    p = init_params(seed).astype(np.float64)
    p += 0.02 * np.random.default_rng(seed + 100).normal(size=p.size)  # nonzero biases too
    return p.astype(np.float32)


def test_parameter_layout_matches_the_oracle():
    assert N_PARAMS == O.N_PARAMS == 5187


@pytest.mark.parametrize("rows", [1, 17, 1000])
def test_forward_matches_bf16_oracle(rows):
    p = params(1)
    x, _ = data(rows, 2)
    net = Mlp(p)
    try:
        y = net.forward(x)
    finally:
        net.destroy()
    ref, _ = O.forward(p, x, bf16_points=True)
    scale = max(1.0, float(np.abs(ref).max()))
    err = float(np.abs(y - ref).max()) / scale
    print(f"forward rows={rows}: max error {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("rows", [32, 777, 4096])
def test_gradients_match_oracle(rows):
    p = params(3)
    x, t = data(rows, 4)
    net = Mlp(p)
    try:
        loss, g = net.gradients(x, t)
    finally:
        net.destroy()
    ref_loss, ref = O.loss_and_gradients(p, x, t, bf16_points=True)
    _, exact = O.loss_and_gradients(p, x, t)
    assert abs(loss / ref_loss - 1) < 1e-3
    parts = O.unpack(g)
    for name, want in O.unpack(ref).items():
        rel = np.linalg.norm(parts[name] - want) / np.linalg.norm(want)
        print(f"rows={rows} {name}: relative error vs bf16 oracle {rel:.2e}")
        assert rel < 2e-4, name
    assert np.linalg.norm(g - exact) / np.linalg.norm(exact) < 5e-2


def test_adam_step_matches_oracle_update():
    p = params(5)
    x, t = data(512, 6)
    net = Mlp(p)
    try:
        _, g = net.gradients(x, t)
        net.set_data(x, t)
        net.train(1, 0, 1e-3)
        after = net.params
    finally:
        net.destroy()
    want, _, _ = O.adam(p.astype(np.float64), g.astype(np.float64), np.zeros(p.size), np.zeros(p.size), 1, 1e-3)
    np.testing.assert_allclose(after, want, rtol=0, atol=2e-6)


def test_training_fits_and_is_reproducible():
    x, t = data(20000, 7)
    runs = []
    for _ in range(2):
        net = Mlp(init_params(11))
        try:
            net.set_data(x, t)
            losses = net.train(1500, 256, 3e-3, seed=9)
            runs.append((losses, net.params, net.forward(x[:2000])))
        finally:
            net.destroy()
    losses, p, y = runs[0]
    first, last = float(losses[:20].mean()), float(losses[-50:].mean())
    print(f"training loss {first:.4f} -> {last:.4f}")
    assert last < 0.1 * first
    assert np.array_equal(runs[0][1], runs[1][1]), "training is not reproducible"
    fit = float(np.mean((y - t[:2000]) ** 2))
    assert fit < 0.1 * float(np.mean(t[:2000] ** 2))


def test_rejects_bad_shapes():
    net = Mlp()
    try:
        with pytest.raises(ValueError):
            net.params = np.zeros(N_PARAMS - 1, np.float32)
        with pytest.raises(ValueError):
            net.forward(np.zeros((3, 11), np.float32))
    finally:
        net.destroy()
