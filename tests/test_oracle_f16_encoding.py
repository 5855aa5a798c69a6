"""The oracle's f16 population encoding (prelude-f16.wgsl:7-22, stored as common.wgsl:134-135)
against what the format fixes: values already representable as halves pass unchanged; values
below the smallest normal half convert to nearest even (numpy's float16 cast, a library
routine); every other value lands on one of its two neighbouring halves, and over cells it
rounds up with the probability of its dropped fraction (so the stored mean is unbiased). A
mistake in the comparison's direction, the mask or the increment fails one of these."""
from __future__ import annotations

import numpy as np
import pytest

import oracle


def encode_all(values: np.ndarray, i: int = 3) -> np.ndarray:
    """encode_f16(v, i, n) of values[n] (already scaled), one cell each."""
    values = np.asarray(values, np.float32)
    g = np.zeros((19, values.size), np.float32)
    g[i] = values / np.float32(32768)  # exact: a power of two
    return oracle.encode_f16_planes(g)[i]


def test_exact_halves_pass_unchanged():
    rng = np.random.default_rng(1)
    halves = rng.integers(0x0400, 0x7bff, 4096).astype(np.uint16)  # positive normal halves
    halves = np.concatenate([halves, halves | 0x8000])
    values = halves.view(np.float16).astype(np.float32)
    assert np.array_equal(encode_all(values), halves)


def test_subnormal_range_rounds_to_nearest_even():
    rng = np.random.default_rng(2)
    values = (rng.uniform(-1, 1, 20000) * 2.0 ** -14).astype(np.float32)
    values = np.concatenate([values, np.float32([0.0, -0.0, 2.0 ** -25, 3 * 2.0 ** -26, 2.0 ** -14 * 0.99999])])
    expect = values.astype(np.float16).view(np.uint16)
    assert np.array_equal(encode_all(values), expect)


def neighbours(values: np.ndarray):
    """The halves just below and above |v| (toward zero, away from zero), with v's sign."""
    bits = values.view(np.uint32)
    sign = (bits >> 31).astype(np.uint16) << 15
    down = np.abs(values).astype(np.float64)
    low = np.abs(values.astype(np.float16)).astype(np.float64)  # nearest, then step toward zero if above
    low16 = np.abs(values.astype(np.float16)).view(np.uint16)
    low16 = np.where(low > down, low16 - 1, low16).astype(np.uint16)
    return low16 | sign, (low16 + 1).astype(np.uint16) | sign


@pytest.mark.parametrize("scale", [2.0 ** -13, 1.0, 300.0, 20000.0])
def test_normal_values_land_on_a_neighbour(scale):
    rng = np.random.default_rng(3)
    values = (rng.uniform(-1, 1, 50000) * scale).astype(np.float32)
    values = values[np.abs(values) >= 2.0 ** -14]
    got = encode_all(values)
    low, high = neighbours(values)
    assert np.all((got == low) | (got == high))
    # a value between two halves takes both across cells
    inexact = values.astype(np.float16).astype(np.float32) != values
    assert np.any(got[inexact] == low[inexact]) and np.any(got[inexact] == high[inexact])


@pytest.mark.parametrize("fraction", [1 / 8, 1 / 4, 1 / 2, 3 / 4, 8191 / 8192])
def test_rounds_up_with_the_dropped_fraction(fraction):
    # 1 + fraction * 2^-10 lies between the halves 1 and 1 + 2^-10; its dropped 13 bits are
    # fraction * 2^13. Over 2^17 cells the share rounded up is the fraction within 4 sigma.
    value = np.float32(1 + fraction * 2.0 ** -10)
    cells = 1 << 17
    got = encode_all(np.full(cells, value, np.float32))
    low, high = np.float16(1.0).view(np.uint16), (np.float16(1.0).view(np.uint16) + 1)
    assert set(np.unique(got)) <= {low, high}
    share = float(np.mean(got == high))
    sigma = (fraction * (1 - fraction) / cells) ** 0.5
    assert abs(share - fraction) < 4 * sigma + 1e-12, (share, fraction)


def test_cells_and_directions_draw_different_words():
    # The same value in different cells and directions rounds differently somewhere: the word
    # depends on n and i as well as the value.
    value = np.float32(1 + 0.5 * 2.0 ** -10)
    g = np.full((19, 4096), value / np.float32(32768), np.float32)
    out = oracle.encode_f16_planes(g)
    assert len(np.unique(out)) == 2
    assert not np.all(out == out[0]) and not np.all(out == out[:, :1])
