"""Host parts of the yaw sweep: the yaw turn of a placed body and the force coefficients."""
from __future__ import annotations

import math

import numpy as np
import pytest

from pocketwind_amd import bodies
from pocketwind_amd.sweep import DEFAULT_YAWS, coefficients, yawed


def placed_body():
    body, _ = bodies.place_mesh(bodies.bluff_body(), None, 32, (128, 64, 64), False)
    return body


def test_zero_yaw_keeps_the_body():
    body = placed_body()
    assert np.array_equal(yawed(body, 0.0), body)


def test_yaw_is_a_rigid_turn_about_the_vertical_axis():
    body = placed_body()
    v0 = body.reshape(-1, 3).astype(np.float64)
    for yaw in (7.5, 14.0, 90.0):
        v = yawed(body, yaw).reshape(-1, 3).astype(np.float64)
        # heights unchanged, horizontal distances between vertices unchanged
        assert np.array_equal(v[:, 2], v0[:, 2])
        i, j = 3, len(v0) // 2
        assert math.isclose(np.linalg.norm(v[i, :2] - v[j, :2]), np.linalg.norm(v0[i, :2] - v0[j, :2]),
                            rel_tol=1e-6)
        # the turn angle of one chord
        a0 = math.atan2(*(v0[j, :2] - v0[i, :2])[::-1])
        a1 = math.atan2(*(v[j, :2] - v[i, :2])[::-1])
        turned = math.degrees((a0 - a1 + math.pi) % (2 * math.pi) - math.pi)
        assert abs(turned - yaw) < 1e-3


def test_positive_yaw_turns_the_nose_towards_plus_y():
    body = placed_body()
    v0 = body.reshape(-1, 3)
    nose = int(np.argmin(v0[:, 0]))
    v = yawed(body, 10.0).reshape(-1, 3)
    assert v[nose, 1] > v0[nose, 1]
    # 90 degrees swaps the horizontal extents
    ext0 = v0.max(0) - v0.min(0)
    ext = yawed(body, 90.0).reshape(-1, 3)
    ext = ext.max(0) - ext.min(0)
    assert np.allclose(ext[[1, 0, 2]], ext0, rtol=1e-5)


def test_coefficients_and_default_yaws():
    cd, cs = coefficients(np.array([0.5, -0.25, 0.0]), 0.05, 400.0)
    assert cd == pytest.approx(0.5 / (0.5 * 0.05 ** 2 * 400))
    assert cs == pytest.approx(-0.5 * cd)
    assert list(DEFAULT_YAWS) == [0, 2, 4, 6, 8, 10, 12, 14] and len(DEFAULT_YAWS) == 8
