"""Host side of the instability guard, following tests/unit/nan-guard.test.ts:30-66: the
density bounds (both endpoints stable) and the instability message."""
from __future__ import annotations

import math

import pytest

from pocketwind_amd.gpu_solver import NAN_GUARD_INTERVAL, is_unstable_density, unstable_message


@pytest.mark.parametrize("rho", [0.5, 1.0, 1.5])
def test_accepts_physical_density(rho):
    assert not is_unstable_density(rho)


@pytest.mark.parametrize("rho", [math.nan, math.inf, -math.inf, 0.5 - 2 ** -25, 1.5 + 2 ** -23, 0.0, -1.0])
def test_rejects_non_finite_or_out_of_range_density(rho):
    assert is_unstable_density(rho)


def test_message_names_the_step_and_both_parts():
    assert NAN_GUARD_INTERVAL == 200
    assert unstable_message(600) == {
        "step": 600,
        "what": "The run went unstable at step 600 and was reset",
        "next": "Lower the speed, or choose a higher tier under Advanced",
    }
