"""A reduced seeded tunnel set: every tunnel stepped in the batch gives the force samples of the
same tunnel stepped alone, bit for bit."""
from __future__ import annotations

import pytest

from pocketwind_amd import tunnel_set

pytestmark = pytest.mark.gpu


def test_reduced_tunnel_set_batch_equals_singles():
    rec = tunnel_set.run(count=8, steps=400, seed=11, timing_steps=50, log=lambda m: print(m, flush=True))
    assert rec["identical_tunnels"] == 8
    assert rec["finite_tunnels"] == 8 and rec["max_cd_difference"] == 0.0
    assert rec["batch_mlups"] > 0 and rec["single_mlups"] > 0
