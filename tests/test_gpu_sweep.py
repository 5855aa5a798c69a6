"""The yaw-sweep pipeline on a small tunnel: noisy mesh, device voxelisation per yaw, batched
f16 REG + LES tunnels with the learned closure, particles every step, drag per yaw."""
from __future__ import annotations

import math

import numpy as np
import pytest

from pocketwind_amd import sweep

pytestmark = pytest.mark.gpu


def test_small_sweep_runs_end_to_end():
    rec = sweep.run(layout=(128, 64, 64), yaws=(0, 10), length_cells=32, flow_throughs=1.0,
                    average_flow_throughs=0.5, particles=20000, mesh_levels=3,
                    log=lambda m: print(m, flush=True))
    assert rec["tunnels"] == 2 and rec["cells"] == 2 * 128 * 64 * 64
    assert rec["triangles"] == 212 * 4 ** 3
    assert rec["particles"] == 20000 and rec["particles_finite"]
    assert rec["force_samples"] == rec["average_steps"] // 8
    r0, r10 = rec["rows"]
    for r in rec["rows"]:
        assert math.isfinite(r["cd"]) and 0.1 < r["cd"] < 5
    # a body symmetric about y has no mean side force at zero yaw; turned, it has one
    assert abs(r10["cs"]) > 2 * abs(r0["cs"]) + 0.01
    assert rec["batched_step_mlups"] > 0 and rec["particle_updates_per_s"] > 0
