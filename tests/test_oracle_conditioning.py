"""Conditioning of the population-level comparison under the LES wall law, in the FP64 oracle
alone: why the row-wavefront parity test (test_gpu_row_parity.py) runs REG + LES on a sphere.

The wall law of src/lbm/shaders/stream_collide.wgsl:18-80 (collision.ts:162-224) switches
discontinuously between the prescribed-stress collision and the bulk collision (slip below
1e-7, a flat-face norm below 0.05, a negative post-collision population). On the 64 x 32 x 32
layout of the spec's LCG blob (gpu-vs-cpu.spec.ts:76-116) one nearly enclosed wall cell sits
close to such a switch: a relative perturbation of 1e-7 (the size of f32 rounding) in the
oracle's own state grows above 1e-6 within 100 steps. On a sphere in the same tunnel, and
without the wall law (TRT), the same perturbation decays.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from grids import blob_kinds, grid_from_kinds
from test_gpu_row_parity import LAYOUT, sphere_tunnel_kinds
from test_gpu_solver import normalized_peak


def growth(kinds, operator, les, steps=(50, 100)):
    grid = grid_from_kinds(LAYOUT, kinds)
    cells = grid.cells

    def make():
        return oracle.Solver(LAYOUT, 0.8, operator, les, oracle.BB, flags=grid.geometry[:cells],
                             equilibrium_state=(1.0, (0.06, 0.0, 0.0)), threads=4)
    a, b = make(), make()
    p = b.populations
    p[:] = p * (1 + 1e-7 * np.random.default_rng(1).standard_normal(p.shape))
    done, out = 0, []
    for step in steps:
        a.step(step - done)
        b.step(step - done)
        done = step
        out.append(normalized_peak(b.populations, a.populations))
    return out


@pytest.mark.slow
def test_wall_law_amplifies_perturbations_on_the_64_long_blob():
    assert max(growth(blob_kinds(LAYOUT), oracle.REG, True)) > 1e-6


@pytest.mark.slow
def test_perturbations_decay_on_the_sphere_and_without_the_wall_law():
    assert max(growth(sphere_tunnel_kinds(), oracle.REG, True)) < 3e-7
    assert max(growth(blob_kinds(LAYOUT), oracle.TRT, False)) < 3e-7
