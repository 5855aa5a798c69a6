"""The instability guard fused into the macro-writing step (nan_guard.wgsl, nan-guard.ts).

* A stable run passes every checkpoint and keeps its step count.
* NaN injected after step 400 is caught at the step-600 checkpoint and the run is reset to
  equilibrium (nan-guard.ts:121-122: "step 400 is caught at 600"), for every storage.
* The fused check flags the same checkpoint as the reference's separate pass would: stepping
  an over-driven tunnel in 200-step batches and scanning every cell's macro density on the
  host finds the first unstable checkpoint at the step the guarded run reports.
"""
from __future__ import annotations

import numpy as np
import pytest

from pocketwind_amd.gpu_solver import NAN_GUARD_INTERVAL, GpuSolver, SolverParams, is_unstable_density
from pocketwind_amd.lattice import EQ, SOLID

from grids import grid_from_kinds

pytestmark = pytest.mark.gpu

LAYOUT = (64, 32, 32)


def tunnel():
    nx, ny, nz = LAYOUT
    idx = np.arange(nx * ny * nz)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    kinds = np.zeros(idx.size, np.uint32)
    kinds[(x >= 16) & (x < 24) & (y >= 12) & (y < 20) & (z >= 12) & (z < 20)] = SOLID
    kinds[(x == 0) | (x == nx - 1)] = EQ
    return grid_from_kinds(LAYOUT, kinds)


def make(params):
    grid = tunnel()
    solver = GpuSolver.create(grid.layout, grid, params)
    assert isinstance(solver, GpuSolver)
    return solver


def test_stable_run_passes_every_checkpoint():
    solver = make(SolverParams(0.6, 0.05, False))
    try:
        assert solver.advance_guarded(450) is None
        assert solver.step_count == 450
        assert solver.advance_guarded(350) is None
        assert solver.step_count == 800
        assert np.all(np.isfinite(solver.read_macro()))
    finally:
        solver.destroy()


@pytest.mark.parametrize("precision", ["f32", "f16", "bf16"])
def test_injected_nan_is_caught_at_the_next_checkpoint_and_reset(precision):
    params = SolverParams(0.6, 0.05, False, precision=precision)
    solver = make(params)
    fresh = make(params)
    try:
        message = solver.advance_guarded(1000, inject_nan=400)
        assert message is not None and message["step"] == 600
        assert message["what"] == "The run went unstable at step 600 and was reset"
        assert solver.step_count == 0
        assert np.array_equal(solver.read_populations(), fresh.read_populations())
        # The reset run steps on and stays stable.
        assert solver.advance_guarded(200) is None
    finally:
        solver.destroy()
        fresh.destroy()


def test_fused_check_matches_a_scan_of_the_macros():
    params = SolverParams(0.5005, 0.35, False)
    guarded = make(params)
    scanned = make(params)
    try:
        caught = None
        while caught is None and guarded.step_count < 20 * NAN_GUARD_INTERVAL:
            caught = guarded.advance_guarded(NAN_GUARD_INTERVAL)
        assert caught is not None, "the over-driven tunnel never went unstable"
        first = None
        for k in range(1, caught["step"] // NAN_GUARD_INTERVAL + 1):
            scanned.step(NAN_GUARD_INTERVAL, True)
            if any(map(is_unstable_density, scanned.read_macro()[:, 0].tolist())):
                first = k * NAN_GUARD_INTERVAL
                break
        assert first == caught["step"]
    finally:
        guarded.destroy()
        scanned.destroy()
