"""GPU FP64 momentum-exchange drag against the oracle (restates tests/browser/force.spec.ts).

* seeded 32^3 blob, TRT, checkpoints 0/1/8/16/50/100: relative error < 1e-4 on one vector
  scale floored at f32 resolution; the sample after reset equals the step-0 sample;
* 70 000 synthetic boundary entries (cell 61 x 69 999 + cell 57) with patched populations:
  fx = 69 999 * 0.01, fy = 0.02, fz = 0 exactly;
* an empty boundary list gives an exactly zero force;
* the slot semantics: 'busy' with nothing encoded, an occupied slot skips encodes.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from grids import blob_grid, blob_kinds, grid_from_kinds
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.lattice import FLOOR, SOLID, SUPPORT

pytestmark = pytest.mark.gpu

F32_EPS = 2.0 ** -23
PARAMS = SolverParams(0.8, 0.06, False)


def assert_force(sample, expected, step, tol=1e-4):
    assert sample.step == step
    scale = max(F32_EPS, *np.abs(expected))
    err = max(abs(a - b) for a, b in zip((sample.fx, sample.fy, sample.fz), expected))
    assert err / scale < tol, (sample, expected)


@pytest.mark.parametrize("walls,precision,tol", [("bb", "f32", 1e-4), ("bouzidi", "f32", 1e-4),
                                                 ("bb", "f16", 5e-3)])
def test_blob_force_matches_oracle(walls, precision, tol):
    grid = blob_grid(walls)
    cells = grid.cells
    cpu = oracle.Solver(grid.layout, 0.8, oracle.TRT, False,
                        oracle.BOUZIDI if walls == "bouzidi" else oracle.BB,
                        flags=grid.geometry if walls == "bouzidi" else grid.geometry[:cells],
                        equilibrium_state=(1.0, (0.06, 0.0, 0.0)), threads=4)
    checkpoints = [0, 1, 8, 16, 50, 100]
    expected, done = [], 0
    for step in checkpoints:
        cpu.step(step - done)
        done = step
        expected.append(cpu.momentum_exchange())
    solver = GpuSolver.create(grid.layout, grid, SolverParams(0.8, 0.06, False, walls=walls,
                                                              precision=precision))
    try:
        assert solver.read_force() == "busy"
        for step, ref in zip(checkpoints, expected):
            solver.step(step - solver.step_count, False)
            solver.encode_force()
            solver.encode_force()  # the slot is occupied: skipped
            solver.step(1, False)
            solver.encode_force()  # still occupied
            sample = solver.read_force()
            assert sample != "busy"
            assert_force(sample, ref, step, tol)
            assert solver.read_force() == "busy"
        assert expected[-1][0] > 0
        solver.reset()
        solver.encode_force()
        assert_force(solver.read_force(), expected[0], 0, tol)
    finally:
        solver.destroy()


def test_reset_discards_unread_sample():
    grid = blob_grid("bb")
    solver = GpuSolver.create(grid.layout, grid, PARAMS)
    try:
        solver.step(8, False)
        solver.encode_force()
        solver.reset()
        assert solver.read_force() == "busy"
    finally:
        solver.destroy()


def test_force_is_bit_identical_run_to_run():
    grid = blob_grid("bb")
    samples = []
    for _ in range(2):
        solver = GpuSolver.create(grid.layout, grid, PARAMS)
        solver.step(40, False)
        solver.encode_force()
        samples.append(solver.read_force())
        solver.destroy()
    assert samples[0] == samples[1]


def test_70000_synthetic_boundary_entries():
    layout = (5, 5, 5)
    kinds = np.zeros(125, np.uint32)
    kinds[62] = SOLID
    kinds[36] = SUPPORT
    kinds[56] = FLOOR
    grid = grid_from_kinds(layout, kinds)
    grid.boundary_cells = np.full(70000, 61, np.uint32)
    grid.boundary_cells[69999] = 57
    solver = GpuSolver.create(layout, grid, SolverParams(0.8, 0.0, False))
    try:
        pops = solver.read_populations()
        for direction, cell, value in ((1, 61, 0.01), (3, 57, 0.02), (4, 61, 0.5), (6, 61, 0.25)):
            pops[direction, cell] = value
        solver.write_populations(pops)
        solver.encode_force()
        s = solver.read_force()
        assert s.step == 0
        assert abs(s.fx - 69999 * 0.01) / (69999 * 0.01) < 1e-4
        assert abs(s.fy - 0.02) / 0.02 < 1e-4
        assert abs(s.fz) == 0
    finally:
        solver.destroy()


def test_empty_boundary_gives_zero_force():
    grid = grid_from_kinds((4, 4, 4), np.zeros(64, np.uint32))
    solver = GpuSolver.create((4, 4, 4), grid, PARAMS)
    try:
        for step in (0, 8):
            solver.step(step - solver.step_count, False)
            solver.encode_force()
            s = solver.read_force()
            assert (abs(s.fx), abs(s.fy), abs(s.fz)) == (0, 0, 0)
        solver.reset()
        solver.encode_force()
        s = solver.read_force()
        assert (abs(s.fx), abs(s.fy), abs(s.fz)) == (0, 0, 0)
    finally:
        solver.destroy()
