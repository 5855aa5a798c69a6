"""VTK blocks made on the device from the macro buffer equal the host writer's bytes for the
same macros (bit for bit, velocity scaling included), after a run with a body."""
from __future__ import annotations

import numpy as np
import pytest

from pocketwind_amd.field_export import LatticeField, field_vtk_bytes, solver_field_vtk
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.lattice import EQ, SOLID

from grids import grid_from_kinds

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("speed", [20.0, 37.3])
def test_device_blocks_equal_the_host_writer(speed):
    layout = (48, 24, 20)
    nx, ny, nz = layout
    idx = np.arange(nx * ny * nz)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    kinds = np.zeros(idx.size, np.uint32)
    kinds[(x >= 12) & (x < 18) & (y >= 9) & (y < 15) & (z >= 7) & (z < 13)] = SOLID
    kinds[(x == 0) | (x == nx - 1)] = EQ
    grid = grid_from_kinds(layout, kinds)
    solver = GpuSolver.create(grid.layout, grid, SolverParams(0.6, 0.05, True, operator="reg"))
    try:
        solver.step(60, True)
        device = solver_field_vtk(solver, 0.004, speed)
        macro = solver.read_macro()
    finally:
        solver.destroy()
    host = field_vtk_bytes(LatticeField(nx, ny, nz, 0.004, speed, macro[:, 0].copy(), macro[:, 1:4].reshape(-1).copy()))
    assert len(device) == len(host)
    assert device == host
    assert np.any(macro[:, 1] != 0.05)
