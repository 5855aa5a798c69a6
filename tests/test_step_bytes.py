"""The step's algorithmic bytes (pocketwind_amd.lattice.step_bytes), which the measurement reports
against HBM peak: populations read and written, one descriptor per 64-cell run on row-path
lattices, flags only for runs that touch a body, every flag on per-cell lattices."""
from __future__ import annotations

import numpy as np

from pocketwind_amd.lattice import EQ, SOLID, VoxelGrid, general_runs, step_bytes


def grid(nx, ny, nz, flags):
    cells = nx * ny * nz
    geometry = np.zeros(2 * cells, np.uint32)
    geometry[:cells] = flags
    return VoxelGrid(nx, ny, nz, geometry)


def test_plain_runs_read_one_descriptor_each():
    g = grid(64, 2, 1, np.zeros(128, np.uint32))
    assert general_runs(g.flags) == 0
    assert step_bytes(g, "f16") == 2 * 19 * 2 * 128 + 4 * 2
    assert step_bytes(g, "f32") == 2 * 19 * 4 * 128 + 4 * 2


def test_inlet_and_outlet_cells_keep_a_run_plain():
    f = np.zeros(128, np.uint32)
    f[0] = f[63] = f[64] = f[127] = EQ
    assert general_runs(f) == 0
    f[1] = EQ  # an EQ cell inside a run makes it general
    assert general_runs(f) == 1


def test_runs_touching_a_body_read_their_flags():
    f = np.zeros(128, np.uint32)
    f[70] = SOLID
    f[71] = 1 << 4  # a wall link
    g = grid(64, 2, 1, f)
    assert general_runs(g.flags) == 1
    assert step_bytes(g, "f16") == 2 * 19 * 2 * 128 + 4 * 2 + 4 * 64


def test_per_cell_lattices_read_every_flag():
    g = grid(32, 2, 2, np.zeros(128, np.uint32))
    assert step_bytes(g, "f16") == 2 * 19 * 2 * 128 + 4 * 128
