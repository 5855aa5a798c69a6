"""D3Q19 lattice, cell kinds and grid types shared by the host API.

Mirrors src/lbm/lattice.ts:2-52, src/geometry/voxel-grid.ts:1-15 and the tier table of
src/lbm/layout.ts:7-12.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np

DIRECTIONS = np.array(
    [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
     [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0],
     [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]], dtype=np.int64)
WEIGHTS = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def opp(i: int) -> int:
    """Index of the reversed velocity (lattice.ts:47-52)."""
    if not isinstance(i, (int, np.integer)) or i < 0 or i >= 19:
        raise ValueError("Direction index must be an integer from 0 to 18.")
    return 0 if i == 0 else (i + 1 if i % 2 == 1 else i - 1)


FLUID, SOLID, SUPPORT, FLOOR, EQ = 0, 1, 2, 3, 4
CELL_TYPE_MASK = 7
NO_Q = 0xFFFFFFFF


class Layout(NamedTuple):
    nx: int
    ny: int
    nz: int

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


TIERS = {
    "S": Layout(128, 64, 64),
    "M": Layout(192, 96, 96),
    "L": Layout(256, 128, 128),
    "XL": Layout(384, 192, 192),
}


@dataclass
class VoxelGrid:
    """Packed geometry [flags N][q offsets N][f32 q per link] plus counts (voxel-grid.ts:7-15)."""

    nx: int
    ny: int
    nz: int
    geometry: np.ndarray
    boundary_cells: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    solid_cells: int = 0
    frontal_cells: int = 0
    blockage: float = 0.0
    watertight: bool = True
    missed_links: int = 0

    @property
    def layout(self) -> Layout:
        return Layout(self.nx, self.ny, self.nz)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def flags(self) -> np.ndarray:
        return self.geometry[: self.cells]

    @property
    def fractions(self) -> np.ndarray:
        return self.geometry[2 * self.cells:].view(np.float32)


def general_runs(flags) -> int:
    """64-cell runs of a row-path lattice (nx a multiple of 64) whose cell flags a step reads:
    every run except those of fluid cells without wall links, with at most an EQ cell first and
    last (the step's run descriptors cover those)."""
    runs = np.asarray(flags, np.uint32).reshape(-1, 64)
    edge_ok = ((runs[:, 0] == 0) | (runs[:, 0] == EQ)) & ((runs[:, 63] == 0) | (runs[:, 63] == EQ))
    return int(np.count_nonzero(~(edge_ok & np.all(runs[:, 1:63] == 0, axis=1))))


def step_bytes(grid, precision: str) -> int:
    """Bytes one step has to move on `grid`: 19 populations read and 19 written per cell; on a
    row-path lattice one 4-byte descriptor per 64-cell run plus the 4-byte flags of the runs that
    touch a body, else every cell's flag. The reference counts a flag for every cell
    (src/app/bench-page.ts:44-45: 156 B per cell at f32, 80 B at f16)."""
    elem = 4 if precision == "f32" else 2
    cells = grid.nx * grid.ny * grid.nz
    pops = 2 * 19 * elem * cells
    if grid.nx % 64:
        return pops + 4 * cells
    return pops + 4 * (cells // 64) + 4 * 64 * general_runs(grid.flags)


def grid_from_kinds(layout, kinds) -> VoxelGrid:
    """Packed geometry of a lattice given by its cell kinds, with halfway walls (every q offset
    NO_Q): each fluid cell's flag carries the wall mask of its pulls (bit 3 + i when the cell at
    x - c_i, wrapping, is SOLID, SUPPORT or FLOOR), and the boundary list is the fluid cells with
    any wall link, as tests/browser/force.spec.ts:40-72 builds its grids."""
    nx, ny, nz = (int(v) for v in layout)
    cells = nx * ny * nz
    kinds = np.asarray(kinds, dtype=np.uint32).reshape(-1)
    if kinds.size != cells:
        raise ValueError("One kind per cell is needed.")
    geometry = np.zeros(2 * cells, dtype=np.uint32)
    geometry[:cells] = kinds
    geometry[cells:] = NO_Q
    # Rolling the wall field by c brings the source x - c of every cell to x.
    grid3 = kinds.reshape(nz, ny, nx) & CELL_TYPE_MASK
    wall = (grid3 >= SOLID) & (grid3 <= FLOOR)
    mask = np.zeros((nz, ny, nx), dtype=np.uint32)
    for i, (cx, cy, cz) in enumerate(DIRECTIONS):
        mask |= np.roll(wall, (int(cz), int(cy), int(cx)), axis=(0, 1, 2)).astype(np.uint32) << np.uint32(3 + i)
    fluid = kinds == FLUID
    geometry[:cells][fluid] = mask.reshape(-1)[fluid]
    boundary = np.flatnonzero(fluid & (mask.reshape(-1) != 0)).astype(np.uint32)
    return VoxelGrid(nx, ny, nz, geometry, boundary, solid_cells=int(np.sum(kinds == SOLID)))


def sphere_kinds(layout, centre, diameter) -> np.ndarray:
    """Cell kinds of an analytic sphere in a channel: SOLID where the cell centre lies strictly
    inside the sphere, the planes x = 0 and x = nx - 1 EQ, FLUID elsewhere."""
    nx, ny, nz = (int(v) for v in layout)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
    kinds = np.where(r2 < (diameter / 2) ** 2, SOLID, FLUID).astype(np.uint32)
    kinds[:, :, 0] = EQ
    kinds[:, :, nx - 1] = EQ
    return kinds.reshape(-1)
