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
