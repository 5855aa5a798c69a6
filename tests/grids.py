"""Seeded test geometries restated from the reference's GPU parity specs.

blob_grid        tests/browser/gpu-vs-cpu.spec.ts:76-116 (LCG seed 0x4a3b2c1d, radius 5 + 2u)
link_fractions   tests/browser/gpu-vs-cpu.spec.ts:57-74  (seed 0x1b873593, every 7th q = 0.5)
grid_from_kinds  tests/browser/force.spec.ts:40-72 (pocketwind_amd.lattice, re-exported)
sphere_kinds     an analytic sphere in a channel with EQ inlet and outlet planes (no reference
                 counterpart; the free-slip / zero-gradient channel cases; pocketwind_amd.lattice)
"""
from __future__ import annotations

import math

import numpy as np

from pocketwind_amd.lattice import (CELL_TYPE_MASK, DIRECTIONS, EQ, FLOOR, FLUID, SOLID, SUPPORT,
                                    VoxelGrid, grid_from_kinds, sphere_kinds)  # noqa: F401 (re-exported)


def lcg(seed: int) -> int:
    return (seed * 1664525 + 1013904223) & 0xFFFFFFFF


def pull_wall_mask(kinds: np.ndarray, layout, cell: int) -> int:
    nx, ny, nz = layout
    x, y, z = cell % nx, (cell // nx) % ny, cell // (nx * ny)
    mask = 0
    for i, (cx, cy, cz) in enumerate(DIRECTIONS):
        src = ((x - cx) % nx) + nx * (((y - cy) % ny) + ny * ((z - cz) % nz))
        if SOLID <= (int(kinds[src]) & CELL_TYPE_MASK) <= FLOOR:
            mask |= 1 << (3 + i)
    return mask


def blob_kinds(layout=(32, 32, 32)) -> np.ndarray:
    nx, ny, nz = layout
    kinds = np.zeros(nx * ny * nz, dtype=np.uint32)
    seed = 0x4A3B2C1D
    for cell in range(kinds.size):
        x, y, z = cell % nx, (cell // nx) % ny, cell // (nx * ny)
        seed = lcg(seed)
        radius = 5 + (seed / 2**32) * 2
        if math.hypot(x - 15.5, y - 15.5, z - 15.5) <= radius:
            kinds[cell] = SOLID
        if x < 2 or x >= nx - 2:
            kinds[cell] = EQ
    kinds[15 + nx * (15 + ny * 15)] = SUPPORT
    kinds[16 + nx * (15 + ny * 15)] = FLOOR
    return kinds


def link_fractions(grid: VoxelGrid, initial_seed: int = 0x1B873593) -> VoxelGrid:
    cells = grid.cells
    geometry = grid.geometry.copy()
    seed = initial_seed
    fractions = []
    for cell in range(cells):
        flag = int(geometry[cell])
        if (flag & CELL_TYPE_MASK) != FLUID or flag >> 3 == 0:
            continue
        geometry[cells + cell] = len(fractions)
        for i in range(1, 19):
            if not flag & (1 << (3 + i)):
                continue
            seed = lcg(seed)
            fractions.append(0.5 if len(fractions) % 7 == 0 else 0.05 + 0.95 * (seed / 2**32))
    packed = np.concatenate([geometry, np.asarray(fractions, dtype=np.float32).view(np.uint32)])
    return VoxelGrid(grid.nx, grid.ny, grid.nz, packed, grid.boundary_cells.copy(),
                     solid_cells=grid.solid_cells)


def blob_grid(walls: str = "bb") -> VoxelGrid:
    layout = (32, 32, 32)
    grid = grid_from_kinds(layout, blob_kinds(layout))
    return link_fractions(grid) if walls == "bouzidi" else grid
