"""voxelize(): the reference's voxelizer API over libpocketwind's GPU voxelizer.

Mirrors src/geometry/voxelize.ts:172-199: ``voxelize(body, supports, layout, floor,
link_fractions)`` returns a :class:`VoxelGrid` with the packed geometry
[flags N][q offsets N][f32 q per link], the body boundary-cell list and the counts.
Invalid input raises ``ValueError`` (the reference's RangeError).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .lattice import VoxelGrid


def _mesh(m, name):
    if m is None:
        return None
    arr = np.ascontiguousarray(m, dtype=np.float32).reshape(-1)
    if arr.size % 9:
        raise ValueError(f"{name} positions must match a non-negative triangle count.")
    return arr


def voxelize(body, supports, layout, floor: bool, link_fractions: bool = True) -> VoxelGrid:
    nx, ny, nz = (int(v) for v in layout)
    b = _mesh(body, "Mesh")
    s = _mesh(supports, "Support mesh")
    out = N.VoxelGridC()
    lay = N.Layout(nx, ny, nz)
    status = N.lib().pw_voxelize(
        b.ctypes.data if b.size else None, b.size // 9,
        None if s is None else (s.ctypes.data if s.size else None), 0 if s is None else s.size // 9,
        C.byref(lay), int(bool(floor)), int(bool(link_fractions)), C.byref(out))
    N.check(status)
    try:
        geo = np.ctypeslib.as_array(out.geometry, shape=(out.geometry_words,)).copy()
        bc = (np.ctypeslib.as_array(out.boundary_cells, shape=(out.n_boundary,)).copy()
              if out.n_boundary else np.zeros(0, np.uint32))
        return VoxelGrid(nx, ny, nz, geo, bc, int(out.solid_cells), int(out.frontal_cells),
                         float(out.blockage), bool(out.watertight), int(out.missed_links))
    finally:
        N.lib().pw_voxel_grid_free(C.byref(out))
