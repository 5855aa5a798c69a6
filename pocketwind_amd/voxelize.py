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


class _GridMemory:
    """The library's arrays of one pw_voxel_grid, freed once the last numpy view of them is gone
    (the geometry is 2N words and more; it is handed over without a copy)."""

    def __init__(self, out):
        self._out = out

    def __del__(self):
        try:
            N.lib().pw_voxel_grid_free(C.byref(self._out))
        except Exception:  # interpreter shutdown: the process frees it
            pass


def _view(pointer, count: int, owner: _GridMemory) -> np.ndarray:
    if count == 0:
        return np.zeros(0, np.uint32)
    owned = type("_OwnedWords", (C.c_uint32 * count,), {})
    words = owned.from_address(C.cast(pointer, C.c_void_p).value)
    words._owner = owner  # the numpy array keeps `words`, which keeps the owner
    return np.frombuffer(words, dtype=np.uint32)


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
    owner = _GridMemory(out)
    geo = _view(out.geometry, int(out.geometry_words), owner)
    bc = _view(out.boundary_cells, int(out.n_boundary), owner)
    return VoxelGrid(nx, ny, nz, geo, bc, int(out.solid_cells), int(out.frontal_cells),
                     float(out.blockage), bool(out.watertight), int(out.missed_links))
