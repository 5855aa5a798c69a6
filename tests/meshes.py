"""Test meshes restated from tests/unit/voxelize/fixtures.ts: the unit cube read from the
reference's binary STL fixture (tests/golden/cube-binary.stl, a copy of
tests/fixtures/stl/cube-binary.stl) scaled into a box, optionally with its +x face removed,
and the icosphere subdivided from the icosahedron."""
from __future__ import annotations

import math
import os
import struct

import numpy as np

_STL = os.path.join(os.path.dirname(__file__), "golden", "cube-binary.stl")


def _unit_cube() -> np.ndarray:
    data = open(_STL, "rb").read()
    (count,) = struct.unpack_from("<I", data, 80)
    rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", 9), ("a", "<u2")]),
                        count=count, offset=84)
    return rec["v"].reshape(-1).astype(np.float32)


def box(lo, hi, open_face=False) -> np.ndarray:
    unit = _unit_cube().astype(np.float64).reshape(-1, 3)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    pos = (lo + unit * (hi - lo)).astype(np.float32).reshape(-1)
    if not open_face:
        return pos
    hx = np.float32(hi[0])
    keep = [t for t in pos.reshape(-1, 9) if not (t[0] == hx and t[3] == hx and t[6] == hx)]
    return np.asarray(keep, np.float32).reshape(-1)


def _unit(p):
    n = math.sqrt(p[0] ** 2 + p[1] ** 2 + p[2] ** 2)
    return (p[0] / n, p[1] / n, p[2] / n)


def sphere(centre, radius, subdivisions=5) -> np.ndarray:
    phi = (1 + math.sqrt(5)) / 2
    verts = [_unit(v) for v in [(-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0), (0, -1, phi),
                                (0, 1, phi), (0, -1, -phi), (0, 1, -phi), (phi, 0, -1), (phi, 0, 1),
                                (-phi, 0, -1), (-phi, 0, 1)]]
    idx = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
           (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
           (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    faces = [(verts[a], verts[b], verts[c]) for a, b, c in idx]

    def mid(a, b):
        return _unit((a[0] + b[0], a[1] + b[1], a[2] + b[2]))

    for _ in range(subdivisions):
        nxt = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nxt
    out = np.array([[v[k] * radius + centre[k] for v in f for k in range(3)] for f in faces])
    return out.astype(np.float32).reshape(-1)
