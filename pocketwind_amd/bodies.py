"""Triangle meshes of the tunnel bodies and their placement in lattice cells.

Presets restate src/geometry/presets.ts:4-189 (sphere 24 x 16 lat-long, Ahmed 25-degree body
lofted from rectangular stations, its four stilts, cube); placement restates
src/geometry/transform.ts:118-143 for flow along +x with z up. Coordinates are evaluated in
float64 and stored as float32, as the reference's Float32Array meshes are. ``bluff_body``
is the 1.0 x 0.37 x 0.28 slanted body of the 256x128x128 benchmark (front edges rounded at
0.1 L, 25-degree rear slant over the last 0.2 L), built with the same loft.

``noisy_bluff_body`` subdivides that body into a fine triangle mesh (217 088 triangles)
with seeded vertex noise, the procedural mesh of the full-pipeline workload.
"""
from __future__ import annotations

import math

import numpy as np


def _loft(stations):
    """Join rectangular stations (x, y_half, z_min, z_max) into a closed solid
    (presets.ts:48-90)."""
    def corners(s):
        x, yh, z0, z1 = s
        return [(x, -yh, z0), (x, yh, z0), (x, yh, z1), (x, -yh, z1)]

    tris = []

    def quad(a, b, c, d):
        tris.append((a, b, c))
        tris.append((a, c, d))

    rings = [corners(s) for s in stations]
    for near, far in zip(rings, rings[1:]):
        for k in range(4):
            n = (k + 1) % 4
            quad(near[k], far[k], far[n], near[n])
    first, last = rings[0], rings[-1]
    quad(first[0], first[1], first[2], first[3])
    quad(last[0], last[3], last[2], last[1])
    return np.asarray(tris, dtype=np.float64).astype(np.float32).reshape(-1)


def _rounded_nose(radius, half_width, bottom, top, segments=12):
    out = []
    for seg in range(segments + 1):
        angle = (seg / segments) * (math.pi / 2)
        inset = radius * (1 - math.sin(angle))
        out.append((radius * (1 - math.cos(angle)), half_width - inset, bottom + inset, top - inset))
    return out


def ahmed_mesh() -> np.ndarray:
    """presets.ts:92-115, millimetres."""
    nose = _rounded_nose(100, 389 / 2, 50, 50 + 288)
    slant = (25 * math.pi) / 180
    start = 1044 - 222 * math.cos(slant)
    drop = 222 * math.sin(slant)
    return _loft(nose + [(start, 389 / 2, 50, 338), (1044, 389 / 2, 50, 338 - drop)])


def stilts_mesh() -> np.ndarray:
    """presets.ts:118-141: four 30 mm stilts, 16-sided prisms from z = 0 to 50."""
    tris = []
    for x in (202, 672):
        for y in (-163.5, 163.5):
            r, h = 15.0, 50.0

            def rim(z, seg):
                theta = (seg / 16) * 2 * math.pi
                return (x + r * math.cos(theta), y + r * math.sin(theta), z)

            for seg in range(16):
                nxt = seg + 1
                a, b, c, d = rim(0, seg), rim(0, nxt), rim(h, nxt), rim(h, seg)
                tris += [(a, b, c), (a, c, d)]
                tris.append(((x, y, 0), rim(0, nxt), rim(0, seg)))
                tris.append(((x, y, h), rim(h, seg), rim(h, nxt)))
    return np.asarray(tris, dtype=np.float64).astype(np.float32).reshape(-1)


def cube_mesh() -> np.ndarray:
    return _loft([(0, 50, 0, 100), (100, 50, 0, 100)])


def sphere_mesh() -> np.ndarray:
    """presets.ts:152-177: radius 50, 24 segments, 16 rings; the poles have one vertex."""
    def vertex(ring, seg):
        phi = (ring / 16) * math.pi
        theta = (seg / 24) * 2 * math.pi
        return (50 * math.sin(phi) * math.cos(theta), 50 * math.sin(phi) * math.sin(theta),
                50 * math.cos(phi))

    tris = []
    for ring in range(16):
        for seg in range(24):
            nxt = (seg + 1) % 24
            top, bottom = vertex(ring, seg), vertex(ring + 1, seg)
            bottom_next, top_next = vertex(ring + 1, nxt), vertex(ring, nxt)
            if ring != 0:
                tris.append((top, bottom, top_next))
            if ring != 15:
                tris.append((top_next, bottom, bottom_next))
    return np.asarray(tris, dtype=np.float64).astype(np.float32).reshape(-1)


PRESETS = {"sphere": sphere_mesh, "ahmed-25": ahmed_mesh, "cube": cube_mesh}


def preset_mesh(name: str):
    """presetMesh (presets.ts:179-185); None for an unknown preset."""
    maker = PRESETS.get(name)
    return None if maker is None else maker()


def preset_supports(name: str):
    return stilts_mesh() if name == "ahmed-25" else None


def bluff_body(length=1.0, width=0.37, height=0.28, radius=0.1, slant_length=0.2,
               slant_deg=25.0, segments=12) -> np.ndarray:
    """The 256x128x128 benchmark body: a box whose four front edges are rounded at `radius`
    (every nose station is the full section inset equally) and whose roof falls at
    `slant_deg` over the last `slant_length` of its length."""
    nose = _rounded_nose(radius, width / 2, 0.0, height, segments)
    drop = slant_length * math.tan(math.radians(slant_deg))
    return _loft(nose + [(length - slant_length, width / 2, 0.0, height),
                         (length, width / 2, 0.0, height - drop)])


def subdivide(mesh: np.ndarray, levels: int) -> np.ndarray:
    """Split every triangle into four at its edge midpoints, `levels` times."""
    t = np.asarray(mesh, np.float64).reshape(-1, 3, 3)
    for _ in range(levels):
        a, b, c = t[:, 0], t[:, 1], t[:, 2]
        ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
        t = np.stack([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                      np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
    return t.astype(np.float32).reshape(-1)


def noisy_bluff_body(levels: int = 5, segments: int = 24, amplitude: float = 0.002,
                     seed: int = 1234) -> np.ndarray:
    """Fine slanted bluff body with seeded vertex noise, watertight: every distinct vertex
    moves once, so shared vertices stay shared. A loft of s nose segments has 8 (s + 2) + 4
    triangles; the default (24 segments, 5 subdivisions) gives 212 * 4^5 = 217 088."""
    t = subdivide(bluff_body(segments=segments), levels).reshape(-1, 3)
    keys, inverse = np.unique(t, axis=0, return_inverse=True)
    rng = np.random.default_rng(seed)
    moved = keys.astype(np.float64) + rng.uniform(-amplitude, amplitude, keys.shape)
    return moved[inverse.reshape(-1)].astype(np.float32).reshape(-1)


def mesh_bounds(mesh: np.ndarray):
    v = np.asarray(mesh, np.float32).reshape(-1, 3).astype(np.float64)
    return v.min(0), v.max(0)


def place_mesh(mesh, supports, length_cells: float, layout, floor: bool):
    """placeMesh (transform.ts:118-143), flow +x, up z: nose at 0.25 nx, centred in y, centred
    in z or standing on the floor plane at z = 1/2; scaled so the longest extent of the body
    is `length_cells`. Returns (body, supports) as float32 arrays."""
    nx, ny, nz = layout
    if min(layout) < 2:
        raise ValueError("Grid dimensions must be integers of at least two cells.")
    if not (length_cells > 0 and math.isfinite(length_cells)):
        raise ValueError("Body length in cells must be finite and positive.")
    lo, hi = mesh_bounds(mesh)
    if not np.all(np.isfinite(lo)) or not np.all(np.isfinite(hi)):
        raise ValueError("Mesh coordinates must be finite.")
    lowest = lo[2]
    if supports is not None:
        lowest = min(lowest, mesh_bounds(supports)[0][2])
    longest = max(hi - lo)
    if longest == 0:
        raise ValueError("Mesh must have a nonzero extent.")
    scale = length_cells / longest
    nose, cy, cz = lo[0], (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2

    def transform(m):
        v = np.asarray(m, np.float32).reshape(-1, 3).astype(np.float64)
        out = np.empty_like(v)
        out[:, 0] = 0.25 * nx + 1 * (v[:, 0] - nose) * scale
        out[:, 1] = ny / 2 + 1 * (v[:, 1] - cy) * scale
        out[:, 2] = (0.5 + (v[:, 2] - lowest) * scale) if floor else (nz / 2 + (v[:, 2] - cz) * scale)
        return out.astype(np.float32).reshape(-1)

    return transform(mesh), (None if supports is None else transform(supports))
