"""ZIP and 3MF archives for the mesh-import tests, built byte by byte so that flags, methods and
sizes can be set as the reference's fixture generator sets them (scripts/make-3mf-fixtures.ts:
version 20, no timestamps or attributes, raw deflate by default), plus the cube model and the
subdivided-icosahedron ellipsoid the reference's 3MF tests use."""
from __future__ import annotations

import math
import struct
import zlib

MODEL_PATH = "3D/3dmodel.model"
CORE_NS = "http://schemas.microsoft.com/3dmanufacturing/core/2015/02"
PRODUCTION_NS = "http://schemas.microsoft.com/3dmanufacturing/production/2015/06"
RELS = ('<Relationships xmlns="http://schemas.openxmlformats.org/package/2006/relationships">'
        '<Relationship Id="start" Type="http://schemas.microsoft.com/3dmanufacturing/2013/01/3dmodel" '
        'Target="/3D/3dmodel.model"/></Relationships>')
CONTENT_TYPES = ('<Types xmlns="http://schemas.openxmlformats.org/package/2006/content-types">'
                 '<Default Extension="rels" ContentType="application/vnd.openxmlformats-package.relationships+xml"/>'
                 '<Default Extension="model" ContentType="application/vnd.ms-package.3dmanufacturing-3dmodel+xml"/>'
                 '</Types>')

CUBE_VERTICES = [(0, 0, 0), (300, 0, 0), (300, 300, 0), (0, 300, 0),
                 (0, 0, 300), (300, 0, 300), (300, 300, 300), (0, 300, 300)]
CUBE_FACES = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
              (1, 2, 6), (1, 6, 5), (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7)]


def raw_deflate(data: bytes) -> bytes:
    c = zlib.compressobj(9, zlib.DEFLATED, -15)
    return c.compress(data) + c.flush()


def make_archive(parts) -> bytes:
    """parts: dicts with name, text, optional method (8) and flags (0)."""
    local_blobs, central_blobs = [], []
    offset = 0
    for p in parts:
        name = p["name"].encode()
        raw = p["text"].encode()
        method = p.get("method", 8)
        flags = p.get("flags", 0)
        body = raw_deflate(raw) if method == 8 else raw
        crc = zlib.crc32(raw) & 0xFFFFFFFF
        local = struct.pack("<IHHHHHIIIHH", 0x04034B50, 20, flags, method, 0, 0, crc, len(body), len(raw),
                            len(name), 0)
        central = struct.pack("<IHHHHHHIIIHHHHHII", 0x02014B50, 20, 20, flags, method, 0, 0, crc, len(body),
                              len(raw), len(name), 0, 0, 0, 0, 0, offset)
        local_blobs += [local, name, body]
        central_blobs += [central, name]
        offset += len(local) + len(name) + len(body)
    directory = b"".join(central_blobs)
    end = struct.pack("<IHHHHIIH", 0x06054B50, 0, 0, len(parts), len(parts), len(directory), offset, 0)
    return b"".join(local_blobs) + directory + end


def mesh_xml(vertices, faces) -> str:
    v = "".join(f'<vertex x="{x}" y="{y}" z="{z}"/>' for x, y, z in vertices)
    t = "".join(f'<triangle v1="{a}" v2="{b}" v3="{c}"/>' for a, b, c in faces)
    return f"<mesh><vertices>{v}</vertices><triangles>{t}</triangles></mesh>"


CUBE_OBJECT = f'<object id="1" type="model">{mesh_xml(CUBE_VERTICES, CUBE_FACES)}</object>'


def model_xml(objects: str = CUBE_OBJECT, build: str = '<item objectid="1"/>', unit: str = "millimeter") -> str:
    return (f'<?xml version="1.0" encoding="UTF-8"?><model xmlns="{CORE_NS}" xmlns:p="{PRODUCTION_NS}" '
            f'unit="{unit}"><resources>{objects}</resources><build>{build}</build></model>')


def model_archive(text: str, method: int = 8) -> bytes:
    return make_archive([{"name": "[Content_Types].xml", "text": CONTENT_TYPES},
                         {"name": "_rels/.rels", "text": RELS},
                         {"name": MODEL_PATH, "text": text, "method": method}])


def _js_number(x: float) -> str:
    """Plain-notation shortest text of a coordinate (the generator interpolates numbers)."""
    r = repr(float(x))
    return r[:-2] if r.endswith(".0") else r


def ellipsoid_archive(levels: int = 7, radii=(450.0, 150.0, 150.0)) -> bytes:
    """A closed ellipsoid: the icosahedron refined `levels` times with shared midpoints projected
    to the unit sphere, then scaled by `radii` (20 * 4^7 = 327680 triangles at 7 levels)."""
    phi = (1 + math.sqrt(5)) / 2

    def unit(v):
        n = math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
        return (v[0] / n, v[1] / n, v[2] / n)

    vertices = [unit(v) for v in [(-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0), (0, -1, phi),
                                  (0, 1, phi), (0, -1, -phi), (0, 1, -phi), (phi, 0, -1), (phi, 0, 1),
                                  (-phi, 0, -1), (-phi, 0, 1)]]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(levels):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                va, vb = vertices[a], vertices[b]
                vertices.append(unit(((va[0] + vb[0]) / 2, (va[1] + vb[1]) / 2, (va[2] + vb[2]) / 2)))
                cache[key] = len(vertices) - 1
            return cache[key]

        refined = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            refined += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = refined
    scaled = [(_js_number(x * radii[0]), _js_number(y * radii[1]), _js_number(z * radii[2])) for x, y, z in vertices]
    obj = f'<object id="1" type="model">{mesh_xml(scaled, faces)}</object>'
    return model_archive(model_xml(obj))
