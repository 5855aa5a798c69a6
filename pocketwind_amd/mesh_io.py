"""Mesh import: binary and ASCII STL, OBJ and 3MF into float32 triangle positions that feed the
voxelizer (bodies.place_mesh, then voxelize).

parse_stl        parseStl (src/geometry/stl.ts:118-150)
parse_obj        parseObj (src/geometry/obj.ts:65-88)
crc32, read_zip, inflate_zip_part
                 src/geometry/zip.ts:27-191 (single-disk, stored or raw-deflate members)
parse_three_mf   parseThreeMf (src/geometry/threemf.ts:246-350): build items and component
                 instances merged in millimetres
parse_model      parseModel (src/geometry/model.ts:28-52): format detection by content

Failures are returned as ModelParseError values with the reference's kinds and byte offsets
(the reference returns error objects rather than throwing); an invalid triangle limit raises
ValueError, as the reference raises RangeError.
"""
from __future__ import annotations

import re
import struct
import zlib
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Union

import numpy as np

DEFAULT_MAX_TRIANGLES = 2_000_000
MAX_SAFE_INTEGER = 2 ** 53 - 1

# ECMAScript WhiteSpace and LineTerminator code points: the \s, \S and trim() of the formats'
# grammars. Python's own \s differs (it adds U+001C-U+001F and U+0085, and lacks U+FEFF).
_WS_CHARS = "\t\n\v\f\r \u00a0\u1680\u2000-\u200a\u2028\u2029\u202f\u205f\u3000\ufeff"
_WS = f"[{_WS_CHARS}]"
_NOT_WS = f"[^{_WS_CHARS}]"
_WS_STRIP = "\t\n\v\f\r \u00a0\u1680" + "".join(chr(c) for c in range(0x2000, 0x200B)) + \
            "\u2028\u2029\u202f\u205f\u3000\ufeff"
_DECIMAL = re.compile(r"[+-]?(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:[eE][+-]?[0-9]+)?\Z")


def _js_trim(text: str) -> str:
    return text.strip(_WS_STRIP)


def _fround_finite(value: float) -> bool:
    with np.errstate(over="ignore", invalid="ignore"):
        return bool(np.isfinite(np.float32(value)))


def _check_limit(max_triangles) -> int:
    if isinstance(max_triangles, bool) or not isinstance(max_triangles, (int, np.integer)) \
            or not 0 <= max_triangles <= MAX_SAFE_INTEGER:
        raise ValueError("maxTriangles must be a non-negative safe integer.")
    return int(max_triangles)


@dataclass(frozen=True)
class Mesh:
    """Mesh (stl.ts:1): 9 float32 coordinates per triangle."""

    positions: np.ndarray
    triangle_count: int


@dataclass(frozen=True)
class ModelParseError:
    """StlParseError / ModelParseError (stl.ts:3-7, model.ts:5-8)."""

    kind: str
    byte_offset: int
    message: str
    part: Optional[str] = None


@dataclass(frozen=True)
class ParsedModel:
    """ParsedModel (model.ts:10-15)."""

    mesh: Mesh
    format: str
    unit_mm: Optional[float]
    item_count: int


# --------------------------------------------------------------------------------------------
# STL

_STL_MESSAGES = {
    "truncated": "The STL ends before the mesh is complete.",
    "not-stl": "The file does not contain a valid STL mesh.",
    "too-many-triangles": "The STL exceeds the triangle limit.",
    "empty": "The STL contains no triangles.",
    "non-finite": "A vertex cannot be represented as a finite float32 coordinate.",
}
_STL_RECORD = np.dtype([("normal", "<f4", 3), ("vertices", "<f4", 9), ("attribute", "<u2")])
# The tokens of one ASCII facet after its `facet` keyword: None is a coordinate, "" a token
# that is skipped (the three normal components).
_FACET = ("normal", "", "", "", "outer", "loop") + ("vertex", None, None, None) * 3 + ("endloop", "endfacet")
_NON_NUMBER = re.compile(r"[+-]?(?:nan|infinity)\Z", re.IGNORECASE)
_SOLID_PREFIX = re.compile(f"solid(?:{_WS}|\\Z)")
_SOLID_LINE = re.compile(r"solid[^\r\n]*(?:\r\n|[\r\n]|\Z)")
_TOKEN = re.compile(f"{_NOT_WS}+")


def _stl_error(kind: str, offset: int) -> ModelParseError:
    return ModelParseError(kind, offset, _STL_MESSAGES[kind])


# WHATWG windows-1252: one character per byte, so token positions are byte offsets.
_CP1252_HIGH = {b: bytes([b]).decode("cp1252", errors="ignore") or chr(b) for b in range(0x80, 0xA0)}
_CP1252_TABLE = {b: c for b, c in _CP1252_HIGH.items() if c != chr(b)}


def _windows_1252(data: bytes) -> str:
    return data.decode("latin-1").translate(_CP1252_TABLE)


def _binary_stl(data: bytes, count: int) -> Union[Mesh, ModelParseError]:
    if count == 0:
        return _stl_error("empty", 80)
    records = np.frombuffer(data, dtype=_STL_RECORD, count=count, offset=84)
    positions = np.ascontiguousarray(records["vertices"]).reshape(-1).astype(np.float32)
    bad = ~np.isfinite(positions)
    if bad.any():
        k = int(np.argmax(bad))
        return _stl_error("non-finite", 84 + 50 * (k // 9) + 12 + 4 * (k % 9))
    return Mesh(positions, count)


def _ascii_stl(text: str, max_triangles: int) -> Union[Mesh, ModelParseError]:
    header = _SOLID_LINE.match(text)
    if not header:
        return _stl_error("not-stl", 0)
    tokens = _TOKEN.finditer(text, header.end())
    values: List[float] = []
    triangles = 0
    for facet in tokens:
        word = facet.group()
        if word == "endsolid":
            return _finish_ascii(text, facet.end(), facet.start(), values, triangles)
        if word != "facet":
            return _stl_error("not-stl", facet.start())
        if triangles >= max_triangles:
            return _stl_error("too-many-triangles", facet.start())
        for expected in _FACET:
            token = next(tokens, None)
            if token is None:
                return _stl_error("truncated", len(text))
            t = token.group()
            if expected is None:
                if _NON_NUMBER.match(t):  # NaN or an infinity (Number() of the other spellings is NaN)
                    return _stl_error("non-finite", token.start())
                if not _DECIMAL.match(t):
                    return _stl_error("not-stl", token.start())
                value = float(t)
                if not _fround_finite(value):
                    return _stl_error("non-finite", token.start())
                values.append(value)
            elif expected and t != expected:
                return _stl_error("not-stl", token.start())
        triangles += 1
    return _stl_error("truncated", len(text))


def _finish_ascii(text: str, name_end: int, endsolid_at: int, values, triangles) -> Union[Mesh, ModelParseError]:
    # The solid's name may follow `endsolid` on its line; nothing but space after that line.
    line_end = re.compile(r"[\r\n]").search(text, name_end)
    if line_end:
        extra = _TOKEN.search(text, line_end.start())
        if extra:
            return _stl_error("not-stl", extra.start())
    if triangles == 0:
        return _stl_error("empty", endsolid_at)
    return Mesh(np.array(values, dtype=np.float64).astype(np.float32), triangles)


def parse_stl(data: bytes, max_triangles: int = DEFAULT_MAX_TRIANGLES) -> Union[Mesh, ModelParseError]:
    """parseStl (stl.ts:118-150): an exact binary record layout wins over ASCII."""
    limit = _check_limit(max_triangles)
    data = bytes(data)
    has_count = len(data) >= 84
    count = struct.unpack_from("<I", data, 80)[0] if has_count else 0
    expected = 84 + 50 * count
    if has_count and expected == len(data):
        if count > limit:
            return _stl_error("too-many-triangles", 80)
        return _binary_stl(data, count)
    if _SOLID_PREFIX.match(_windows_1252(data[:6])):
        return _ascii_stl(_windows_1252(data), limit)
    if has_count and expected > len(data):
        if count > limit:
            return _stl_error("too-many-triangles", 80)
        return _stl_error("truncated", len(data))
    return _stl_error("not-stl", 0)


# --------------------------------------------------------------------------------------------
# OBJ

_OBJ_LINE = re.compile(r"[^\r\n]*(?:\r\n|[\r\n]|\Z)")
_OBJ_REF = re.compile(r"[+-]?[0-9]+(?:/[+-]?[0-9]+(?:/[+-]?[0-9]+)?|//[+-]?[0-9]+)?\Z")
_SPLIT_WS = re.compile(f"{_WS}+")


def _obj_lines(text: str):
    """The matches of String.matchAll(/[^\\r\\n]*(?:\\r\\n|[\\r\\n]|$)/g): each line with its
    terminator, then the empty match at the end of the text."""
    pos = 0
    while pos <= len(text):
        m = _OBJ_LINE.match(text, pos)
        yield m.group()
        pos = m.end() if m.end() > pos else pos + 1


def _obj_error(kind: str, offset: int, message: str) -> ModelParseError:
    return ModelParseError(kind, offset, message)


def parse_obj(data: bytes, max_triangles: int = DEFAULT_MAX_TRIANGLES) -> Union[Mesh, ModelParseError]:
    """parseObj (obj.ts:65-88): `v` and `f` records; a polygon becomes a fan from its first
    vertex; indices resolve when the face is declared (negative ones count back)."""
    limit = _check_limit(max_triangles)
    text = bytes(data).decode("utf-8", errors="replace")
    vertices: List[List[float]] = []
    positions: List[float] = []
    offset = 0
    for line in _obj_lines(text):
        fields = _SPLIT_WS.split(_js_trim(line.split("#", 1)[0]))
        tag, fields = fields[0], fields[1:]
        if tag == "v":
            coords = fields[:3]
            if len(coords) != 3 or not all(_DECIMAL.match(c) for c in coords):
                return _obj_error("not-stl", offset, "An OBJ vertex needs three numeric coordinates.")
            vertex = [float(c) for c in coords]
            if not all(_fround_finite(v) for v in vertex):
                return _obj_error("non-finite", offset, "An OBJ vertex exceeds the finite float32 range.")
            vertices.append(vertex)
        elif tag == "f":
            if len(fields) < 3:
                return _obj_error("not-stl", offset, "An OBJ face needs at least three vertices.")
            face: List[float] = []
            for ref in fields:
                if not _OBJ_REF.match(ref):
                    return _obj_error("not-stl", offset, "An OBJ face contains an invalid vertex reference.")
                index = int(ref.split("/")[0])
                k = index - 1 if index > 0 else len(vertices) + index
                if index == 0 or abs(index) > MAX_SAFE_INTEGER or not 0 <= k < len(vertices):
                    return _obj_error("not-stl", offset, "An OBJ face references a missing vertex.")
                face.extend(vertices[k])
            if len(positions) // 9 + len(face) // 3 - 2 > limit:
                return _obj_error("too-many-triangles", offset, "The OBJ exceeds the triangle limit.")
            for j in range(3, len(face) - 3, 3):
                positions.extend(face[0:3])
                positions.extend(face[j:j + 6])
        offset += len(line.encode("utf-8"))
    if not positions:
        return _obj_error("empty", 0, "The OBJ contains no faces.")
    return Mesh(np.array(positions, dtype=np.float64).astype(np.float32), len(positions) // 9)


# --------------------------------------------------------------------------------------------
# ZIP

EOCD_BYTES = 22
MAX_COMMENT_BYTES = 65_535
MAX_EXPANDED_BYTES = 512 * 1024 * 1024


@dataclass
class ZipPart:
    """ZipPart (zip.ts:3-10): one member, located by its central-directory entry."""

    name: str
    method: int
    size: int
    crc: int
    byte_offset: int
    compressed: bytes


def crc32(data: bytes) -> int:
    """ZIP's IEEE CRC-32 (reflected polynomial 0xEDB88320), as zip.ts:27-38."""
    return zlib.crc32(bytes(data)) & 0xFFFFFFFF


def _zip_error(kind: str, offset: int, message: str, part: Optional[str] = None) -> ModelParseError:
    return ModelParseError(kind, offset, message, part)


def _has_zip64_field(data: bytes, start: int, end: int) -> bool:
    """An extra-field run that holds a ZIP64 record (id 1) or does not parse."""
    o = start
    while o < end:
        if o + 4 > end:
            return True
        field_id, length = struct.unpack_from("<HH", data, o)
        if field_id == 1 or o + 4 + length > end:
            return True
        o += 4 + length
    return False


def read_zip(data: bytes) -> Union[Dict[str, ZipPart], ModelParseError]:
    """readZip (zip.ts:51-149): index the members of a single-disk archive from its central
    directory, rejecting ZIP64, encryption, methods other than stored/deflate, inconsistent
    local headers and more than 512 MiB of declared expansion."""
    data = bytes(data)
    n = len(data)
    u16 = lambda o: struct.unpack_from("<H", data, o)[0]
    u32 = lambda o: struct.unpack_from("<I", data, o)[0]
    end = n - EOCD_BYTES
    lowest = max(0, end - MAX_COMMENT_BYTES)
    while end >= lowest:
        if u32(end) == 0x06054B50 and end + EOCD_BYTES + u16(end + 20) == n:
            break
        end -= 1
    if end < lowest:
        return _zip_error("truncated", n, "The ZIP directory is missing.")
    count = u16(end + 10)
    directory_size = u32(end + 12)
    offset = u32(end + 16)
    if (count == 0xFFFF or directory_size == 0xFFFFFFFF or offset == 0xFFFFFFFF
            or (end >= 20 and u32(end - 20) == 0x07064B50)
            or u16(end + 4) != 0 or u16(end + 6) != 0 or u16(end + 8) != count):
        return _zip_error("zip-unsupported", end, "ZIP64 and multi-disk archives are unsupported.")
    if offset + directory_size != end:
        return _zip_error("truncated", offset, "The ZIP directory has an invalid size or offset.")
    parts: Dict[str, ZipPart] = {}
    expanded = 0
    for _ in range(count):
        if offset + 46 > end or u32(offset) != 0x02014B50:
            return _zip_error("truncated", offset, "A ZIP directory entry is incomplete.")
        flags, method = u16(offset + 8), u16(offset + 10)
        compressed_size, size = u32(offset + 20), u32(offset + 24)
        name_len, extra_len, comment_len = u16(offset + 28), u16(offset + 30), u16(offset + 32)
        following = offset + 46 + name_len + extra_len + comment_len
        local = u32(offset + 42)
        if following > end:
            return _zip_error("truncated", offset, "A ZIP entry extends beyond its directory.")
        if (flags & 0x41 or method not in (0, 8) or size == 0xFFFFFFFF or compressed_size == 0xFFFFFFFF
                or local == 0xFFFFFFFF or u16(offset + 34) != 0
                or _has_zip64_field(data, offset + 46 + name_len, offset + 46 + name_len + extra_len)):
            return _zip_error("zip-unsupported", offset, "Use a ZIP without encryption or ZIP64, stored or deflated.")
        if local + 30 > offset or u32(local) != 0x04034B50:
            return _zip_error("truncated", local, "A ZIP local header is missing.")
        if u16(local + 6) & 0x41 or u16(local + 8) != method:
            return _zip_error("zip-unsupported", local, "ZIP local compression flags disagree with the directory.")
        local_name_len, local_extra_len = u16(local + 26), u16(local + 28)
        start = local + 30 + local_name_len + local_extra_len
        if start + compressed_size > offset:
            return _zip_error("truncated", start, "A ZIP member is incomplete.")
        if _has_zip64_field(data, start - local_extra_len, start):
            return _zip_error("zip-unsupported", local, "ZIP64 members are unsupported.")
        name = data[offset + 46:offset + 46 + name_len].decode("utf-8", errors="replace")
        local_name = data[local + 30:local + 30 + local_name_len].decode("utf-8", errors="replace")
        if name != local_name or name in parts:
            return _zip_error("zip-unsupported", local, "ZIP member names disagree or are duplicated.")
        expanded += size
        if expanded > MAX_EXPANDED_BYTES:
            return _zip_error("zip-unsupported", offset, "The expanded ZIP exceeds 512 MiB.")
        parts[name] = ZipPart(name, method, size, u32(offset + 16), start, data[start:start + compressed_size])
        offset = following
    if offset != end:
        return _zip_error("truncated", offset, "The ZIP directory count is incorrect.")
    return parts


def inflate_zip_part(part: ZipPart) -> Union[bytes, ModelParseError]:
    """inflateZipPart (zip.ts:152-191): expand a member and check its size and CRC."""
    fail = lambda message: _zip_error("truncated", part.byte_offset, message, part.name)
    out = part.compressed
    if part.method == 8:
        inflater = zlib.decompressobj(-15)
        try:
            out = inflater.decompress(part.compressed, part.size + 1)
        except zlib.error:
            return fail("The ZIP member contains invalid raw-deflate data.")
        if len(out) > part.size:
            return fail("The ZIP member expands beyond its declared size.")
        if not inflater.eof or inflater.unused_data:
            return fail("The ZIP member contains invalid raw-deflate data.")
        if len(out) != part.size:
            return fail("The ZIP member expands to an incorrect size.")
    if len(out) != part.size or crc32(out) != part.crc:
        return fail("The ZIP member size or checksum is incorrect.")
    return out


# --------------------------------------------------------------------------------------------
# 3MF

class _BadXml(Exception):
    pass


class _TooManyTriangles(Exception):
    pass


UNITS_MM = {"micron": 0.001, "millimeter": 1.0, "centimeter": 10.0, "inch": 25.4, "foot": 304.8, "meter": 1000.0}
IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)
MODEL_PATH = "3D/3dmodel.model"

_NAME = r"[A-Za-z0-9_.:-]"
_XML_TOKEN = re.compile(
    r"<!--[\s\S]*?-->|<\?[\s\S]*?\?>|<!\[CDATA\[[\s\S]*?\]\]>"
    r"|</?[A-Za-z_]" + _NAME + r"*(?:[^<>\"']|\"[^\"]*\"|'[^']*')*>")
_XML_NAME = re.compile(r"</?(" + _NAME + r"+)")
_XML_CLOSE = re.compile(r"</" + _NAME + f"+{_WS}*>\\Z")
_XML_ATTR = re.compile(f"{_WS}+([A-Za-z_]{_NAME}*){_WS}*={_WS}*(?:\"([^\"<]*)\"|'([^'<]*)')")
_XML_TAG_END = re.compile(f"{_WS}*/?>\\Z")
_XML_ENTITY = re.compile(r"&([^;]*);|&")
_XML_NAMED = {"amp": "&", "lt": "<", "gt": ">", "quot": '"', "apos": "'"}
_XML_NUMERIC = re.compile(r"#(?:[0-9]+|x[0-9a-fA-F]+)\Z")


def _unescape(value: str) -> str:
    def entity(m):
        name = m.group(1)
        if name in _XML_NAMED:
            return _XML_NAMED[name]
        if name and _XML_NUMERIC.match(name):
            code = int(name[2:], 16) if name.startswith("#x") else int(name[1:])
            if 0 < code <= 0x10FFFF and not 0xD800 <= code <= 0xDFFF:
                return chr(code)
        raise _BadXml("The XML contains an unknown or invalid entity.")
    return _XML_ENTITY.sub(entity, value)


def _attributes(token: str, name: str) -> Dict[str, str]:
    attrs: Dict[str, str] = {}
    cursor = len(name) + 1
    while True:
        m = _XML_ATTR.match(token, cursor)
        if not m:
            break
        key = m.group(1)
        if key in attrs:
            raise _BadXml("The XML has duplicate attributes.")
        value = m.group(2) if m.group(2) is not None else (m.group(3) or "")
        attrs[key] = _unescape(value) if "&" in value else value
        cursor = m.end()
    if not _XML_TAG_END.match(token, cursor):
        raise _BadXml("The XML attributes must be quoted.")
    return attrs


def scan_xml(text: str, visit: Callable[[str, Dict[str, str]], None]) -> None:
    """The scanner of threemf.ts:58-93: one root element, balanced tags, quoted attributes,
    comments, processing instructions and CDATA skipped, anything else (a DOCTYPE, stray '<')
    rejected. `visit` receives each start tag's path of local names and its attributes."""
    stack: List[str] = []
    paths: List[str] = []
    offset = 0
    roots = 0
    for m in _XML_TOKEN.finditer(text):
        gap = text[offset:m.start()]
        if "<" in gap or (not stack and _js_trim(gap)):
            raise _BadXml("The XML contains malformed text.")
        token = m.group()
        offset = m.end()
        if token.startswith("<?") or token.startswith("<!"):
            continue
        name = _XML_NAME.match(token).group(1)
        if token.startswith("</"):
            if not _XML_CLOSE.match(token):
                raise _BadXml("The XML closing tag is malformed.")
            if not stack or stack.pop() != name:
                raise _BadXml("The XML elements are not balanced.")
            paths.pop()
            continue
        if not stack:
            roots += 1
        local = name[name.rfind(":") + 1:]
        path = f"{paths[-1]}/{local}" if paths else local
        visit(path, _attributes(token, name))
        if not token.endswith("/>"):
            stack.append(name)
            paths.append(path)
    if stack or roots != 1 or _js_trim(text[offset:]):
        raise _BadXml("The XML document is incomplete.")


def _number(value: Optional[str]) -> float:
    if value is None or not _DECIMAL.match(value) or not np.isfinite(float(value)):
        raise _BadXml("A model coordinate or transform is not a finite number.")
    return float(value)


@dataclass
class _Reference:
    object_id: str
    path: str
    transform: tuple


def _reference(attrs: Dict[str, str]) -> _Reference:
    object_id = attrs.get("objectid")
    if not object_id:
        raise _BadXml("A component or build item is missing its object ID.")
    if "transform" in attrs:
        transform = tuple(_number(v) for v in _SPLIT_WS.split(_js_trim(attrs["transform"])))
    else:
        transform = IDENTITY
    if len(transform) != 12:
        raise _BadXml("A 3MF transform must contain twelve numbers.")
    path = next((v for k, v in attrs.items() if k.endswith(":path")), "")
    return _Reference(object_id, path, transform)


def _triangle_index(value: Optional[str]) -> int:
    if value is None or not re.fullmatch(r"[0-9]+", value) or int(value) > MAX_SAFE_INTEGER:
        raise _BadXml("A triangle index is invalid.")
    return int(value)


@dataclass
class _Object:
    vertices: List[float]
    indices: List[int]
    components: List[_Reference]


@dataclass
class _ModelPart:
    objects: Dict[str, _Object]
    items: List[_Reference]
    unit_mm: float


def _parse_part(text: str, max_triangles: int) -> _ModelPart:
    part = _ModelPart({}, [], 1.0)
    state = {"current": None, "model": False}

    def visit(path: str, attrs: Dict[str, str]) -> None:
        current = state["current"]
        if path == "model":
            state["model"] = True
            unit = attrs.get("unit", "millimeter")
            if unit not in UNITS_MM:
                raise _BadXml("The 3MF unit is unsupported.")
            part.unit_mm = UNITS_MM[unit]
        elif path == "model/resources/object":
            oid = attrs.get("id")
            if not oid or oid in part.objects:
                raise _BadXml("An object ID is missing or duplicated.")
            state["current"] = part.objects[oid] = _Object([], [], [])
        elif path == "model/resources/object/mesh/vertices/vertex":
            if current is not None:
                current.vertices.extend((_number(attrs.get("x")), _number(attrs.get("y")), _number(attrs.get("z"))))
        elif path == "model/resources/object/mesh/triangles/triangle":
            if current is not None:
                if len(current.indices) // 3 >= max_triangles:
                    raise _TooManyTriangles("The 3MF exceeds the triangle limit.")
                current.indices.extend((_triangle_index(attrs.get("v1")), _triangle_index(attrs.get("v2")),
                                        _triangle_index(attrs.get("v3"))))
        elif path == "model/resources/object/components/component":
            if current is not None:
                current.components.append(_reference(attrs))
        elif path == "model/build/item":
            part.items.append(_reference(attrs))

    scan_xml(text, visit)
    if not state["model"]:
        raise _BadXml("The model part has no model root.")
    for obj in part.objects.values():
        if obj.components and obj.indices:
            raise _BadXml("An object cannot contain both a mesh and components.")
        if obj.indices and 3 * max(obj.indices) >= len(obj.vertices):
            raise _BadXml("A triangle references a missing vertex.")
    return part


_PERCENT_RUN = re.compile(r"(?:%[0-9A-Fa-f]{2})+")


def _decode_uri_component(value: str) -> str:
    """decodeURIComponent: runs of %XX escapes decode as strict UTF-8; a lone '%' is an error."""
    out = []
    pos = 0
    for m in _PERCENT_RUN.finditer(value):
        if "%" in value[pos:m.start()]:
            raise _BadXml("URI malformed")
        out.append(value[pos:m.start()])
        raw = bytes(int(m.group()[i + 1:i + 3], 16) for i in range(0, len(m.group()), 3))
        try:
            out.append(raw.decode("utf-8"))
        except UnicodeDecodeError:
            raise _BadXml("URI malformed") from None
        pos = m.end()
    if "%" in value[pos:]:
        raise _BadXml("URI malformed")
    out.append(value[pos:])
    return "".join(out)


def _part_path(base: str, target: str) -> str:
    """partPath (threemf.ts:185-196): an archive path relative to `base`, or absolute."""
    decoded = _decode_uri_component(target)
    if re.search(r"[?#\\]", decoded) or re.match(r"[a-z]+:", decoded, re.IGNORECASE):
        raise _BadXml("A model part path is invalid.")
    segments = [] if decoded.startswith("/") else base.split("/")[:-1]
    for segment in decoded.split("/"):
        if segment == "..":
            if not segments or not segments.pop():
                raise _BadXml("A model part path leaves the archive root.")
        elif segment and segment != ".":
            segments.append(segment)
    return "/".join(segments)


def _read_text(parts: Dict[str, ZipPart], path: str) -> Union[str, ModelParseError]:
    part = parts.get(path)
    if part is None:
        return ModelParseError("no-model-part", 0, "The referenced 3MF part is missing.", path)
    data = inflate_zip_part(part)
    if isinstance(data, ModelParseError):
        return data
    try:
        return data.decode("utf-8-sig")
    except UnicodeDecodeError as e:
        raise _BadXml(f"The encoded data was not valid for encoding utf-8: {e.reason}") from None


_REL_START = re.compile(r"(?:/3dmodel|/StartPart|^StartPart)\Z")


def _instance_vertices(obj: _Object, chain: List[tuple], unit_mm: float) -> Union[np.ndarray, ModelParseError]:
    v = np.asarray(obj.vertices, dtype=np.float64).reshape(-1, 3)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    for (a, b, c, d, e, f, g, h, i, tx, ty, tz) in chain:
        # Row vectors: component transforms act before the transforms that enclose them.
        x, y, z = x * a + y * d + z * g + tx, x * b + y * e + z * h + ty, x * c + y * f + z * i + tz
    with np.errstate(over="ignore", invalid="ignore"):
        out = np.stack([x * unit_mm, y * unit_mm, z * unit_mm], axis=1).astype(np.float32)
    if not np.all(np.isfinite(out)):
        return ModelParseError("non-finite", 0, "A transformed 3MF vertex exceeds the finite float32 range.")
    return out


def parse_three_mf(data: bytes, max_triangles: int = DEFAULT_MAX_TRIANGLES):
    """parseThreeMf (threemf.ts:246-350): (Mesh, unit_mm, item_count) in millimetres, or a
    ModelParseError. The start part comes from _rels/.rels (else 3D/3dmodel.model); build
    items and nested components (also in other parts) are instanced with their transforms."""
    limit = _check_limit(max_triangles)
    parts = read_zip(data)
    if isinstance(parts, ModelParseError):
        return parts
    active = {"part": "_rels/.rels"}
    try:
        root_path = MODEL_PATH
        if "_rels/.rels" in parts:
            rels = _read_text(parts, "_rels/.rels")
            if isinstance(rels, ModelParseError):
                return rels
            found = []

            def visit_rel(path, attrs):
                if path == "Relationships/Relationship" and _REL_START.search(attrs.get("Type", "")):
                    if attrs.get("TargetMode") == "External" or not attrs.get("Target"):
                        raise _BadXml("The start model must be an archive part.")
                    found.append(_part_path("", attrs["Target"]))

            scan_xml(rels, visit_rel)
            if found:
                root_path = found[-1]
        cache: Dict[str, _ModelPart] = {}

        def load(path: str) -> Union[_ModelPart, ModelParseError]:
            active["part"] = path
            if path in cache:
                return cache[path]
            text = _read_text(parts, path)
            if isinstance(text, ModelParseError):
                return text
            cache[path] = _parse_part(text, limit)
            return cache[path]

        root = load(root_path)
        if isinstance(root, ModelParseError):
            return root
        unit_mm = root.unit_mm
        instances = []
        total = [0]

        def expand(ref: _Reference, path: str, transforms: List[tuple], ancestors: set):
            target = _part_path(path, ref.path) if ref.path else path
            key = f"{target}#{ref.object_id}"
            if key in ancestors or len(ancestors) >= 128:
                raise _BadXml("The component hierarchy is cyclic or too deep.")
            model = load(target)
            if isinstance(model, ModelParseError):
                return model
            if model.unit_mm != unit_mm:
                raise _BadXml("All 3MF model parts must use the same unit.")
            obj = model.objects.get(ref.object_id)
            if obj is None:
                raise _BadXml("A build item or component references a missing object.")
            chain = [ref.transform] + transforms
            ancestors.add(key)
            for child in obj.components:
                err = expand(child, target, chain, ancestors)
                if err is not None:
                    return err
            ancestors.discard(key)
            total[0] += len(obj.indices) // 3
            if total[0] > limit:
                raise _TooManyTriangles("The 3MF exceeds the triangle limit.")
            instances.append((obj, chain))
            return None

        for item in root.items:
            err = expand(item, root_path, [], set())
            if err is not None:
                return err
        if total[0] == 0:
            return ModelParseError("empty", 0, "The 3MF build contains no triangles.", root_path)
        chunks = []
        for obj, chain in instances:
            vertices = _instance_vertices(obj, chain, unit_mm)
            if isinstance(vertices, ModelParseError):
                return vertices
            if obj.indices:
                chunks.append(vertices[np.asarray(obj.indices, dtype=np.int64)].reshape(-1))
        positions = np.concatenate(chunks).astype(np.float32) if chunks else np.zeros(0, np.float32)
        return Mesh(positions, total[0]), unit_mm, len(root.items)
    except _TooManyTriangles as e:
        return ModelParseError("too-many-triangles", 0, str(e), active["part"])
    except _BadXml as e:
        return ModelParseError("bad-xml", 0, str(e), active["part"])


# --------------------------------------------------------------------------------------------
# Dispatch

_OBJ_VERTEX_LINE = re.compile(f"(?:\\A|(?<=[\\n\\r\\u2028\\u2029])){_WS}*v[ \\t]+")
_CAD_NAME = re.compile(r"\.(?:step|stp|iges|igs)\Z", re.IGNORECASE)


def _text_format(data: bytes) -> Optional[str]:
    if len(data) >= 84 and 84 + struct.unpack_from("<I", data, 80)[0] * 50 == len(data):
        return "stl"
    text = data.decode("utf-8", errors="replace")
    if text.startswith("\ufeff"):  # TextDecoder drops one byte-order mark
        text = text[1:]
    if re.match(f"solid(?:{_WS}|\\Z)", text):
        return "stl"
    return "obj" if _OBJ_VERTEX_LINE.search(text) else None


def parse_model(data: bytes, file_name: str, max_triangles: int = DEFAULT_MAX_TRIANGLES) \
        -> Union[ParsedModel, ModelParseError]:
    """parseModel (model.ts:28-52): a ZIP is 3MF; otherwise exact binary STL, `solid` text or
    OBJ `v` lines decide, whatever the file name says; STEP/IGES names get an export hint."""
    limit = _check_limit(max_triangles)
    data = bytes(data)
    if len(data) >= 4 and struct.unpack_from("<I", data, 0)[0] == 0x04034B50:
        result = parse_three_mf(data, limit)
        if isinstance(result, ModelParseError):
            return result
        mesh, unit_mm, items = result
        return ParsedModel(mesh, "3mf", unit_mm, items)
    detected = _text_format(data)
    if detected is None and _CAD_NAME.search(file_name):
        return ModelParseError("unsupported-cad", 0, "Export this CAD model as STL, 3MF or OBJ before importing it.")
    fmt = detected or "stl"
    mesh = parse_obj(data, limit) if fmt == "obj" else parse_stl(data, limit)
    if isinstance(mesh, ModelParseError):
        return mesh
    return ParsedModel(mesh, fmt, None, 1)
