"""Mesh import (STL, OBJ, ZIP, 3MF, content dispatch), restating tests/unit/stl.test.ts,
tests/unit/obj/obj.test.ts, tests/unit/zip/zip.test.ts, tests/unit/threemf/threemf.test.ts
and tests/unit/model/model.test.ts on the reference's own fixture files (copied under
tests/golden/mesh/), and an imported closed ellipsoid voxelised watertight by the oracle."""
from __future__ import annotations

import os
import struct
from dataclasses import replace

import numpy as np
import pytest

from archives import (CUBE_OBJECT, MODEL_PATH, ellipsoid_archive, make_archive, model_archive, model_xml)
from pocketwind_amd.mesh_io import (Mesh, ModelParseError, ParsedModel, crc32, inflate_zip_part, parse_model,
                                    parse_obj, parse_stl, parse_three_mf, read_zip)

HERE = os.path.join(os.path.dirname(__file__), "golden", "mesh")


def fixture(name: str) -> bytes:
    return open(os.path.join(HERE, name), "rb").read()


def bounds(mesh: Mesh, lo, hi):
    p = mesh.positions.reshape(-1, 3).astype(np.float64)
    assert np.all(np.abs(p.min(0) - lo) <= 1e-3) and np.all(np.abs(p.max(0) - hi) <= 1e-3)


def kind(result):
    return result.kind if isinstance(result, ModelParseError) else None


# ---------------------------------------------------------------------------------- STL

@pytest.mark.parametrize("name", ["cube-binary", "cube-ascii-crlf", "solid-header-binary"])
def test_stl_reads_each_cube_with_the_same_vertices(name):
    mesh = parse_stl(fixture(f"{name}.stl"))
    ref = parse_stl(fixture("cube-binary.stl"))
    assert mesh.triangle_count == 12 and mesh.positions.size == 108
    assert np.array_equal(mesh.positions, ref.positions)
    assert set(mesh.positions.tolist()) == {0.0, 1.0}


def test_stl_errors_of_the_fixture_files():
    truncated = fixture("truncated.stl")
    assert parse_stl(truncated) == ModelParseError("truncated", len(truncated), parse_stl(truncated).message)
    assert (kind(parse_stl(fixture("random.stl"))), parse_stl(fixture("random.stl")).byte_offset) == ("not-stl", 0)
    assert (kind(parse_stl(fixture("empty.stl"))), parse_stl(fixture("empty.stl")).byte_offset) == ("empty", 80)
    nan = parse_stl(fixture("nan-vertex.stl"))
    assert (nan.kind, nan.byte_offset) == ("non-finite", 96)


@pytest.mark.parametrize("name", ["cube-binary", "cube-ascii-crlf"])
def test_stl_triangle_limit(name):
    assert kind(parse_stl(fixture(f"{name}.stl"), 11)) == "too-many-triangles"


def test_stl_ignores_binary_normals_and_attributes_and_rejects_trailing_bytes():
    data = bytearray(fixture("cube-binary.stl"))
    struct.pack_into("<f", data, 84, float("nan"))
    struct.pack_into("<H", data, 132, 65535)
    assert parse_stl(bytes(data)).triangle_count == 12
    assert kind(parse_stl(fixture("cube-binary.stl") + b"\0")) == "not-stl"


def test_stl_ascii_errors_at_their_offsets():
    assert kind(parse_stl(b"solid empty\nendsolid empty\n")) == "empty"
    data = b"solid cube\nfacet normal 0 0 0\nouter loop\nvertex 0 0"
    assert (parse_stl(data).kind, parse_stl(data).byte_offset) == ("truncated", len(data))
    text = "solid cube\nfacet normal 0 0 0\nouter loop\nwrong 0 0 0"
    r = parse_stl(text.encode())
    assert (r.kind, r.byte_offset) == ("not-stl", text.index("wrong"))


@pytest.mark.parametrize("value", ["NaN", "Infinity", "-Infinity", "1e400", "1e39"])
def test_stl_rejects_non_finite_ascii_vertices(value):
    text = f"solid cube\nfacet normal 0 0 0\nouter loop\nvertex {value} 0 0"
    r = parse_stl(text.encode())
    assert (r.kind, r.byte_offset) == ("non-finite", text.index(value))


@pytest.mark.parametrize("limit", [-1, 0.5, float("nan"), float("inf")])
def test_stl_rejects_invalid_limits(limit):
    with pytest.raises(ValueError):
        parse_stl(fixture("empty.stl"), limit)


def test_stl_offsets_are_bytes_after_a_utf8_name_and_bad_tokens_are_reported():
    prefix = "solid café\nfacet normal 0 0 0\nouter loop\nvertex "
    r = parse_stl(f"{prefix}NaN 0 0".encode())
    assert (r.kind, r.byte_offset) == ("non-finite", len(prefix.encode()))
    prefix = "solid cube\nfacet normal 0 0 0\nouter loop\nvertex "
    r = parse_stl(f"{prefix}wrong 0 0".encode())
    assert (r.kind, r.byte_offset) == ("not-stl", len(prefix))


def test_stl_rejects_content_after_endsolid_and_reads_signed_scientific_values():
    text = fixture("cube-ascii-crlf.stl").decode()
    r = parse_stl((text + "facet").encode())
    assert (r.kind, r.byte_offset) == ("not-stl", len(text))
    mesh = parse_stl(text.replace("vertex 0 0 0", "vertex -1e-2 +.5 1.", 1).encode())
    assert mesh.positions[:3].tolist() == [float(np.float32(-0.01)), 0.5, 1.0]


def test_stl_default_cap_applies_before_reading_records():
    data = bytearray(84)
    struct.pack_into("<I", data, 80, 2_000_001)
    r = parse_stl(bytes(data))
    assert (r.kind, r.byte_offset) == ("too-many-triangles", 80)


# ---------------------------------------------------------------------------------- OBJ

VERTS = "v 0 0 0\r\nv 1 0 0\r\nv 1 1 0\r\nv 0 1 0\r\n"


def obj(text: str, limit=None):
    return parse_obj(text.encode()) if limit is None else parse_obj(text.encode(), limit)


def test_obj_fans_the_quad_cube():
    mesh = parse_obj(fixture("quad-cube.obj"))
    assert mesh.triangle_count == 12
    bounds(mesh, [0, 0, 0], [300, 300, 300])


@pytest.mark.parametrize("face", ["1 2 3 4", "1/9/8 2/7/6 3/5/4 4/3/2", "1//4 2//3 3//2 4//1", "-4 -3 -2 -1"])
def test_obj_resolves_vertex_references(face):
    mesh = obj(f"{VERTS}vt 0 0\nvn 0 0 1\ng cube\nusemtl red\n f {face} # face\n")
    assert mesh.triangle_count == 2
    assert mesh.positions.tolist() == [0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 0, 0, 1, 0]


def test_obj_negative_indices_resolve_at_the_face():
    assert obj(f"{VERTS}f -4 -3 -2\nv 9 9 9").triangle_count == 1


@pytest.mark.parametrize("face", ["0 1 2", "1 2 5", "-5 -2 -1", "1.5 2 3", "1 2", "1/no 2 3", "1/ 2/ 3/"])
def test_obj_rejects_invalid_faces(face):
    assert kind(obj(f"{VERTS}f {face}")) == "not-stl"


def test_obj_rejects_malformed_and_overflowing_coordinates():
    assert kind(obj("v 1 2")) == "not-stl"
    assert kind(obj("v 1e50 0 0")) == "non-finite"
    assert kind(obj("v NaN 0 0")) == "not-stl"


def test_obj_offsets_count_bytes_after_a_bom_utf8_comment_and_mixed_endings():
    prefix = "\ufeff# café\r\n" + VERTS
    r = obj(f"{prefix}f 0 1 2")
    assert (r.kind, r.byte_offset) == ("not-stl", len(prefix.encode()))


def test_obj_limits_and_empty():
    assert kind(obj(f"{VERTS}f 1 2 3 4", 1)) == "too-many-triangles"
    assert kind(obj(VERTS)) == "empty"


# ---------------------------------------------------------------------------------- ZIP

@pytest.mark.parametrize("method", [0, 8])
def test_zip_reads_stored_and_deflated_members(method):
    parts = read_zip(make_archive([{"name": "part.xml", "text": "mesh & café", "method": method},
                                   {"name": "empty", "text": "", "method": method}]))
    assert inflate_zip_part(parts["part.xml"]) == "mesh & café".encode()
    assert inflate_zip_part(parts["empty"]) == b""


def test_crc32_check_value():
    assert crc32(b"123456789") == 0xCBF43926


@pytest.mark.parametrize("name", ["zip64", "encrypted", "unsupported-method"])
def test_zip_rejects_unsupported_archives(name):
    assert kind(read_zip(fixture(f"{name}.3mf"))) == "zip-unsupported"


def test_zip_rejects_every_truncated_prefix():
    data = make_archive([{"name": "a", "text": "abc"}])
    for n in range(len(data)):
        assert kind(read_zip(data[:n])) == "truncated"


def test_zip_honours_the_comment_length_over_a_signature_inside_the_comment():
    archive = make_archive([{"name": "a", "text": "abc"}])
    data = bytearray(archive + bytes(30))
    struct.pack_into("<H", data, len(archive) - 2, 30)
    struct.pack_into("<I", data, len(archive) + 3, 0x06054B50)
    assert isinstance(read_zip(bytes(data)), dict)


@pytest.mark.parametrize("method", [0, 8])
def test_zip_rejects_corrupt_members(method):
    part = read_zip(make_archive([{"name": "a", "text": "abc", "method": method}]))["a"]
    bad = replace(part, compressed=b"\xff" * len(part.compressed))
    r = inflate_zip_part(bad)
    assert (r.kind, r.part) == ("truncated", "a")


def test_zip_rejects_sizes_that_disagree_either_way():
    part = read_zip(make_archive([{"name": "a", "text": "abc"}]))["a"]
    for size in (1, 10):
        assert kind(inflate_zip_part(replace(part, size=size))) == "truncated"


def test_zip_data_descriptors_use_the_central_directory_sizes():
    data = bytearray(make_archive([{"name": "a", "text": "abc", "flags": 8}]))
    struct.pack_into("<III", data, 14, 0, 0, 0)
    assert inflate_zip_part(read_zip(bytes(data))["a"]) == b"abc"


def test_zip_rejects_a_forged_oversized_archive_before_expansion():
    data = bytearray(make_archive([{"name": "a", "text": "abc"}]))
    central = struct.unpack_from("<I", data, len(data) - 6)[0]
    struct.pack_into("<I", data, central + 24, 513 * 1024 * 1024)
    assert kind(read_zip(bytes(data))) == "zip-unsupported"


# ---------------------------------------------------------------------------------- 3MF

@pytest.mark.parametrize("name,count,lo,hi,items,unit", [
    ("single-object", 12, [0, 0, 0], [300, 300, 300], 1, 1.0),
    ("stored", 12, [0, 0, 0], [300, 300, 300], 1, 1.0),
    ("inch", 12, [0, 0, 0], [7620, 7620, 7620], 1, 25.4),
    ("components", 24, [-300, 0, 0], [980, 310, 330], 2, 1.0),
    ("external-part", 12, [10, 20, 30], [310, 320, 330], 1, 1.0)])
def test_3mf_reads_fixtures_with_merged_bounds(name, count, lo, hi, items, unit):
    mesh, unit_mm, item_count = parse_three_mf(fixture(f"{name}.3mf"))
    assert (mesh.triangle_count, unit_mm, item_count) == (count, unit, items)
    bounds(mesh, lo, hi)


@pytest.mark.parametrize("unit,mm", [("micron", 0.001), ("millimeter", 1.0), ("centimeter", 10.0),
                                     ("inch", 25.4), ("foot", 304.8), ("meter", 1000.0)])
def test_3mf_units_convert_to_millimetres(unit, mm):
    mesh, unit_mm, _ = parse_three_mf(model_archive(model_xml(CUBE_OBJECT, '<item objectid="1"/>', unit)))
    assert unit_mm == mm
    bounds(mesh, [0, 0, 0], [300 * mm] * 3)


@pytest.mark.parametrize("name,expected", [("zip64", "zip-unsupported"), ("encrypted", "zip-unsupported"),
                                           ("unsupported-method", "zip-unsupported"),
                                           ("missing-model", "no-model-part")])
def test_3mf_reports_archive_problems(name, expected):
    assert kind(parse_three_mf(fixture(f"{name}.3mf"))) == expected


def test_3mf_start_relationship_prefixes_single_quotes_and_escaped_paths():
    text = model_xml().replace('"', "'").replace("<model", "<c:model").replace("</model>", "</c:model>") \
        .replace("xmlns=", "xmlns:c=")
    data = make_archive([
        {"name": "_rels/.rels", "text": '<r:Relationships xmlns:r="urn:rels"><r:Relationship Type="StartPart" '
                                        'Target="/3D/a&amp;b.model"/></r:Relationships>'},
        {"name": "3D/a&b.model", "text": text}])
    assert parse_three_mf(data)[0].triangle_count == 12


def test_3mf_default_path_and_unit():
    data = make_archive([{"name": MODEL_PATH, "text": model_xml().replace(' unit="millimeter"', "")}])
    mesh, unit_mm, _ = parse_three_mf(data)
    assert (unit_mm, mesh.triangle_count) == (1.0, 12)


def test_3mf_ignores_metadata_comments_colours_and_unused_resources():
    text = model_xml(f'{CUBE_OBJECT}<object id="2"/>').replace(
        "<resources>", '<!-- <bad> --><metadata name="x"><![CDATA[a < b]]></metadata><resources><colorgroup id="3"/>')
    assert parse_three_mf(model_archive(text))[0].triangle_count == 12


@pytest.mark.parametrize("text", [
    "<model><resources></model>",
    "<!DOCTYPE model><model/>",
    "<model unit=millimeter/>",
    model_xml().replace('v1="0"', 'v1="900"'),
    model_xml().replace('x="0"', 'x="no"'),
    model_xml().replace('id="1"', 'id="1" id="2"'),
    model_xml(CUBE_OBJECT, '<item objectid="9"/>'),
    model_xml(CUBE_OBJECT, '<item objectid="1" transform="1 2"/>'),
    model_xml(CUBE_OBJECT, '<item objectid="1"/>', "yard"),
    model_xml('<object id="1"><components><component objectid="1"/></components></object>')])
def test_3mf_rejects_invalid_models(text):
    r = parse_three_mf(model_archive(text))
    assert (r.kind, r.part) == ("bad-xml", MODEL_PATH)


def test_3mf_reports_a_missing_external_part():
    text = model_xml('<object id="1"><components><component objectid="2" p:path="Objects/missing.model"/>'
                     '</components></object>')
    r = parse_three_mf(model_archive(text))
    assert (r.kind, r.part) == ("no-model-part", "3D/Objects/missing.model")


def test_3mf_limits_merged_instances_as_well_as_meshes():
    assert kind(parse_three_mf(fixture("single-object.3mf"), 11)) == "too-many-triangles"
    two = model_archive(model_xml(CUBE_OBJECT, '<item objectid="1"/><item objectid="1"/>'))
    assert kind(parse_three_mf(two, 12)) == "too-many-triangles"


def test_3mf_rejects_overflow_after_transforms():
    text = model_xml(CUBE_OBJECT, '<item objectid="1" transform="1e40 0 0 0 1 0 0 0 1 0 0 0"/>')
    assert kind(parse_three_mf(model_archive(text))) == "non-finite"


def test_3mf_closed_ellipsoid_imports_and_voxelises_watertight():
    import oracle
    mesh, _, _ = parse_three_mf(ellipsoid_archive())
    assert mesh.triangle_count == 327680
    bounds(mesh, [-450, -150, -150], [450, 150, 150])
    p = mesh.positions.astype(np.float64).reshape(-1, 3)
    positions = (p / 40 + np.array([16, 8, 8])).astype(np.float32).reshape(-1)
    grid = oracle.voxelize(positions, None, (32, 16, 16), False, False)
    assert grid.solid_cells > 0 and grid.watertight


# ---------------------------------------------------------------------------------- dispatch

@pytest.mark.parametrize("name,fmt,unit", [
    ("cube-binary.stl", "stl", None), ("cube-ascii-crlf.stl", "stl", None), ("solid-header-binary.stl", "stl", None),
    ("single-object.3mf", "3mf", 1.0), ("quad-cube.obj", "obj", None), ("inch.3mf", "3mf", 25.4)])
def test_model_format_comes_from_the_content(name, fmt, unit):
    r = parse_model(fixture(name), "misnamed.step")
    assert isinstance(r, ParsedModel)
    assert (r.format, r.unit_mm, r.item_count, r.mesh.triangle_count) == (fmt, unit, 1, 12)


@pytest.mark.parametrize("name", ["part.step", "part.STP", "part.iges", "part.IGS"])
def test_model_cad_names_get_an_export_hint(name):
    r = parse_model(b"ISO-10303-21;", name)
    assert r.kind == "unsupported-cad" and "Export" in r.message


def test_model_keeps_stl_errors_and_limits_across_formats():
    assert kind(parse_model(fixture("truncated.stl"), "bad.stl")) == "truncated"
    for name in ("cube-binary.stl", "single-object.3mf", "quad-cube.obj"):
        assert kind(parse_model(fixture(name), name, 11)) == "too-many-triangles"
    for limit in (-1, float("nan"), 0.5):
        with pytest.raises(ValueError):
            parse_model(b"", "x", limit)


def test_model_obj_text_in_a_binary_stl_header_keeps_stl():
    data = bytearray(fixture("cube-binary.stl"))
    data[:8] = b"v 0 0 0\n"
    assert parse_model(bytes(data), "cube.obj").format == "stl"
