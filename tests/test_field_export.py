"""VTK field export on the host, restating tests/unit/vtk.test.ts and
tests/unit/io/field-export.test.ts, plus the Number-to-text rule of the SPACING line."""
from __future__ import annotations

import struct

import numpy as np
import pytest

from pocketwind_amd.field_export import (VECTORS_LINE, Field, LatticeField, field_file_name, field_vtk_bytes,
                                         js_number, lattice_to_field, write_vtk)

N8 = 8
CELLS = N8 ** 3


def vtk_field(cell_size=0.002):
    cell = np.arange(CELLS, dtype=np.float64)
    rho = (cell + 0.5).astype(np.float32)
    velocity = np.stack([cell, -cell, cell * 0.25], axis=1).astype(np.float32).reshape(-1)
    return Field(N8, N8, N8, cell_size, rho, velocity)


def be(data: bytes, offset: int) -> float:
    return struct.unpack_from(">f", data, offset)[0]


def test_writes_the_legacy_header_lines():
    data = write_vtk(vtk_field())
    marker = b"LOOKUP_TABLE default\n"
    rho_start = data.index(marker) + len(marker)
    assert data[:rho_start].decode().split("\n") == [
        "# vtk DataFile Version 3.0", "PocketWind field", "BINARY", "DATASET STRUCTURED_POINTS",
        "DIMENSIONS 8 8 8", "ORIGIN 0 0 0", "SPACING 0.002 0.002 0.002", "POINT_DATA 512",
        "SCALARS rho float 1", "LOOKUP_TABLE default", ""]


def test_writes_densities_then_velocities_as_big_endian_float32_and_ends_with_one_newline():
    data = write_vtk(vtk_field())
    marker = b"LOOKUP_TABLE default\n"
    rho_start = data.index(marker) + len(marker)
    velocity_start = rho_start + CELLS * 4 + 1 + len(VECTORS_LINE)
    for cell in range(100):
        assert be(data, rho_start + 4 * cell) == cell + 0.5
    assert data[rho_start + CELLS * 4] == 0x0A
    assert data[rho_start + CELLS * 4 + 1:velocity_start] == b"VECTORS velocity float\n"
    for cell in range(100):
        o = velocity_start + 12 * cell
        assert (be(data, o), be(data, o + 4), be(data, o + 8)) == (cell, -cell, cell * 0.25)
    assert len(data) == velocity_start + CELLS * 12 + 1 and data[-1] == 0x0A


def test_rejects_arrays_dimensions_and_spacing_that_cannot_describe_the_grid():
    f = vtk_field()
    with pytest.raises(ValueError):
        write_vtk(Field(8, 8, 8, 0.002, f.rho[:-1], f.velocity))
    with pytest.raises(ValueError):
        write_vtk(Field(8, 8, 8, 0.002, f.rho, f.velocity[:-1]))
    for bad in (Field(0, 8, 8, 0.002, f.rho, f.velocity), Field(8, 8, 2.5, 0.002, f.rho, f.velocity),
                Field(8, 8, 8, 0.0, f.rho, f.velocity), Field(8, 8, 8, float("nan"), f.rho, f.velocity)):
        with pytest.raises(ValueError):
            write_vtk(bad)


def lattice_input(speed=20.0):
    cell = np.arange(CELLS)
    rho = (1 + cell / 1024).astype(np.float32)
    velocity = np.tile(np.array([0.05, 0.025, -0.0125], np.float32), CELLS)
    return LatticeField(N8, N8, N8, 0.004, speed, rho, velocity)


def test_lattice_velocity_scales_so_the_inlet_speed_maps_to_the_physical_speed():
    source = lattice_input()
    field = lattice_to_field(source)
    # 20 m/s at a lattice speed of 0.05 is 400 m/s per lattice unit.
    assert field.velocity[:3].tolist() == [20.0, 10.0, -5.0]
    assert np.array_equal(field.rho, source.rho)
    assert source.velocity[0] == np.float32(0.05)


def test_field_file_header_size_and_values_read_back():
    header = ("# vtk DataFile Version 3.0\nPocketWind field\nBINARY\nDATASET STRUCTURED_POINTS\n"
              "DIMENSIONS 8 8 8\nORIGIN 0 0 0\nSPACING 0.004 0.004 0.004\nPOINT_DATA 512\n"
              "SCALARS rho float 1\nLOOKUP_TABLE default\n").encode()
    data = field_vtk_bytes(lattice_input())
    assert data[:len(header)] == header
    velocity_start = len(header) + CELLS * 4 + 1 + len(VECTORS_LINE)
    for cell in range(100):
        assert be(data, len(header) + 4 * cell) == np.float32(1 + cell / 1024)
        o = velocity_start + 12 * cell
        assert (be(data, o), be(data, o + 4), be(data, o + 8)) == (20.0, 10.0, -5.0)
    assert len(data) == velocity_start + CELLS * 12 + 1


def test_rejects_a_speed_that_cannot_scale_the_field():
    for speed in (0.0, float("nan"), -1.0, float("inf")):
        with pytest.raises(ValueError):
            lattice_to_field(lattice_input(speed))


def test_names_the_download_after_the_run_and_removes_reserved_characters():
    assert field_file_name("Ahmed 25") == "Ahmed 25-field.vtk"
    assert field_file_name("a/b:c") == "a-b-c-field.vtk"
    assert field_file_name("  ") == "untitled-field.vtk"
    assert field_file_name('x*y?z"<>|\\') == "x-y-z------field.vtk"


@pytest.mark.parametrize("value,text", [
    (0.002, "0.002"), (1.0, "1"), (100.0, "100"), (1e-7, "1e-7"), (1e-6, "0.000001"), (1.5e21, "1.5e+21"),
    (1e21, "1e+21"), (123456789012345680000.0, "123456789012345680000"), (0.1 + 0.2, "0.30000000000000004"),
    (5e-324, "5e-324"), (-0.25, "-0.25")])
def test_spacing_text_follows_number_to_string(value, text):
    assert js_number(value) == text
