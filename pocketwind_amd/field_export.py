"""Field export as a legacy VTK STRUCTURED_POINTS file (src/io/vtk.ts, src/io/field-export.ts).

write_vtk            writeVtk (vtk.ts:27-80): header lines, big-endian float32 densities, a
                     newline, the VECTORS line, big-endian float32 velocity triples, a newline
lattice_to_field     latticeToField (field-export.ts:22-30): velocity * speedMps / U_LU, rounded
                     to float32 as Float32Array.map does
field_vtk_bytes      fieldVtkBytes (field-export.ts:33-35)
field_file_name      fieldFileName (field-export.ts:37-40)
solver_field_vtk     the same file for a solver's macro buffer, with the binary blocks made on
                     the device in one pass (pw_solver_field_vtk_blocks)
"""
from __future__ import annotations

import ctypes as C
import math
import re
from dataclasses import dataclass
from decimal import Decimal

import numpy as np

from . import _native as N
from .drag import U_LU

VECTORS_LINE = b"VECTORS velocity float\n"


@dataclass(frozen=True)
class Field:
    """Field (vtk.ts:1-8): rho per cell, velocity interleaved x, y, z per cell."""

    nx: int
    ny: int
    nz: int
    cell_size_m: float
    rho: np.ndarray
    velocity: np.ndarray


@dataclass(frozen=True)
class LatticeField:
    """LatticeField (field-export.ts:4-17): lattice-unit rho and velocity plus the inlet speed in
    m/s that fixes the velocity scale."""

    nx: int
    ny: int
    nz: int
    cell_size_m: float
    speed_mps: float
    rho: np.ndarray
    velocity: np.ndarray


def js_number(x: float) -> str:
    """Number.prototype.toString for a finite double (ECMA-262 Number::toString): the shortest
    round-trip digits, plain notation for decimal exponents in [-6, 21), otherwise d.ddde+n."""
    if x == 0:
        return "0"
    if x < 0:
        return "-" + js_number(-x)
    sign, digits, exp = Decimal(repr(float(x))).normalize().as_tuple()
    s = "".join(map(str, digits))
    k = len(s)
    n = exp + k  # value = 0.s * 10^n
    if k <= n <= 21:
        return s + "0" * (n - k)
    if 0 < n <= 21:
        return s[:n] + "." + s[n:]
    if -6 < n <= 0:
        return "0." + "0" * (-n) + s
    e = n - 1
    mantissa = s if k == 1 else s[0] + "." + s[1:]
    return f"{mantissa}e{'+' if e >= 0 else '-'}{abs(e)}"


def _check_dimensions(nx, ny, nz) -> int:
    if not all(isinstance(v, (int, np.integer)) and not isinstance(v, bool) and v > 0 for v in (nx, ny, nz)):
        raise ValueError("Field dimensions must be positive integers.")
    cells = int(nx) * int(ny) * int(nz)
    if cells > 2 ** 53 - 1:
        raise ValueError("The field is too large to address.")
    return cells


def _header(nx: int, ny: int, nz: int, cell_size_m: float, cells: int) -> bytes:
    spacing = js_number(cell_size_m)
    return "\n".join([
        "# vtk DataFile Version 3.0",
        "PocketWind field",
        "BINARY",
        "DATASET STRUCTURED_POINTS",
        f"DIMENSIONS {nx} {ny} {nz}",
        "ORIGIN 0 0 0",
        f"SPACING {spacing} {spacing} {spacing}",
        f"POINT_DATA {cells}",
        "SCALARS rho float 1",
        "LOOKUP_TABLE default",
        "",
    ]).encode()


def _check_spacing(cell_size_m) -> None:
    if not isinstance(cell_size_m, (int, float, np.floating)) or not math.isfinite(cell_size_m) or cell_size_m <= 0:
        raise ValueError("cellSizeM must be a positive finite number.")


def _assemble(header: bytes, rho_be: bytes, velocity_be: bytes) -> bytes:
    return b"".join([header, rho_be, b"\n", VECTORS_LINE, velocity_be, b"\n"])


def write_vtk(field: Field) -> bytes:
    """writeVtk (vtk.ts:27-80). ValueError plays RangeError."""
    cells = _check_dimensions(field.nx, field.ny, field.nz)
    rho = np.asarray(field.rho, dtype=np.float32).reshape(-1)
    velocity = np.asarray(field.velocity, dtype=np.float32).reshape(-1)
    if rho.size != cells:
        raise ValueError(f"rho must hold {cells} values.")
    if velocity.size != 3 * cells:
        raise ValueError(f"velocity must hold {3 * cells} values.")
    _check_spacing(field.cell_size_m)
    return _assemble(_header(field.nx, field.ny, field.nz, float(field.cell_size_m), cells),
                     rho.astype(">f4").tobytes(), velocity.astype(">f4").tobytes())


def _velocity_scale(speed_mps) -> float:
    if not isinstance(speed_mps, (int, float, np.floating)) or not math.isfinite(speed_mps) or speed_mps <= 0:
        raise ValueError("speedMps must be a positive finite number.")
    return float(speed_mps) / U_LU


def lattice_to_field(field: LatticeField) -> Field:
    """latticeToField (field-export.ts:22-30): the caller's arrays are not changed."""
    scale = _velocity_scale(field.speed_mps)
    velocity = (np.asarray(field.velocity, dtype=np.float32).astype(np.float64) * scale).astype(np.float32)
    return Field(field.nx, field.ny, field.nz, field.cell_size_m, field.rho, velocity)


def field_vtk_bytes(field: LatticeField) -> bytes:
    """fieldVtkBytes (field-export.ts:33-35)."""
    return write_vtk(lattice_to_field(field))


_RESERVED = re.compile(r'[\\/:*?"<>|]')
# String.prototype.trim's WhiteSpace and LineTerminator code points.
_JS_SPACE = ("\t\n\v\f\r \u00a0\u1680\u2000\u2001\u2002\u2003\u2004\u2005\u2006\u2007\u2008"
             "\u2009\u200a\u2028\u2029\u202f\u205f\u3000\ufeff")


def field_file_name(name: str) -> str:
    """fieldFileName (field-export.ts:37-40)."""
    safe = _RESERVED.sub("-", name).strip(_JS_SPACE)
    return f"{'untitled' if safe == '' else safe}-field.vtk"


def solver_field_vtk(solver, cell_size_m: float, speed_mps: float) -> bytes:
    """fieldVtkBytes for the solver's macro buffer: the densities and the scaled velocities are
    converted to big-endian float32 on the device and copied back as the file's two blocks."""
    layout = solver.layout
    cells = _check_dimensions(layout.nx, layout.ny, layout.nz)
    _check_spacing(cell_size_m)
    scale = _velocity_scale(speed_mps)
    rho_be = np.empty(4 * cells, np.uint8)
    velocity_be = np.empty(12 * cells, np.uint8)
    N.check(N.lib().pw_solver_field_vtk_blocks(solver._live(), C.c_double(scale), rho_be.ctypes.data, rho_be.size,
                                               velocity_be.ctypes.data, velocity_be.size))
    return _assemble(_header(layout.nx, layout.ny, layout.nz, float(cell_size_m), cells), rho_be.tobytes(),
                     velocity_be.tobytes())
