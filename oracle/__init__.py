"""FP64 CPU oracle for pocketwind's D3Q19 wind-tunnel hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``pocketwind_amd`` never does. It binds ``oracle/libpworacle.so`` (built by
``oracle/Makefile`` from the C restatement in this directory) with ctypes.

Reference correspondence (file:line, relative to the pocketwind repository):
  Solver            src/lbm/reference/cpu-solver.ts:69-323
  equilibrium etc.  src/lbm/reference/collision.ts:21-224
  voxelize          src/geometry/voxelize.ts:172-199, wall-links.ts:99-131
  nearest_link_hit  src/geometry/triangle-bins.ts:136-162 (via nearestLinkHit)
  preset_mesh       src/geometry/presets.ts:179-184
  place_mesh        src/geometry/transform.ts:118-143
  particle_*        src/render/particle-layout.ts:14-112, shaders/particles_update.wgsl
  encode_f16        src/lbm/shaders/prelude-f16.wgsl:7-22, common.wgsl:134-135
                    (pinned by tests/test_oracle_f16_encoding.py: exact halves pass, subnormals
                    round to nearest even as numpy's float16 cast, normal values land on one of
                    their two neighbours with the dropped fraction's probability; the reference's
                    tests hold no known encodings, so the hash's bit mapping follows the cited
                    constants without an external known answer)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpworacle.so")

TRT, REG, BGK = 0, 1, 2
BB, BOUZIDI = 0, 1
FLUID, SOLID, SUPPORT, FLOOR, EQ = 0, 1, 2, 3, 4
NO_Q = 0xFFFFFFFF

C19 = np.array(
    [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
     [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0],
     [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]], dtype=np.int64)
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)


class _VoxelGrid(C.Structure):
    _fields_ = [("geometry", _u32), ("geometry_words", C.c_size_t),
                ("boundary_cells", _u32), ("n_boundary", C.c_size_t),
                ("solid_cells", C.c_uint64), ("frontal_cells", C.c_uint64),
                ("blockage", C.c_double), ("watertight", C.c_int), ("missed_links", C.c_uint64)]


def _declare(L):
    L.pwo_opp.argtypes = [C.c_int]
    L.pwo_opp.restype = C.c_int
    L.pwo_equilibrium.argtypes = [_d, C.c_double, _d]
    L.pwo_moments.argtypes = [_d, _d, _d]
    L.pwo_collide.argtypes = [_d, _d, C.c_double, C.c_double, C.c_int, C.c_int, _d]
    L.pwo_friction_velocity.argtypes = [C.c_double, C.c_double]
    L.pwo_friction_velocity.restype = C.c_double
    L.pwo_collide_wall.argtypes = [_d, _d, C.c_double, _d, C.POINTER(C.c_int), C.c_int,
                                   C.c_double, _d]
    L.pwo_collide_wall.restype = C.c_int
    L.pwo_solver_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_int,
                                    C.c_int, C.c_int, _u32, C.c_size_t, _d, _d, _d,
                                    C.c_char_p, C.c_size_t]
    L.pwo_solver_create.restype = C.c_void_p
    L.pwo_solver_create_bc.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, _u32, C.c_size_t, _d, _d,
                                       _d, C.c_char_p, C.c_size_t]
    L.pwo_solver_create_bc.restype = C.c_void_p
    L.pwo_solver_source.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
    L.pwo_solver_source.restype = C.c_int64
    L.pwo_solver_destroy.argtypes = [C.c_void_p]
    L.pwo_solver_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.pwo_solver_current.argtypes = [C.c_void_p]
    L.pwo_solver_current.restype = _d
    L.pwo_solver_previous.argtypes = [C.c_void_p]
    L.pwo_solver_previous.restype = _d
    L.pwo_solver_macroscopic.argtypes = [C.c_void_p, C.c_size_t, _d]
    L.pwo_solver_macro_all.argtypes = [C.c_void_p, _d]
    L.pwo_solver_momentum_exchange.argtypes = [C.c_void_p, _d]
    L.pwo_solver_link_count.argtypes = [C.c_void_p]
    L.pwo_solver_link_count.restype = C.c_size_t
    L.pwo_voxelize.argtypes = [_f, C.c_size_t, _f, C.c_size_t, C.c_uint32, C.c_uint32,
                               C.c_uint32, C.c_int, C.c_int, C.POINTER(_VoxelGrid),
                               C.c_char_p, C.c_size_t]
    L.pwo_voxelize.restype = C.c_int
    L.pwo_build_wall_links.argtypes = [_u32, _f, C.c_size_t, _f, C.c_size_t, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                       C.POINTER(_VoxelGrid)]
    L.pwo_voxel_grid_free.argtypes = [C.POINTER(_VoxelGrid)]
    L.pwo_nearest_link_hit.argtypes = [_f, C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint32,
                                       _d, _d, _d]
    L.pwo_nearest_link_hit.restype = C.c_int
    L.pwo_preset_mesh.argtypes = [C.c_int, _f]
    L.pwo_preset_mesh.restype = C.c_size_t
    L.pwo_place_mesh.argtypes = [_f, C.c_size_t, _f, C.c_size_t, C.c_double, C.c_uint32,
                                 C.c_uint32, C.c_uint32, C.c_int, _f, _f]
    L.pwo_place_mesh.restype = C.c_int
    L.pwo_particle_hash.argtypes = [C.c_uint32]
    L.pwo_particle_hash.restype = C.c_uint32
    L.pwo_particle_needs_reseed.argtypes = [_d, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint32]
    L.pwo_particle_needs_reseed.restype = C.c_int
    L.pwo_seed_region.argtypes = [_d, _d, C.c_uint32, C.c_uint32, C.c_uint32, _d]
    L.pwo_seed_region.restype = C.c_int
    L.pwo_particle_seeds.argtypes = [_u32, C.c_uint32, C.c_uint32, C.c_uint32, _d, _d,
                                     C.c_uint32, C.c_uint32, _f, _f, _d, C.c_char_p, C.c_size_t]
    L.pwo_particle_seeds.restype = C.c_int
    L.pwo_particles_update.argtypes = [_f, C.c_uint32, _f, _u32, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_float, _f, _f]
    L.pwo_encode_f16.argtypes = [C.c_float, C.c_uint32, C.c_uint32]
    L.pwo_encode_f16.restype = C.c_uint16
    L.pwo_encode_f16_planes.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
    L.pwo_encode_f16_planes.restype = None
    L.pwo_decode_f16.argtypes = [C.c_uint16]
    L.pwo_decode_f16.restype = C.c_float
    L.pwo_encode_bf16.argtypes = [C.c_float, C.c_uint32, C.c_uint32]
    L.pwo_encode_bf16.restype = C.c_uint16
    L.pwo_decode_bf16.argtypes = [C.c_uint16]
    L.pwo_decode_bf16.restype = C.c_float


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


# ---------------------------------------------------------------- collision.ts
def opp(i: int) -> int:
    return lib().pwo_opp(i)


def equilibrium(rho: float, u) -> np.ndarray:
    out = np.zeros(19)
    uu = np.asarray(u, dtype=np.float64)
    lib().pwo_equilibrium(_ptr(out, _d), rho, _ptr(uu, _d))
    return out


def moments(g) -> tuple[float, np.ndarray]:
    g = np.ascontiguousarray(g, dtype=np.float64)
    rho = C.c_double()
    u = np.zeros(3)
    lib().pwo_moments(_ptr(g, _d), C.byref(rho), _ptr(u, _d))
    return rho.value, u


def collide(g, geq, rho, tau0, op=TRT, les=False, force_terms=None) -> np.ndarray:
    g = np.array(g, dtype=np.float64)
    geq = np.ascontiguousarray(geq, dtype=np.float64)
    ft = np.zeros(19) if force_terms is None else np.ascontiguousarray(force_terms, np.float64)
    lib().pwo_collide(_ptr(g, _d), _ptr(geq, _d), rho, tau0, op, int(les), _ptr(ft, _d))
    return g


def friction_velocity(speed: float, viscosity: float) -> float:
    return lib().pwo_friction_velocity(speed, viscosity)


def collide_wall(g, geq, rho, u, wall_incoming, viscosity, force_terms=None):
    """Returns (applied, g_after)."""
    g = np.array(g, dtype=np.float64)
    geq = np.ascontiguousarray(geq, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    inc = (C.c_int * len(wall_incoming))(*wall_incoming)
    ft = np.zeros(19) if force_terms is None else np.ascontiguousarray(force_terms, np.float64)
    ok = lib().pwo_collide_wall(_ptr(g, _d), _ptr(geq, _d), rho, _ptr(u, _d), inc,
                                len(wall_incoming), viscosity, _ptr(ft, _d))
    return bool(ok), g


# ---------------------------------------------------------------- cpu-solver.ts
class Solver:
    """FP64 CpuSolver (cpu-solver.ts:69-323). ``populations`` views the latest state."""

    def __init__(self, layout, tau0, operator=TRT, les=False, walls=BB, flags=None,
                 body_force=None, equilibrium_state=None, initial=None, threads=1,
                 sides="periodic", outlet="equilibrium"):
        self.nx, self.ny, self.nz = (int(v) for v in layout)
        self.cells = self.nx * self.ny * self.nz
        self.threads = threads
        fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        bf = None if body_force is None else np.ascontiguousarray(body_force, np.float64)
        eq = None
        if equilibrium_state is not None:
            rho, u = equilibrium_state
            eq = np.array([rho, *u], dtype=np.float64)
        init = None if initial is None else np.ascontiguousarray(initial, np.float64).reshape(-1)
        err = C.create_string_buffer(256)
        bc_sides = {"periodic": 0, "free-slip": 1}[sides]
        bc_outlet = {"equilibrium": 0, "zero-gradient": 1}[outlet]
        self._h = lib().pwo_solver_create_bc(
            self.nx, self.ny, self.nz, float(tau0), int(operator), int(les), int(walls),
            bc_sides, bc_outlet, _ptr(fl, _u32), 0 if fl is None else fl.size, _ptr(bf, _d),
            _ptr(eq, _d), _ptr(init, _d), err, 256)
        if not self._h:
            raise ValueError(err.value.decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pwo_solver_destroy(h)
            self._h = None

    def step(self, count: int = 1):
        lib().pwo_solver_step(self._h, int(count), int(self.threads))

    def source(self, cell: int, i: int) -> int:
        """Flat state index fluid cell ``cell`` pulls direction i from (-1 for other cells)."""
        return int(lib().pwo_solver_source(self._h, int(cell), int(i)))

    def _view(self, p):
        return np.ctypeslib.as_array(p, shape=(19 * self.cells,))

    @property
    def populations(self) -> np.ndarray:
        return self._view(lib().pwo_solver_current(self._h))

    @property
    def previous(self) -> np.ndarray:
        return self._view(lib().pwo_solver_previous(self._h))

    def macroscopic(self, cell: int):
        out = np.zeros(4)
        lib().pwo_solver_macroscopic(self._h, int(cell), _ptr(out, _d))
        return out[0], out[1:].copy()

    def macro_all(self) -> np.ndarray:
        out = np.zeros(4 * self.cells)
        lib().pwo_solver_macro_all(self._h, _ptr(out, _d))
        return out.reshape(self.cells, 4)

    def momentum_exchange(self) -> np.ndarray:
        out = np.zeros(3)
        lib().pwo_solver_momentum_exchange(self._h, _ptr(out, _d))
        return out

    @property
    def link_count(self) -> int:
        return lib().pwo_solver_link_count(self._h)


# ---------------------------------------------------------------- voxelizer
@dataclass
class VoxelGrid:
    nx: int
    ny: int
    nz: int
    geometry: np.ndarray
    boundary_cells: np.ndarray
    solid_cells: int
    frontal_cells: int
    blockage: float
    watertight: bool
    missed_links: int

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    @property
    def fractions(self) -> np.ndarray:
        return self.geometry[2 * self.cells:].view(np.float32)


def _grid_from(g: _VoxelGrid, nx, ny, nz) -> VoxelGrid:
    geo = np.ctypeslib.as_array(g.geometry, shape=(g.geometry_words,)).copy()
    bc = (np.ctypeslib.as_array(g.boundary_cells, shape=(g.n_boundary,)).copy()
          if g.n_boundary else np.zeros(0, np.uint32))
    out = VoxelGrid(nx, ny, nz, geo, bc, int(g.solid_cells), int(g.frontal_cells),
                    float(g.blockage), bool(g.watertight), int(g.missed_links))
    lib().pwo_voxel_grid_free(C.byref(g))
    return out


def voxelize(body, supports, layout, floor: bool, link_fractions: bool = True) -> VoxelGrid:
    nx, ny, nz = layout
    b = np.ascontiguousarray(body, dtype=np.float32).reshape(-1)
    s = None if supports is None else np.ascontiguousarray(supports, np.float32).reshape(-1)
    g = _VoxelGrid()
    err = C.create_string_buffer(256)
    rc = lib().pwo_voxelize(_ptr(b, _f), b.size // 9, _ptr(s, _f), 0 if s is None else s.size // 9,
                            nx, ny, nz, int(floor), int(link_fractions), C.byref(g), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return _grid_from(g, nx, ny, nz)


def build_wall_links(flags, body, supports, layout, floor, measure=True) -> VoxelGrid:
    nx, ny, nz = layout
    fl = np.array(flags, dtype=np.uint32)
    b = np.ascontiguousarray(body, dtype=np.float32).reshape(-1)
    s = None if supports is None else np.ascontiguousarray(supports, np.float32).reshape(-1)
    g = _VoxelGrid()
    lib().pwo_build_wall_links(_ptr(fl, _u32), _ptr(b, _f), b.size // 9, _ptr(s, _f),
                               0 if s is None else s.size // 9, nx, ny, nz, int(floor),
                               int(measure), C.byref(g))
    return _grid_from(g, nx, ny, nz)


def nearest_link_hit(mesh, layout, origin, direction):
    m = np.ascontiguousarray(mesh, dtype=np.float32).reshape(-1)
    o = np.asarray(origin, np.float64)
    d = np.asarray(direction, np.float64)
    t = C.c_double()
    hit = lib().pwo_nearest_link_hit(_ptr(m, _f), m.size // 9, *layout, _ptr(o, _d),
                                     _ptr(d, _d), C.byref(t))
    return t.value if hit else None


PRESETS = {"sphere": 0, "ahmed-25": 1, "cube": 2, "ahmed-stilts": 3}


def preset_mesh(name: str) -> np.ndarray:
    kind = PRESETS[name]
    n = lib().pwo_preset_mesh(kind, None)
    out = np.zeros(9 * n, np.float32)
    lib().pwo_preset_mesh(kind, _ptr(out, _f))
    return out


def place_mesh(mesh, supports, length_cells, layout, floor):
    m = np.ascontiguousarray(mesh, np.float32).reshape(-1)
    s = None if supports is None else np.ascontiguousarray(supports, np.float32).reshape(-1)
    ob = np.zeros_like(m)
    os_ = None if s is None else np.zeros_like(s)
    rc = lib().pwo_place_mesh(_ptr(m, _f), m.size // 9, _ptr(s, _f), 0 if s is None else s.size // 9,
                              float(length_cells), *layout, int(floor), _ptr(ob, _f), _ptr(os_, _f))
    if rc != 0:
        raise ValueError("Mesh placement failed.")
    return ob, os_


# ---------------------------------------------------------------- particles
def particle_hash(v: int) -> int:
    return lib().pwo_particle_hash(v & 0xFFFFFFFF)


def particle_needs_reseed(position, age, layout, flag) -> bool:
    p = np.asarray(position, np.float64)
    return bool(lib().pwo_particle_needs_reseed(_ptr(p, _d), float(age), *layout, int(flag) & 0xFFFFFFFF))


def seed_region(bmin, bmax, layout):
    out = np.zeros(4)
    rc = lib().pwo_seed_region(_ptr(np.asarray(bmin, np.float64), _d),
                               _ptr(np.asarray(bmax, np.float64), _d), *layout, _ptr(out, _d))
    if rc != 0:
        raise ValueError("Particle seeds need a positive cross-section.")
    return out


def particle_seeds(flags, layout, bmin, bmax, count, seed):
    fl = np.ascontiguousarray(flags, np.uint32)
    data = np.zeros(8 * count, np.float32)
    fb = np.zeros(3, np.float32)
    region = np.zeros(4)
    err = C.create_string_buffer(256)
    rc = lib().pwo_particle_seeds(_ptr(fl, _u32), *layout, _ptr(np.asarray(bmin, np.float64), _d),
                                  _ptr(np.asarray(bmax, np.float64), _d), count, seed,
                                  _ptr(data, _f), _ptr(fb, _f), _ptr(region, _d), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return data, region, fb


def particles_update(particles, macro, flags, layout, seed, steps, cross_section, fallback_yz):
    p = np.array(particles, np.float32)
    m = np.ascontiguousarray(macro, np.float32).reshape(-1)
    fl = np.ascontiguousarray(flags, np.uint32)
    cs = np.asarray(cross_section, np.float32)
    fb = np.asarray(fallback_yz, np.float32)
    lib().pwo_particles_update(_ptr(p, _f), p.size // 8, _ptr(m, _f), _ptr(fl, _u32), *layout,
                               seed, float(steps), _ptr(cs, _f), _ptr(fb, _f))
    return p


# ---------------------------------------------------------------- storage encodings
def encode_f16(value_scaled: float, i: int, n: int) -> int:
    return lib().pwo_encode_f16(value_scaled, i, n)


def encode_f16_planes(g: np.ndarray) -> np.ndarray:
    """Every deviation of a (19, cells) f32 state stored as the reference stores it
    (common.wgsl:134-135): encode_f16(g * 2^15, i, n)."""
    g = np.ascontiguousarray(g, np.float32).reshape(19, -1)
    out = np.empty(g.shape, np.uint16)
    lib().pwo_encode_f16_planes(g.ctypes.data, g.shape[1], out.ctypes.data)
    return out


def decode_f16(bits: int) -> float:
    return lib().pwo_decode_f16(bits)


def encode_bf16(value: float, i: int, n: int) -> int:
    return lib().pwo_encode_bf16(value, i, n)


def decode_bf16(bits: int) -> float:
    return lib().pwo_decode_bf16(bits)
