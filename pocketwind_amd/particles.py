"""Streakline particles over a solver's live velocity (src/render/particles.ts:30-138).

``Particles.create(solver, body_bounds, count, seed)`` seeds inlet particles as
particleSeeds does (src/render/particle-layout.ts:66-112); ``update(steps)`` advects them on
the solver's stream by trilinear sampling of the macro field written by the last step batch.
The reference caps ``count`` at 65 536; ``max_count`` raises that cap up to 2^24.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

MAX_PARTICLES = 65536  # src/render/particle-layout.ts:9


class Particles:
    def __init__(self, handle, solver, count):
        self._h = C.c_void_p(handle)
        self.solver = solver
        self.count = count

    @staticmethod
    def create(solver, body_bounds, count: int, seed: int, max_count: int = MAX_PARTICLES):
        lo, hi = (np.asarray(v, np.float64) for v in body_bounds)
        if not isinstance(count, (int, np.integer)) or count < 1 or count > max_count:
            raise ValueError("Particle count, seed or geometry buffer is invalid.")
        if not isinstance(seed, (int, np.integer)) or seed < 0 or seed > 0xFFFFFFFF:
            raise ValueError("Particle count, seed or geometry buffer is invalid.")
        out = C.c_void_p()
        N.check(N.lib().pw_particles_create(solver._live(), int(count), int(seed),
                                            lo.ctypes.data_as(C.POINTER(C.c_double)),
                                            hi.ctypes.data_as(C.POINTER(C.c_double)), C.byref(out)))
        return Particles(out.value, solver, int(count))

    @staticmethod
    def create_in_batch(batch, tunnel: int, body_bounds, count: int, seed: int,
                        max_count: int = MAX_PARTICLES):
        """The same particles over tunnel ``tunnel`` of a TunnelBatch (its last macro field)."""
        lo, hi = (np.asarray(v, np.float64) for v in body_bounds)
        if not isinstance(count, (int, np.integer)) or count < 1 or count > max_count:
            raise ValueError("Particle count, seed or geometry buffer is invalid.")
        if not isinstance(seed, (int, np.integer)) or seed < 0 or seed > 0xFFFFFFFF:
            raise ValueError("Particle count, seed or geometry buffer is invalid.")
        out = C.c_void_p()
        N.check(N.lib().pw_particles_create_batch(batch._h, int(tunnel), int(count), int(seed),
                                                  lo.ctypes.data_as(C.POINTER(C.c_double)),
                                                  hi.ctypes.data_as(C.POINTER(C.c_double)),
                                                  C.byref(out)))
        return Particles(out.value, batch, int(count))

    @staticmethod
    def create_on_field(layout, geometry, macro_ptr: int, body_bounds, count: int, seed: int,
                        max_count: int = MAX_PARTICLES):
        """createParticles over a caller's macro buffer (src/render/particles.ts:30-138):
        ``geometry`` is the grid's host flag array, ``macro_ptr`` the device address of 4
        floats per cell. The caller finishes writing the buffer before each update."""
        lo, hi = (np.asarray(v, np.float64) for v in body_bounds)
        if not isinstance(count, (int, np.integer)) or count < 1 or count > max_count:
            raise ValueError("Particle count, seed or geometry buffer is invalid.")
        if not isinstance(seed, (int, np.integer)) or seed < 0 or seed > 0xFFFFFFFF:
            raise ValueError("Particle count, seed or geometry buffer is invalid.")
        geo = np.ascontiguousarray(geometry, np.uint32)
        out = C.c_void_p()
        N.check(N.lib().pw_particles_create_field(C.byref(N.Layout(*layout)), geo.ctypes.data_as(N._u32p),
                                                  geo.size, C.c_void_p(int(macro_ptr)), int(count),
                                                  int(seed), lo.ctypes.data_as(C.POINTER(C.c_double)),
                                                  hi.ctypes.data_as(C.POINTER(C.c_double)),
                                                  C.byref(out)))
        return Particles(out.value, None, int(count))

    def synchronize(self) -> None:
        N.check(N.lib().pw_particles_synchronize(self._h))

    def update(self, steps: float) -> None:
        with np.errstate(over="ignore"):
            finite = np.isfinite(np.float32(steps))
        if not finite or steps < 0:
            raise ValueError("Particle advection needs non-negative finite solver steps.")
        N.check(N.lib().pw_particles_update(self._h, float(steps)))

    def read(self) -> np.ndarray:
        out = np.empty((self.count, 8), np.float32)
        N.check(N.lib().pw_particles_read(self._h, out.ctypes.data, out.size))
        return out

    def write(self, data) -> None:
        arr = np.ascontiguousarray(data, np.float32)
        N.check(N.lib().pw_particles_write(self._h, arr.ctypes.data, arr.size))

    def info(self):
        region = (C.c_float * 4)()
        fallback = (C.c_float * 3)()
        N.check(N.lib().pw_particles_info(self._h, region, fallback))
        return np.array(region[:], np.float32), np.array(fallback[:], np.float32)

    def destroy(self) -> None:
        if self._h:
            N.lib().pw_particles_destroy(self._h)
            self._h = None
