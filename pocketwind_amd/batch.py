"""TunnelBatch: many independent tunnels advanced by one launch per step.

Each tunnel follows the GpuSolver semantics (src/lbm/gpu-solver.ts) with its own geometry,
relaxation time and inlet speed; operator, walls, LES, storage and wall model are shared (the
learned closure's network too, ``set_closure``). Even/odd step pairs replay a captured hipGraph. There is no reference counterpart: the reference runs one
solver at a time (src/app/solver-slot.ts:14-44).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .gpu_solver import FORCE_EVERY_STEPS, SolverParams, _pack
from .lattice import VoxelGrid


class TunnelBatch:
    def __init__(self, handle, grids, params):
        self._h = C.c_void_p(handle)
        self.grids = grids
        self.params = params

    @staticmethod
    def create(grids: list[VoxelGrid], tau0, u_in, shared: SolverParams) -> "TunnelBatch":
        n = len(grids)
        if n == 0 or len(tau0) != n or len(u_in) != n:
            raise ValueError("A batch needs one tau0 and one u_in per tunnel.")
        geos = [np.ascontiguousarray(g.geometry, np.uint32) for g in grids]
        bcs = [np.ascontiguousarray(g.boundary_cells, np.uint32) for g in grids]
        lays = (N.Layout * n)(*[N.Layout(g.nx, g.ny, g.nz) for g in grids])
        gp = (C.c_void_p * n)(*[a.ctypes.data for a in geos])
        gw = (C.c_size_t * n)(*[a.size for a in geos])
        bp = (C.c_void_p * n)(*[a.ctypes.data if a.size else None for a in bcs])
        bn = (C.c_size_t * n)(*[a.size for a in bcs])
        taus = (C.c_double * n)(*[float(t) for t in tau0])
        us = (C.c_double * n)(*[float(u) for u in u_in])
        prm = _pack(shared)
        out = C.c_void_p()
        N.check(N.lib().pw_batch_create(n, lays, gp, gw, bp, bn, taus, us, C.byref(prm), C.byref(out)))
        return TunnelBatch(out.value, grids, shared)

    def step(self, count: int, write_macro_on_last: bool = False) -> None:
        N.check(N.lib().pw_batch_step(self._h, int(count), int(bool(write_macro_on_last))))

    @property
    def step_count(self) -> int:
        return int(N.lib().pw_batch_step_count(self._h))

    @property
    def cells(self) -> int:
        return int(N.lib().pw_batch_cells(self._h))

    def step_sampled(self, count: int, every: int = FORCE_EVERY_STEPS) -> np.ndarray:
        """(tunnels, samples, 4) array of (step, fx, fy, fz)."""
        first = self.step_count + 1
        m = sum(1 for k in range(first, first + count) if k % every == 0)
        out = np.zeros((len(self.grids), max(m, 1), 4), np.float64)
        N.check(N.lib().pw_batch_step_sampled(self._h, int(count), int(every), out.ctypes.data, out.size))
        return out[:, :m]

    def read_macro(self, tunnel: int) -> np.ndarray:
        g = self.grids[tunnel]
        out = np.empty((g.nx * g.ny * g.nz, 4), np.float32)
        N.check(N.lib().pw_batch_read_macro(self._h, int(tunnel), out.ctypes.data, out.size))
        return out

    @property
    def variant(self) -> str:
        """The step variant chosen at create: cache policy / chunks per XCD / block [/ persistent grid]."""
        chunks = C.c_uint32()
        code = int(N.lib().pw_batch_step_variant(self._h, C.byref(chunks)))
        persistent = int(N.lib().pw_batch_persistent_grid(self._h))
        block = int(N.lib().pw_batch_block_size(self._h))
        name = f"{ {1: 'cell', 2: 'cell-nt', 3: 'cell-nt-stores'}.get(code, '?') }/xcd-chunks-{chunks.value}/block-{block}"
        return name + (f"/persistent-{persistent}" if persistent else "")

    def read_populations(self, tunnel: int) -> np.ndarray:
        g = self.grids[tunnel]
        dt = {"f32": np.float32, "f16": np.float16, "bf16": np.uint16}[self.params.precision]
        out = np.empty((19, g.nx * g.ny * g.nz), dt)
        N.check(N.lib().pw_batch_read_populations(self._h, int(tunnel), out.ctypes.data, out.nbytes))
        return out

    def set_closure(self, net, norm) -> None:
        """Learned wall closure for every tunnel (the batch needs wall_model "learned")."""
        norm = np.ascontiguousarray(norm, np.float32)
        N.check(N.lib().pw_batch_set_closure(self._h, net._h, norm.ctypes.data, norm.size))

    def force(self) -> np.ndarray:
        """(tunnels, 3) momentum-exchange forces of the current state (waits for queued work)."""
        out = np.empty((len(self.grids), 3), np.float64)
        N.check(N.lib().pw_batch_force(self._h, out.ctypes.data, out.size))
        return out

    def set_closure_interval(self, every: int) -> None:
        N.check(N.lib().pw_batch_set_closure_interval(self._h, int(every)))

    def set_closure_placement(self, placement: str) -> None:
        """"kernel" (the default) or "step": GpuSolver.set_closure_placement for every row tunnel."""
        placements = ("kernel", "step")
        if placement not in placements:
            raise ValueError(f"Unknown closure placement {placement!r}.")
        N.check(N.lib().pw_batch_set_closure_placement(self._h, placements.index(placement)))

    def time_steps(self, count: int) -> float:
        ms = C.c_float()
        N.check(N.lib().pw_batch_time_steps(self._h, int(count), C.byref(ms)))
        return float(ms.value)

    def destroy(self) -> None:
        if self._h:
            N.lib().pw_batch_destroy(self._h)
            self._h = None
