"""GpuSolver: the reference's solver step/readback API over libpocketwind's C-ABI.

Mirrors src/lbm/gpu-solver.ts:19-465 and src/lbm/gpu-force.ts:11-216 -- the same method
names (snake_case), argument meaning and error behaviour:

* invalid input raises ``ValueError`` (the reference's RangeError);
* an allocation failure is *returned* as :class:`AllocationFailure` with nothing left
  allocated (gpu-solver.ts:330-334);
* ``read_force()`` returns ``'busy'`` when no sample is encoded (gpu-force.ts:192-205).

Device work is queued on the solver's own HIP stream. Readbacks wait for it.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Literal, Optional, Union

import numpy as np

from . import _native as N
from .lattice import CELL_TYPE_MASK, FLUID, Layout, VoxelGrid

FORCE_EVERY_STEPS = 8  # src/lbm/params.ts:12

_OPERATORS = {"trt": 0, "reg": 1}
_WALLS = {"bb": 0, "bouzidi": 1}
_STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
_WALL_MODELS = {"law": 0, "learned": 1}
_SIDES = {"periodic": 0, "free-slip": 1}
_OUTLETS = {"equilibrium": 0, "zero-gradient": 1}

NAN_GUARD_INTERVAL = 200  # nan-guard.ts:6


def is_unstable_density(rho: float) -> bool:
    """nan-guard.ts: non-finite or outside [0.5, 1.5] (both endpoints are stable)."""
    return not math.isfinite(rho) or rho < 0.5 or rho > 1.5


def unstable_message(step: int) -> dict:
    """The 'unstable' event detail (nan-guard.ts resetUnstableRun, app/messages.ts)."""
    return {"step": step, "what": f"The run went unstable at step {step} and was reset",
            "next": "Lower the speed, or choose a higher tier under Advanced"}


@dataclass(frozen=True)
class SolverParams:
    """SolverParams (gpu-solver.ts:19-30); ``precision`` adds bf16 to f32/f16."""

    tau0: float
    u_in: float
    les: bool
    operator: str = "trt"
    walls: str = "bb"
    precision: str = "f32"
    ramp_steps: int = 0
    wall_model: str = "law"
    # Boundary options with no reference counterpart (the reference's sides are periodic and
    # its outlet is the equilibrium outlet): "free-slip" sides, a "zero-gradient" outlet.
    sides: str = "periodic"
    outlet: str = "equilibrium"
    # Smallest slip speed the wall law acts on; 0: 1e-7 for f32 storage (the reference's), 1e-4
    # for 16-bit storage, whose velocity noise would otherwise switch stagnation cells to the law.
    slip_floor: float = 0.0


@dataclass(frozen=True)
class ForceSample:
    step: int
    fx: float
    fy: float
    fz: float


@dataclass(frozen=True)
class AllocationFailure:
    """Returned, not raised, when device memory runs out (gpu-resources.ts:3-5)."""

    message: str
    kind: str = "allocation-failure"


def _pack(params: SolverParams) -> N.Params:
    for name, table, value in (("operator", _OPERATORS, params.operator),
                               ("walls", _WALLS, params.walls),
                               ("precision", _STORAGE, params.precision),
                               ("wall_model", _WALL_MODELS, params.wall_model),
                               ("sides", _SIDES, params.sides), ("outlet", _OUTLETS, params.outlet)):
        if value not in table:
            raise ValueError(f"Unknown {name} {value!r}.")
    if not isinstance(params.ramp_steps, (int, np.integer)) or params.ramp_steps < 0:
        raise ValueError("Ramp steps must be a non-negative safe integer.")
    return N.Params(float(params.tau0), float(params.u_in), int(bool(params.les)),
                    _OPERATORS[params.operator], _WALLS[params.walls], _STORAGE[params.precision],
                    int(params.ramp_steps), _WALL_MODELS[params.wall_model], _SIDES[params.sides],
                    _OUTLETS[params.outlet], float(params.slip_floor))


class GpuSolver:
    """Pull-stream/collide driver with f32 arithmetic and f32, f16 or bf16 storage."""

    def __init__(self, handle: int, layout: Layout, params: SolverParams, grid: VoxelGrid):
        self._h = C.c_void_p(handle)
        self.layout = layout
        self.params = params
        self.cells = layout.nx * layout.ny * layout.nz
        self._fluid_cell = int(np.flatnonzero((grid.flags & CELL_TYPE_MASK) == FLUID)[0]) \
            if np.any((grid.flags & CELL_TYPE_MASK) == FLUID) else -1
        self._injected_nan = False
        self._destroyed = False

    # -- creation -------------------------------------------------------------------
    @staticmethod
    def create(layout, grid: VoxelGrid, params: SolverParams) -> Union["GpuSolver", AllocationFailure]:
        layout = Layout(*layout)
        if (grid.nx, grid.ny, grid.nz) != tuple(layout):
            raise ValueError("Geometry must match the solver layout and contain flags and q offsets.")
        geo = np.ascontiguousarray(grid.geometry, dtype=np.uint32)
        bc = np.ascontiguousarray(grid.boundary_cells, dtype=np.uint32)
        lay = N.Layout(*layout)
        prm = _pack(params)
        out = C.c_void_p()
        status = N.lib().pw_solver_create(
            C.byref(lay), geo.ctypes.data_as(N._u32p), geo.size, bc.ctypes.data_as(N._u32p),
            bc.size, C.byref(prm), C.byref(out))
        if status == N.PW_ENOMEM:
            return AllocationFailure(N.last_error())
        N.check(status)
        return GpuSolver(out.value, layout, params, grid)

    def _live(self):
        if self._destroyed:
            raise RuntimeError("The solver has been destroyed.")
        return self._h

    # -- stepping -------------------------------------------------------------------
    def step(self, count: int, write_macro_on_last: bool) -> None:
        """Queue `count` steps; only the last writes macros when asked (gpu-solver.ts:378-395)."""
        if not isinstance(count, (int, np.integer)) or count < 0 or count >= 2**32:
            raise ValueError("Step count must be a non-negative safe integer.")
        N.check(N.lib().pw_solver_step(self._live(), int(count), int(bool(write_macro_on_last))))

    @property
    def step_count(self) -> int:
        return int(N.lib().pw_solver_step_count(self._live()))

    def reset(self) -> None:
        N.check(N.lib().pw_solver_reset(self._live()))

    def encode_force(self) -> None:
        N.check(N.lib().pw_solver_encode_force(self._live()))

    def read_force(self) -> Union[ForceSample, Literal["busy"]]:
        out = N.ForceSample()
        status = N.check(N.lib().pw_solver_read_force(self._live(), C.byref(out)))
        if status == N.PW_EBUSY:
            return "busy"
        return ForceSample(int(out.step), out.fx, out.fy, out.fz)

    # -- readback -------------------------------------------------------------------
    @property
    def storage_dtype(self):
        return {"f32": np.float32, "f16": np.float16, "bf16": np.uint16}[self.params.precision]

    def read_populations(self) -> np.ndarray:
        """Latest state as a (19, N) array in storage format (gpu-solver.ts:338-340)."""
        out = np.empty((19, self.cells), dtype=self.storage_dtype)
        N.check(N.lib().pw_solver_read_populations(self._live(), out.ctypes.data, out.nbytes))
        return out

    @property
    def population_buffers(self):
        """The three direction groups A (0-6), B (7-12), C (13-18) of the latest state."""
        pops = self.read_populations()
        return pops[0:7], pops[7:13], pops[13:19]

    def read_populations_f64(self) -> np.ndarray:
        """Decoded deviations g = f - w of the latest state, (19, N) float64."""
        pops = self.read_populations()
        if self.params.precision == "f16":
            return pops.astype(np.float64) / 32768.0
        if self.params.precision == "bf16":
            return (pops.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        return pops.astype(np.float64)

    def write_populations(self, pops: np.ndarray) -> None:
        arr = np.ascontiguousarray(pops, dtype=self.storage_dtype)
        N.check(N.lib().pw_solver_write_populations(self._live(), arr.ctypes.data, arr.nbytes))

    def read_macro(self) -> np.ndarray:
        """macroBuffer as (N, 4) float32: rho, ux, uy, uz (docs/numerics.md)."""
        out = np.empty((self.cells, 4), dtype=np.float32)
        N.check(N.lib().pw_solver_read_macro(self._live(), out.ctypes.data, out.size))
        return out

    macro_buffer = property(read_macro)

    def inject_nan_for_runner(self, step: int) -> None:
        """gpu-solver.ts:358-375: NaN into the first fluid cell once, at the given step."""
        if self._destroyed or self._injected_nan or self._fluid_cell < 0 or step <= 0 \
                or self.step_count != step:
            return
        N.check(N.lib().pw_debug_inject_nan(self._live()))
        self._injected_nan = True

    def advance_guarded(self, count: int, inject_nan: Optional[int] = None):
        """NanGuard.advance (nan-guard.ts:151-196) with the density check fused into the step:
        steps `count` times in batches ending on multiples of NAN_GUARD_INTERVAL; a check that
        finds a density outside [0.5, 1.5] or not finite resets the solver and returns the
        instability message {step, what, next}; otherwise None. `inject_nan` plays the runner's
        inject_nan parameter: NaN goes into the first fluid cell after that step."""
        if not isinstance(count, int) or count < 0:
            raise ValueError("Step count must be a non-negative safe integer.")
        remaining = count
        while remaining > 0:
            until = NAN_GUARD_INTERVAL - self.step_count % NAN_GUARD_INTERVAL
            until_inject = (inject_nan - self.step_count) if (inject_nan and inject_nan > self.step_count) \
                else remaining
            steps = min(remaining, until, until_inject)
            at = C.c_uint64()
            N.check(N.lib().pw_solver_advance_guarded(self._live(), steps, C.byref(at)))
            remaining -= steps
            if at.value:
                self._injected_nan = False
                return unstable_message(int(at.value))
            if inject_nan:
                self.inject_nan_for_runner(inject_nan)
        return None

    def device_state(self) -> N.DeviceState:
        out = N.DeviceState()
        N.check(N.lib().pw_solver_device_state(self._live(), C.byref(out)))
        return out

    KERNELS = {"auto": 0, "cell": 1, "cell-nt": 2, "cell-nt-stores": 3}

    def set_kernel(self, kernel: str) -> None:
        """Force a step-kernel variant ('cell', 'cell-nt', 'cell-nt-stores') at block size 256,
        or re-tune variant and block size ('auto')."""
        N.check(N.lib().pw_solver_set_kernel(self._live(), self.KERNELS[kernel]))

    def force_cell_path(self, on: bool = True) -> None:
        """Test hook: one cell per lane even on row-wavefront grids (nx a multiple of 64)."""
        N.check(N.lib().pw_debug_cell_path(self._live(), int(bool(on))))

    ROUNDINGS = ("hardware", "reference")

    def set_rounding(self, rounding: str) -> None:
        """Rounding of f16 stores: "hardware" (the default) or "reference", the rule of
        prelude-f16.wgsl:7-22 (random word lowbias32 of the value's bits, cell and direction).
        Resets the solver to its initial state under the new rounding."""
        if rounding not in self.ROUNDINGS:
            raise ValueError(f"Unknown f16 rounding {rounding!r}.")
        N.check(N.lib().pw_solver_set_rounding(self._live(), self.ROUNDINGS.index(rounding)))

    def set_variant(self, kernel: str, block: int, xcd_chunks: int) -> None:
        """An explicit step variant: population cache policy, threads per block (64, 128,
        256) and chunks per XCD (a power of two up to 1024), without timing."""
        N.check(N.lib().pw_solver_set_step_variant(self._live(), self.KERNELS[kernel], block, xcd_chunks))

    @property
    def kernel(self) -> str:
        code = int(N.lib().pw_solver_kernel(self._live()))
        return {1: "cell", 2: "cell-nt", 3: "cell-nt-stores"}.get(code, "auto")

    @property
    def block_size(self) -> int:
        return int(N.lib().pw_solver_block_size(self._live()))

    @property
    def xcd_chunks(self) -> int:
        return int(N.lib().pw_solver_xcd_chunks(self._live()))

    # Learned wall closure (no reference counterpart; see include/pw.h).
    def wall_cells(self) -> np.ndarray:
        n = int(N.lib().pw_solver_wall_cell_count(self._live()))
        out = np.empty(n, np.uint32)
        N.check(N.lib().pw_solver_wall_cells(self._live(), out.ctypes.data, n))
        return out

    def closure_features(self, u_ref: float) -> np.ndarray:
        """Per wall cell: 12 closure features, then the wall law's stress vector (15 floats)."""
        n = int(N.lib().pw_solver_wall_cell_count(self._live()))
        out = np.empty((n, 15), np.float32)
        N.check(N.lib().pw_solver_closure_features(self._live(), C.c_float(u_ref), out.ctypes.data, out.size))
        return out

    def wall_forces(self) -> np.ndarray:
        n = int(N.lib().pw_solver_wall_cell_count(self._live()))
        out = np.empty((n, 3), np.float64)
        N.check(N.lib().pw_solver_wall_forces(self._live(), out.ctypes.data, out.size))
        return out

    def set_closure(self, net, norm) -> None:
        norm = np.ascontiguousarray(norm, np.float32)
        N.check(N.lib().pw_solver_set_closure(self._live(), net._h, norm.ctypes.data, norm.size))

    PLACEMENTS = ("kernel", "step", "overlapped", "alongside")

    def set_closure_placement(self, placement: str) -> None:
        """Where the closure runs: "kernel" (its own kernel, every `interval` steps; the default),
        "step" (inside the step's wall wavefronts, every step; REG + LES, nx a multiple of 64),
        "overlapped" (its kernel on a second stream beside the step that reads the same state;
        the output serves the `interval` steps after that one) or "alongside" (the same
        schedule and outputs as "overlapped", evaluated by the first blocks of that step's own
        launch; REG)."""
        if placement not in self.PLACEMENTS:
            raise ValueError(f"Unknown closure placement {placement!r}.")
        N.check(N.lib().pw_solver_set_closure_placement(self._live(), self.PLACEMENTS.index(placement)))

    def prescribe_wall_stress(self, stress, valid, source: bool = False) -> None:
        """Test hook: the wall collision uses these stress vectors (n_wall x 3, wall-cell order)
        where `valid`, the wall law elsewhere, in place of a network's output. With `source`,
        the vectors are momentum sources added to the wall cells after the wall law's
        collision instead (lattice units per step; the body's force then carries minus their
        sum, which the caller accounts for)."""
        stress = np.asarray(stress, np.float32).reshape(-1, 3)
        out = np.zeros((stress.shape[0], 4), np.float32)
        out[:, :3] = stress
        out[:, 3] = np.asarray(valid, bool) * (2.0 if source else 1.0)
        N.check(N.lib().pw_debug_set_wall_stress(self._live(), out.ctypes.data, out.size))

    def set_closure_interval(self, every: int) -> None:
        N.check(N.lib().pw_solver_set_closure_interval(self._live(), int(every)))

    def closure_output(self) -> np.ndarray:
        n = int(N.lib().pw_solver_wall_cell_count(self._live()))
        out = np.empty((n, 4), np.float32)
        N.check(N.lib().pw_solver_closure_output(self._live(), out.ctypes.data, out.size))
        return out

    def synchronize(self) -> None:
        N.check(N.lib().pw_solver_synchronize(self._live()))

    def step_sampled(self, count: int, every: int = FORCE_EVERY_STEPS) -> np.ndarray:
        """Advance `count` steps, sampling the force after every `every`-th step on the device;
        returns an (n, 4) array of (step, fx, fy, fz)."""
        first = self.step_count + 1
        n = sum(1 for k in range(first, first + count) if k % every == 0)
        out = np.zeros((max(n, 1), 4), np.float64)
        N.check(N.lib().pw_solver_step_sampled(self._live(), int(count), int(every), out.ctypes.data,
                                               out.size))
        return out[:n]

    def time_steps(self, count: int) -> float:
        """Milliseconds for `count` steps between two events on the solver's stream."""
        ms = C.c_float()
        N.check(N.lib().pw_solver_time_steps(self._live(), int(count), C.byref(ms)))
        return float(ms.value)

    def destroy(self) -> None:
        if self._destroyed:
            return
        N.lib().pw_solver_destroy(self._h)
        self._destroyed = True

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def copy_bandwidth(nbytes: int = 128 << 20, passes: int = 40) -> float:
    """Measured device copy GB/s (bench-page.ts:204-242)."""
    out = C.c_double()
    N.check(N.lib().pw_copy_bandwidth(int(nbytes), int(passes), C.byref(out)))
    return out.value
