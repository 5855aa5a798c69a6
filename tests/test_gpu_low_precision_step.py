"""One step of the 16-bit storages with the library's own rounding against the f32 step.

The f16 and bf16 storages step REG through the regrouped collision (no per-direction
equilibrium), f32 through the equilibrium form; both are collide-reg.wgsl:3-18 and differ by f32
roundings only. From one developed state (a 16-bit solver's state after some steps, decoded
exactly into an f32 solver), both solvers step once: every stored 16-bit value must be one of the
two representable neighbours of the f32 step's value (stochastic rounding picks one), to within
32 f32 ulps of the cell's largest population deviation for the arithmetic's own difference. A
wrong coefficient or sign in the regrouped collision moves populations by far more. Cases: the benchmark's
instance (REG + LES with the wall law, row path, the 64 x 32 x 32 sphere tunnel) and the per-cell
path (REG + LES on the spec's 32^3 blob, whose SUPPORT/FLOOR cells and nearly enclosed wall cell
exercise the wall law's fallbacks)."""
from __future__ import annotations

import numpy as np
import pytest

from grids import blob_grid, grid_from_kinds
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from test_gpu_row_parity import sphere_tunnel_kinds

pytestmark = pytest.mark.gpu

TAU0, U_IN = 0.8, 0.06


def decode(pops: np.ndarray, precision: str) -> np.ndarray:
    if precision == "f16":
        return pops.astype(np.float64) / 32768.0
    return (pops.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def spacing(values: np.ndarray, precision: str) -> np.ndarray:
    """Gap between consecutive representable stored values around each value (in g units)."""
    if precision == "f16":
        scaled = np.abs(values * 32768.0).astype(np.float16)
        return np.spacing(scaled).astype(np.float64) / 32768.0
    bits = np.abs(values).astype(np.float32).view(np.uint32) & np.uint32(0x7F800000)
    exponent = bits.view(np.float32).astype(np.float64)  # 2^e of the value
    return np.maximum(exponent * 2.0 ** -7, 2.0 ** -133)


def grid_of(case: str):
    if case == "row-sphere":
        return grid_from_kinds((64, 32, 32), sphere_tunnel_kinds((64, 32, 32))), ("cell-nt-stores", 128, 16)
    return blob_grid("bb"), None


@pytest.mark.parametrize("precision", ["f16", "bf16"])
@pytest.mark.parametrize("case", ["row-sphere", "cell-blob"])
def test_one_step_stores_a_neighbour_of_the_f32_step(case, precision):
    grid, variant = grid_of(case)
    # One slip floor for both (the storages' defaults differ).
    low = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, True, operator="reg", precision=precision,
                                                           slip_floor=1e-4))
    f32 = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, True, operator="reg", slip_floor=1e-4))
    try:
        for s in (low, f32):
            if variant:
                s.set_variant(*variant)
        worst = 0.0
        for k in (1, 9, 40):
            low.step(k, True)
            state = decode(low.read_populations(), precision)
            f32.write_populations(state.astype(np.float32))
            low.step(1, True)
            f32.step(1, True)
            got = decode(low.read_populations(), precision)
            want = f32.read_populations().astype(np.float64)
            gap = spacing(want, precision)
            # The two forms' own difference: a few f32 roundings of the terms a cell's populations
            # are summed from (its density deviation, rho u u and Pi, each weighted), which can
            # cancel to a far smaller value, so the allowance scales with the cell's largest
            # population deviation rather than with the value itself. A wrong coefficient or sign
            # moves a value by the size of its terms, thousands of times this allowance.
            scale = np.abs(want).max(axis=0, keepdims=True)
            arith = 32 * np.finfo(np.float32).eps * scale
            excess = np.abs(got - want) - (gap + arith)
            worst = max(worst, float((np.abs(got - want) / (gap + arith)).max()))
            bad = np.argwhere(excess > 0)
            print(case, precision, "after step", low.step_count, "values beyond one stored ulp:", len(bad),
                  f"largest |diff| / allowance {worst:.3f}", flush=True)
            for i, n in bad[:5]:
                print("   direction", i, "cell", n, "flag", hex(int(grid.flags[n])), "f32", want[i, n], "stored", got[i, n],
                      "ulp", gap[i, n], "cell scale", scale[0, n], flush=True)
            assert len(bad) == 0
    finally:
        low.destroy()
        f32.destroy()
