"""f16 storage with the reference's rounding rule (prelude-f16.wgsl:7-22): the device's stored
bits against the oracle's restatement of encode_population, and the reference's f16 bars.

Bits: an f32 solver run on the same grid computes the same f32 values the f16 solver rounds
(the same step code, unscaled; the f16 step's scale is a power of two). The initial states
must satisfy f16 words == encode_f16(f32 values * 2^15, i, n) at every cell and direction; after
k steps the f16 state is decoded (exactly) into the f32 solver, both step once, and the same
identity must hold again. Run on the per-cell path (the spec's 32^3 blob, TRT, halfway and
interpolated walls) and on the row-wavefront path at the benchmark's instance (64 x 32 x 32
sphere tunnel, REG + LES, non-temporal stores, 128 threads, 16 chunks per XCD).

Bars: the spec's storage comparison (gpu-vs-cpu.spec.ts:529-641: TRT, bounce-back, the seeded
32^3 blob, step 100) with this rounding holds the reference's own 2e-3 on g, rho - 1 and u for
every cache policy (the rule draws no per-step seed, so the result is one deterministic draw).
"""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle
from grids import blob_grid, grid_from_kinds, link_fractions
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from test_gpu_row_parity import sphere_tunnel_kinds
from test_gpu_solver import CHECKPOINTS, TAU0, U_IN, normalized_peak, oracle_states

pytestmark = pytest.mark.gpu


def decode(words: np.ndarray) -> np.ndarray:
    return words.view(np.float16).astype(np.float32) / np.float32(32768)


def differing(f16: GpuSolver, f32: GpuSolver) -> int:
    words = f16.read_populations().view(np.uint16)
    expect = oracle.encode_f16_planes(f32.read_populations())
    return int(np.count_nonzero(words != expect))


CASES = {
    "cell/trt/bb": ("blob32", "trt", False, "bb", None),
    "cell/trt/bouzidi": ("blob32", "trt", False, "bouzidi", None),
    "row/reg+les/bb": ("sphere64", "reg", True, "bb", ("cell-nt-stores", 128, 16)),
    "row/trt/bouzidi": ("sphere64", "trt", False, "bouzidi", ("cell-nt-stores", 128, 16)),
}


def case_grid(shape: str, walls: str):
    if shape == "blob32":
        return blob_grid(walls)
    grid = grid_from_kinds((64, 32, 32), sphere_tunnel_kinds((64, 32, 32)))
    return link_fractions(grid) if walls == "bouzidi" else grid


@pytest.mark.parametrize("case", list(CASES))
def test_stored_bits_equal_the_reference_encoding(case):
    shape, operator, les, walls, variant = CASES[case]
    grid = case_grid(shape, walls)
    # One slip floor for both (the storages' defaults differ; the bits compare one arithmetic).
    f16 = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, les, operator=operator, walls=walls,
                                                           precision="f16", slip_floor=1e-4))
    f32 = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, les, operator=operator, walls=walls,
                                                           slip_floor=1e-4))
    try:
        for s in (f16, f32):
            if variant:
                s.set_variant(*variant)
        f16.set_rounding("reference")
        print(case, "initial state", flush=True)
        assert differing(f16, f32) == 0
        for k in (1, 9, 40):
            f16.step(k, True)
            f32.write_populations(decode(f16.read_populations().view(np.uint16)))
            f16.step(1, True)
            f32.step(1, True)
            bad = differing(f16, f32)
            print(case, "after step", f16.step_count, "differing words:", bad, flush=True)
            assert bad == 0
    finally:
        f16.destroy()
        f32.destroy()


@pytest.mark.parametrize("kernel", ["cell", "cell-nt", "cell-nt-stores"])
def test_reference_rounding_holds_the_reference_f16_bars(kernel):
    grid = blob_grid("bb")
    ref = oracle_states(grid, "trt", False, "bb", CHECKPOINTS)
    solver = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, False, precision="f16"))
    try:
        solver.set_kernel(kernel)
        solver.set_rounding("reference")
        solver.step(100, True)
        macro = solver.read_macro().astype(np.float64)
        macro[:, 0] -= 1
        ref_macro, ref_pops, _ = ref[100]
        record = {
            "g": normalized_peak(solver.read_populations_f64(), ref_pops),
            "rho-1": normalized_peak(macro[:, 0], ref_macro[:, 0]),
            "u": normalized_peak(macro[:, 1:], ref_macro[:, 1:]),
        }
        print(kernel, json.dumps({k: f"{v:.3e}" for k, v in record.items()}))
        for name, err in record.items():
            assert err < 2e-3, (kernel, name, record)
    finally:
        solver.destroy()


def test_rounding_option_is_for_f16_only():
    grid = blob_grid("bb")
    solver = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, False))
    try:
        with pytest.raises(ValueError, match="f16"):
            solver.set_rounding("reference")
    finally:
        solver.destroy()
