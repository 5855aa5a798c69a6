"""Population-level parity of the row-wavefront step (nx a multiple of 64, the path every
benchmark grid takes) against the FP64 oracle, with the reference's boundaries (periodic y/z,
equilibrium inlet and outlet planes).

The seeded blob of tests/browser/gpu-vs-cpu.spec.ts:76-116 (LCG seed 0x4a3b2c1d, radius
5 + 2u around (15.5, 15.5, 15.5), EQ planes x < 2 and x >= nx - 2, one SUPPORT and one FLOOR
cell) is laid out at 64 x 32 x 32 so that one wavefront holds one 64-cell x-row; link
fractions as the spec's :57-74. tau0 0.8, u_in 0.06, checkpoints 0/1/10/50/100 with 98 steps
between 1 and 100 that write no macros, normalised peak error < 1e-5 for f32 storage
(gpu-vs-cpu.spec.ts:14-19). Every cache policy, block size and XCD chunking of the step is
run for REG + LES with bounce-back (the benchmark's instance); TRT with bounce-back and with
interpolated walls run every policy and chunking. The 16-bit storages run the benchmark's
variant at the storage bars of test_gpu_solver.py.
"""
from __future__ import annotations

import functools
import json

import numpy as np
import pytest

import oracle
from grids import blob_kinds, grid_from_kinds, link_fractions
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.lattice import CELL_TYPE_MASK, FLOOR, SOLID
from test_gpu_solver import LOW_PRECISION_BARS, normalized_peak

pytestmark = pytest.mark.gpu

LAYOUT = (64, 32, 32)
TAU0, U_IN = 0.8, 0.06
CHECKPOINTS = [0, 1, 10, 50, 100]
POLICIES = ["cell", "cell-nt", "cell-nt-stores"]


@functools.lru_cache(maxsize=None)
def row_blob(walls: str):
    grid = grid_from_kinds(LAYOUT, blob_kinds(LAYOUT))
    return link_fractions(grid) if walls == "bouzidi" else grid


@functools.lru_cache(maxsize=None)
def reference(operator: str, les: bool, walls: str):
    grid = row_blob(walls)
    cells = grid.cells
    cpu = oracle.Solver(grid.layout, TAU0, {"trt": oracle.TRT, "reg": oracle.REG}[operator], les,
                        {"bb": oracle.BB, "bouzidi": oracle.BOUZIDI}[walls],
                        flags=grid.geometry if walls == "bouzidi" else grid.geometry[:cells],
                        equilibrium_state=(1.0, (U_IN, 0.0, 0.0)), threads=8)
    done, out = 0, {}
    for step in CHECKPOINTS:
        cpu.step(step - done)
        done = step
        macro = cpu.macro_all()
        macro[:, 0] -= 1
        out[step] = (macro, cpu.populations.copy().reshape(19, cells))
    return out


def errors(grid, macro, pops, ref):
    """Normalised peak errors over every cell, walls included (their macro is (1, 0, 0, 0) on
    both sides)."""
    ref_macro, ref_pops = ref
    shifted = macro.astype(np.float64)
    shifted[:, 0] -= 1
    return {
        "g": normalized_peak(pops, ref_pops),
        "rho-1": normalized_peak(shifted[:, 0], ref_macro[:, 0]),
        "u": normalized_peak(shifted[:, 1:], ref_macro[:, 1:]),
    }


def run_f32(operator, les, walls, kernel, block, chunks):
    grid = row_blob(walls)
    ref = reference(operator, les, walls)
    params = SolverParams(TAU0, U_IN, les, operator=operator, walls=walls)
    label = f"{operator}{'+les' if les else ''}/{walls}/{kernel}/{block}/{chunks}"
    print(label, flush=True)
    solver = GpuSolver.create(grid.layout, grid, params)
    diag = GpuSolver.create(grid.layout, grid, params)
    try:
        for s in (solver, diag):
            s.set_variant(kernel, block, chunks)
        snaps = {}
        solver.step(0, True)
        snaps[0] = (solver.read_macro(), solver.read_populations())
        solver.step(1, True)
        snaps[1] = (solver.read_macro(), solver.read_populations())
        solver.step(98, False)
        preserved = solver.read_macro()
        solver.step(1, True)
        snaps[100] = (solver.read_macro(), solver.read_populations())
        for step in (10, 50):
            diag.step(step - diag.step_count, True)
            snaps[step] = (diag.read_macro(), diag.read_populations())
        assert np.array_equal(preserved, snaps[1][0]), "macros changed without write_macro"
        record = {step: errors(grid, *snaps[step], ref[step]) for step in CHECKPOINTS}
        print(label, json.dumps({k: {n: f"{e:.2e}" for n, e in v.items()} for k, v in record.items()}))
        for step, errs in record.items():
            for name, err in errs.items():
                assert err < 1e-5, f"{label} step {step}: {name} normalized error {err:.3e}"
        kinds = grid.flags & CELL_TYPE_MASK
        wall = (kinds >= SOLID) & (kinds <= FLOOR)
        assert np.abs(snaps[100][1][:, wall].astype(np.float64)).max() == 0
    finally:
        solver.destroy()
        diag.destroy()


@pytest.mark.parametrize("chunks", [1, 16])
@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("kernel", POLICIES)
def test_row_path_reg_les_matches_oracle(kernel, block, chunks):
    run_f32("reg", True, "bb", kernel, block, chunks)


@pytest.mark.parametrize("chunks", [1, 16])
@pytest.mark.parametrize("kernel", POLICIES)
@pytest.mark.parametrize("walls", ["bb", "bouzidi"])
def test_row_path_trt_matches_oracle(walls, kernel, chunks):
    run_f32("trt", False, walls, kernel, 256, chunks)


@pytest.mark.parametrize("precision", ["f16", "bf16"])
def test_row_path_low_precision_storage_at_the_benchmark_variant(precision):
    """REG + LES with bounce-back, the benchmark's step variant (non-temporal stores, 128
    threads, 16 chunks per XCD), 16-bit storage against the FP64 oracle at step 100."""
    grid = row_blob("bb")
    ref = reference("reg", True, "bb")
    solver = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, True, operator="reg",
                                                              precision=precision))
    try:
        solver.set_variant("cell-nt-stores", 128, 16)
        solver.step(100, True)
        record = errors(grid, solver.read_macro(), solver.read_populations_f64(), ref[100])
        print(precision, json.dumps({k: f"{v:.2e}" for k, v in record.items()}))
        for name, bar in LOW_PRECISION_BARS[precision].items():
            assert record[name] < bar, (precision, name, record)
    finally:
        solver.destroy()
