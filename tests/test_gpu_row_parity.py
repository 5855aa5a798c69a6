"""Population-level parity of the row-wavefront step (nx a multiple of 64, the path every
benchmark grid takes) against the FP64 oracle, with the reference's boundaries (periodic y/z,
equilibrium inlet and outlet planes).

TRT (bounce-back and interpolated walls): the seeded blob generator of
tests/browser/gpu-vs-cpu.spec.ts:76-116 (LCG seed 0x4a3b2c1d, radius 5 + 2u around
(15.5, 15.5, 15.5), EQ planes x < 2 and x >= nx - 2, one SUPPORT and one FLOOR cell) laid out
at 64 x 32 x 32, so that one wavefront holds one 64-cell x-row; link fractions as the spec's
:57-74. REG + LES with the wall law (the benchmark's operator): a sphere of diameter 10 at
(20, 15.3, 16.4) in the same tunnel with the spec's SUPPORT and FLOOR cells. The 64-long blob
has a nearly enclosed wall cell where the wall law's fallbacks (a discontinuous switch between
the prescribed-stress and the bulk collision) amplify a 1e-7 perturbation of the FP64 oracle's
own state to 6e-6 by step 100, so no f32 solver can hold 1e-5 there; on the sphere the same
perturbation decays (test_oracle_conditioning.py shows both).

tau0 0.8, u_in 0.06, checkpoints 0/1/10/50/100 with 98 steps between 1 and 100 that write no
macros, normalised peak error < 1e-5 for f32 storage (gpu-vs-cpu.spec.ts:14-19) on the
populations g, on u and, for TRT, on rho - 1. The macro buffer stores rho itself as f32, whose
spacing near 1 (1.19e-7) is about 1e-5 of the largest |rho - 1| of these flows (1.3e-2), so the
spec's bar on rho - 1 is a one-ulp test: the per-cell path meets it on the spec's own 32^3 blob
with 8.1e-6, and on the sphere both paths reach 1.45e-5 at step 100 (2.3 ulp, at the first
fluid cells behind the inlet plane, identically on the row path at nx = 64 and the cell path
at nx = 63). For REG + LES the density is therefore held at three f32 ulps of the oracle's at
every cell. Every cache policy, block size and XCD chunking runs REG + LES; TRT runs every
policy and chunking. The 16-bit storages run the spec's storage comparison (TRT, bounce-back, the blob) at
the benchmark's variant, at the storage bars of test_gpu_solver.py.

The row path is also compared with the per-cell path on the same grid (the step forced to one
cell per lane): both run the same collision code, so their states must agree bit for bit, in
every storage format. The per-cell path is the one the spec's 32^3 blob pins
(test_gpu_solver.py).
"""
from __future__ import annotations

import functools
import json

import numpy as np
import pytest

import oracle
from grids import blob_kinds, grid_from_kinds, link_fractions, sphere_kinds
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.lattice import CELL_TYPE_MASK, FLOOR, SOLID, SUPPORT
from test_gpu_solver import LOW_PRECISION_BARS, normalized_peak

pytestmark = pytest.mark.gpu

LAYOUT = (64, 32, 32)
TAU0, U_IN = 0.8, 0.06
CHECKPOINTS = [0, 1, 10, 50, 100]
POLICIES = ["cell", "cell-nt", "cell-nt-stores"]


def sphere_tunnel_kinds(layout=LAYOUT):
    """Sphere of diameter 10 at (20, 15.3, 16.4), EQ planes x = 0 and nx - 1, and the spec's
    SUPPORT and FLOOR cells at (15, 15, 4) and (16, 15, 4)."""
    nx, ny = layout[0], layout[1]
    kinds = sphere_kinds(layout, (20, 15.3, 16.4), 10).copy()
    kinds[15 + nx * (15 + ny * 4)] = SUPPORT
    kinds[16 + nx * (15 + ny * 4)] = FLOOR
    return kinds


@functools.lru_cache(maxsize=None)
def row_grid(shape: str, walls: str):
    kinds = blob_kinds(LAYOUT) if shape == "blob" else sphere_tunnel_kinds()
    grid = grid_from_kinds(LAYOUT, kinds)
    return link_fractions(grid) if walls == "bouzidi" else grid


def shape_of(operator: str, les: bool) -> str:
    return "sphere" if les else "blob"


@functools.lru_cache(maxsize=None)
def reference(operator: str, les: bool, walls: str):
    grid = row_grid(shape_of(operator, les), walls)
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
    grid = row_grid(shape_of(operator, les), walls)
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
        held = ("g", "u") if les else ("g", "rho-1", "u")
        for step, errs in record.items():
            for name in held:
                assert errs[name] < 1e-5, f"{label} step {step}: {name} normalized error {errs[name]:.3e}"
            if les:
                rho = snaps[step][0][:, 0].astype(np.float64)
                ref_rho = ref[step][0][:, 0] + 1
                ulps = np.abs(rho - ref_rho) / np.spacing(np.float32(ref_rho)).astype(np.float64)
                assert ulps.max() <= 3, f"{label} step {step}: density {ulps.max():.2f} f32 ulps off"
        kinds = grid.flags & CELL_TYPE_MASK
        wall = (kinds >= SOLID) & (kinds <= FLOOR)
        assert np.abs(snaps[100][1][:, wall].astype(np.float64)).max() == 0
    finally:
        solver.destroy()
        diag.destroy()


@pytest.mark.parametrize("chunks", [1, 16, 256])
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
    """The spec's storage comparison (gpu-vs-cpu.spec.ts:529-641: TRT with bounce-back on the
    seeded blob, step 100) on the row path at the benchmark's step variant (non-temporal stores,
    128 threads, 16 chunks per XCD), 16-bit storage against the FP64 oracle."""
    grid = row_grid("blob", "bb")
    ref = reference("trt", False, "bb")
    solver = GpuSolver.create(grid.layout, grid, SolverParams(TAU0, U_IN, False, operator="trt",
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


@pytest.mark.parametrize("precision", ["f32", "f16", "bf16"])
def test_row_path_equals_the_cell_path(precision):
    """The benchmark's instance (REG + LES, bounce-back, non-temporal stores, 128 threads, 16
    chunks per XCD) on the sphere tunnel: the row-wavefront step and the same step forced to one
    cell per lane give bit-identical populations and macros at steps 1, 10 and 100."""
    grid = row_grid("sphere", "bb")
    params = SolverParams(TAU0, U_IN, True, operator="reg", precision=precision)
    row = GpuSolver.create(grid.layout, grid, params)
    cell = GpuSolver.create(grid.layout, grid, params)
    try:
        for s in (row, cell):
            s.set_variant("cell-nt-stores", 128, 16)
        cell.force_cell_path(True)
        for step in (1, 10, 100):
            row.step(step - row.step_count, True)
            cell.step(step - cell.step_count, True)
            a, b = row.read_populations(), cell.read_populations()
            differ = int(np.count_nonzero(a.view(np.uint16 if a.itemsize == 2 else np.uint32)
                                          != b.view(np.uint16 if b.itemsize == 2 else np.uint32)))
            print(precision, step, "differing population words:", differ, flush=True)
            assert differ == 0
            assert np.array_equal(row.read_macro().view(np.uint32), cell.read_macro().view(np.uint32))
    finally:
        row.destroy()
        cell.destroy()
