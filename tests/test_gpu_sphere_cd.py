"""Sphere at Re 100 against the reference's own FP32 GPU drag coefficients.

Case (src/validation/cases.ts:300-338, src/app/case-run.ts:121-250, runner.ts:234-261): the
24 x 16 lat-long sphere preset, diameter D = nx / 8 cells, nose at 0.25 nx, centred, no floor,
periodic y/z, EQ inlet and pressure outlet, halfway bounce-back, no LES, u_in = 0.05,
nu = 0.05 D / 100, six flow-throughs of ceil(nx / 0.05) steps, a force sample every 8 steps,
Cd = mean fx over the last two flow-throughs / (1/2 u^2 pi (D/2)^2). The reference values
come from public/validation/results.json (tests/golden/reference_results.json). Bar: within
0.5 % of the reference's FP32 result.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

from pocketwind_amd import bodies, drag
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_results.json")))
ROWS = {r["id"]: r for r in GOLDEN["laminar"]}
TIERS = {"S": (128, 64, 64), "M": (192, 96, 96), "L": (256, 128, 128)}


def sphere_cd(tier: str, operator: str, precision: str = "f32"):
    layout = TIERS[tier]
    d = layout[0] / 8
    body, _ = bodies.place_mesh(bodies.sphere_mesh(), None, d, layout, False)
    grid = voxelize(body, None, layout, False, link_fractions=False)
    nu = 0.05 * d / 100
    solver = GpuSolver.create(layout, grid, SolverParams(3 * nu + 0.5, 0.05, False, operator=operator,
                                                         precision=precision))
    try:
        flow = drag.flow_through_steps(layout[0])
        total = 6 * flow
        samples = solver.step_sampled(total, 8)
        fx, fy, fz = drag.window_mean_force(samples, total, 2 * flow)
        cd = fx / (0.5 * 0.05 ** 2 * math.pi * (d / 2) ** 2)
        return cd, grid, samples
    finally:
        solver.destroy()


MEASURED = {}  # (tier, operator) -> Cd of this run, for the convergence check


@pytest.mark.parametrize("operator", ["trt", "reg"])
@pytest.mark.parametrize("tier", ["S", "M", "L"])
def test_sphere_cd_within_half_percent_of_reference(tier, operator):
    row = ROWS[f"sphere-re-100-{tier}-{operator}"]
    cd, grid, samples = sphere_cd(tier, operator)
    MEASURED[(tier, operator)] = cd
    assert grid.frontal_cells == round(row["voxelAreaM2"] / row["cellSizeM"] ** 2)
    assert row["steps"] == 6 * drag.flow_through_steps(TIERS[tier][0])
    err = cd / row["cd"] - 1
    print(f"sphere {tier} {operator}: Cd {cd:.10f} reference {row['cd']:.10f} error {100 * err:+.4f} %")
    assert abs(err) < 0.005
    assert np.all(np.isfinite(samples))


def test_sphere_cd_f16_storage_close_to_f32():
    cd32, _, _ = sphere_cd("S", "trt", "f32")
    cd16, _, _ = sphere_cd("S", "trt", "f16")
    print(f"sphere S trt: Cd f32 {cd32:.8f} f16 {cd16:.8f} ({100 * (cd16 / cd32 - 1):+.4f} %)")
    assert abs(cd16 / cd32 - 1) < 0.01


@pytest.mark.parametrize("operator", ["trt", "reg"])
def test_grid_convergence_of_the_sphere_series_matches_the_reference(operator):
    """The convergence section (results.ts:208-260, h = 1 / nx over S, M, L) from this solver's
    Cd values: monotone like the reference's, with the extrapolated Cd within 0.5 % of the
    reference's and the observed order within 0.25."""
    from pocketwind_amd.validation import grid_convergence
    samples = []
    for tier in ("S", "M", "L"):
        cd = MEASURED.get((tier, operator))
        if cd is None:
            cd = sphere_cd(tier, operator)[0]
        samples.append({"h": 1 / TIERS[tier][0], "value": cd})
    mine = grid_convergence(samples)
    ref = next(r for r in GOLDEN["convergence"] if r["case"] == "sphere-re-100" and r["operator"] == operator)
    print(f"sphere {operator}: convergence {mine} reference {ref}")
    assert mine["kind"] == ref["kind"] == "monotone"
    assert abs(mine["extrapolated"] / ref["extrapolated"] - 1) < 0.005
    assert abs(mine["order"] - ref["order"]) < 0.25
