"""Validation harness on the host: grid convergence and grid band (tests/unit/convergence.test.ts,
tests/unit/forces/grid-band.test.ts), case specifications and result rows
(tests/unit/validation/runner.test.ts), scoring (tests/unit/validation/score-results.test.ts),
each pinned also against the reference's own saved results (public/validation/results.json,
kept as tests/golden/reference_results.json): every stored row's case fields, its Cd, voxel Cd,
Maskell bound, error and pass verdict from its mean force, the convergence section from the
laminar rows, and the grid-check rows' band verdicts."""
from __future__ import annotations

import json
import math
import os
import subprocess
import sys

import pytest

from pocketwind_amd import validation as V

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REF = json.load(open(os.path.join(GOLDEN, "reference_results.json")))
STORED = [r for s in ("laminar", "validation", "confinement", "sensitivity") for r in REF[s]]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def quadratic(h):
    return 1 + 0.5 * h ** 2


# ------------------------------------------------------------------------- convergence

def test_recovers_order_two_and_the_limit_from_four_grids():
    r = V.grid_convergence([{"h": h, "value": quadratic(h)} for h in (1 / 384, 1 / 128, 1 / 256, 1 / 192)])
    assert r["kind"] == "monotone" and abs(r["order"] - 2) < 0.01 and abs(r["extrapolated"] - 1) < 1e-6


def test_safety_factor_and_unequal_ratios():
    r = V.grid_convergence([{"h": 1, "value": 1.5}, {"h": 2, "value": 3}, {"h": 4, "value": 9}])
    assert r["order"] == pytest.approx(2, abs=1e-9) and r["extrapolated"] == pytest.approx(1, abs=1e-9)
    assert r["gciFinePct"] == pytest.approx(125 / 3, abs=1e-10)
    r = V.grid_convergence([{"h": 1, "value": 1.5}, {"h": 2, "value": 3}, {"h": 6, "value": 19}])
    assert abs(r["order"] - 2) < 1e-8 and abs(r["extrapolated"] - 1) < 1e-8


def test_only_the_three_finest_levels_count():
    rows = [{"h": h, "value": quadratic(h)} for h in (1 / 8, 1 / 4, 1 / 2)] + [{"h": 1, "value": 42}]
    r = V.grid_convergence(rows)
    assert r["order"] == pytest.approx(2, abs=1e-9) and r["extrapolated"] == pytest.approx(1, abs=1e-9)


@pytest.mark.parametrize("values,kind", [((1, 2, 1), "oscillatory"), ((1, 3, 4), "divergent")])
def test_non_monotone_sequences(values, kind):
    rows = [{"h": h, "value": v} for h, v in zip((1, 2, 3), values)]
    assert V.grid_convergence(rows) == {"kind": kind}
    assert V.grid_band([{"h": r["h"], "cd": r["value"]} for r in rows]) == {"kind": "no-band", "reason": kind}


def test_insufficient_levels():
    assert V.grid_convergence([{"h": 1, "value": 1.5}, {"h": 2, "value": 3}]) == {"kind": "insufficient"}
    assert V.grid_convergence([{"h": 1, "value": 1}, {"h": 2, "value": 1}, {"h": 4, "value": 2}]) == \
        {"kind": "insufficient"}


def test_grid_band():
    band = V.grid_band([{"h": 4, "cd": 9}, {"h": 1, "cd": 1.5}, {"h": 2, "cd": 3}])
    assert band["kind"] == "band"
    assert band["extrapolatedCd"] == pytest.approx(1, abs=1e-9) and band["gciPct"] == pytest.approx(125 / 3, abs=1e-9)
    assert V.grid_band([{"h": 1, "cd": 2}, {"h": 2, "cd": 2}, {"h": 4, "cd": 2}]) == \
        {"kind": "no-band", "reason": "insufficient"}
    for rows in ([{"h": 1, "cd": math.nan}, {"h": 2, "cd": 3}, {"h": 4, "cd": 9}],
                 [{"h": 0, "cd": 1}, {"h": 2, "cd": 3}, {"h": 4, "cd": 9}]):
        assert V.grid_band(rows) == {"kind": "no-band", "reason": "insufficient"}


def test_convergence_section_of_the_reference_results():
    rows = V.convergence_rows(STORED + REF["validation"])
    got = {(r["case"], r["operator"]): r for r in rows}
    for ref in REF["convergence"]:
        mine = got[(ref["case"], ref["operator"])]
        assert mine["kind"] == ref["kind"] and mine["tiers"] == ref["tiers"]
        for key in ("order", "extrapolated", "gciFinePct"):
            if key in ref:
                assert mine[key] == pytest.approx(ref[key], rel=1e-12)


def test_grid_check_rows_of_the_reference_results():
    for check in REF["gridCheck"]:
        band = V.grid_band([{"h": 1 / r["lengthCells"], "cd": r["cd"]} for r in check["rows"]])
        assert band["kind"] == check["kind"] and band.get("reason") == check.get("reason")


def test_grid_check_lengths_steps_and_reynolds_number_of_the_reference_results():
    """The stored grid checks (the Ahmed body and the sphere from tier S at the app's defaults):
    the body lengths per tier and the Reynolds number fixed by the interactive viscosity floor."""
    from pocketwind_amd import bodies
    from pocketwind_amd.lattice import TIERS
    s = TIERS["S"]
    for check in REF["gridCheck"]:
        mesh = bodies.preset_mesh(check["case"])
        current = V.run_length_cells(check["case"], mesh, (s.nx, s.ny, s.nz))
        lengths = [V.grid_check_length_cells(current, s.nx, r["tier"]) for r in check["rows"]]
        assert lengths == [r["lengthCells"] for r in check["rows"]]
        assert V.U_LU * current / V.NU_MIN_LU == pytest.approx(check["re"], rel=1e-14)
    assert [V.grid_check_steps(t) for t in "SML"] == [(10240, 5120), (15360, 7680), (20480, 10240)]
    assert V.mean_of_last([(8, 1.0), (16, 2.0), (24, 4.0)], 16) == 3.0
    assert V.mean_of_last([(8, 1.0), (16, float("nan"))], 16) is None and V.mean_of_last([], 8) is None


# ------------------------------------------------------------------------- cases

def spec_for(row):
    """The case builder call that made a stored row."""
    g = row["group"]
    if g == "laminar":
        return V.laminar_sphere_case(row["tier"], row["operator"], row["walls"], row["storage"])
    if g == "confinement":
        return V.confinement_case(row["case"].split("-")[0], row["operator"])
    if g == "shape-class":
        return V.shape_case(row["case"], row["tier"])
    if g == "smoke":
        return V.smoke_case()
    if g == "sensitivity":
        return next(s for s in V.sensitivity_cases() if s.id == row["id"])
    return next(s for s in V.build_cases() if s.id == row["id"])


SPEC_FIELDS = ["case", "id", "group", "label", "tier", "grid", "cells", "operator", "walls", "storage", "variant",
               "sensitivity", "shapeClass", "speedMps", "re", "les", "floor", "stilts", "bodyLengthMm",
               "bodyLengthCells", "cellSizeM", "referenceAreaM2", "baselineBlockage", "uLu", "steps",
               "flowThroughSteps", "referenceCd", "isDefault"]


@pytest.mark.parametrize("row", STORED, ids=[r["id"] for r in STORED])
def test_case_specs_match_every_stored_row(row):
    spec = spec_for(row)
    for key in SPEC_FIELDS:
        assert getattr(spec, key) == row[key], key


def test_build_cases_covers_the_stored_and_skipped_rows():
    ids = {s.id for s in V.build_cases()}
    stored = {r["id"] for r in STORED if r["group"] != "smoke"}
    assert stored <= ids
    assert ids - stored == {"ahmed-25-L-operator-trt", "cube-confinement-trt"}


@pytest.mark.parametrize("row", STORED, ids=[r["id"] for r in STORED])
def test_rows_from_the_stored_mean_forces_reproduce_the_stored_verdicts(row):
    spec = spec_for(row)
    # Two equal samples in the averaging window average to the stored force exactly.
    forces = [V.ForceSample(spec.steps - 8, row["fx"], row["fy"], row["fz"]),
              V.ForceSample(spec.steps, row["fx"], row["fy"], row["fz"])]
    run = V.SolverRun(forces, spec.steps, row["msPerStep"], voxelAreaM2=row["voxelAreaM2"], blockage=row["blockage"])
    out = V.row_from_run(spec, run, V.RunContext(row["commit"], row["date"]))
    for key in ("cd", "cdVoxel", "cdRaw", "cdMaskell", "errorPct"):
        if row[key] is None:
            assert out[key] is None, key
        else:
            assert out[key] == pytest.approx(row[key], rel=1e-14, abs=0), key
    assert (out["pass"], out["reasons"]) == (row["pass"], row["reasons"])
    assert (out["fx"], out["fy"], out["fz"]) == (row["fx"], row["fy"], row["fz"])


@pytest.mark.parametrize("row", STORED, ids=[r["id"] for r in STORED])
def test_case_placement_reproduces_every_stored_voxel_area_and_blockage(row):
    """The placed preset (floor, stilts as supports or counted in the drag) voxelised by the
    oracle, which the device voxelizer matches bit for bit, covers exactly the frontal cells of
    the reference's run; the case's Reynolds number and speed are the stored ones."""
    import oracle
    spec = spec_for(row)
    body, supports = V.case_meshes(spec)
    layout = (spec.grid["nx"], spec.grid["ny"], spec.grid["nz"])
    grid = oracle.voxelize(body, supports, layout, spec.floor, link_fractions=False)
    assert grid.frontal_cells * spec.cellSizeM ** 2 == row["voxelAreaM2"]
    assert grid.frontal_cells / (layout[1] * layout[2]) == row["blockage"]
    assert (supports is None) == (spec.shape != "ahmed-25" or row["id"] == "sensitivity-stilts-included")
    re, speed, _ = V.case_scales(spec)
    assert re == pytest.approx(row["re"] if row["re"] is not None else 40 * row["bodyLengthMm"] / 1000 / 1.5e-5,
                               rel=1e-15)
    assert speed == row["speedMps"]


# ------------------------------------------------------------------------- runner unit cases

CONTEXT = V.RunContext("abc1234", "2026-11-06T00:00:00.000Z", {"vendor": "test"}, "test-agent")
RUN = V.SolverRun([V.ForceSample(8, 1, 1, 0.25), V.ForceSample(10, 2, 0.5, 0.25), V.ForceSample(12, 3, -0.5, 0.25)],
                  12, 2.5)


def short_spec(**patch):
    from dataclasses import replace
    base = V.ahmed_case("S", "reg")
    return replace(base, **{"steps": 12, "flowThroughSteps": 2, "averageSteps": 4, "referenceCd": 0.025, **patch})


def test_runner_averages_the_window_and_converts_the_force():
    spec = short_spec()
    row = V.row_from_run(spec, RUN, CONTEXT)
    assert row["cd"] == pytest.approx(2.5 / (0.5 * 1.204 * 40 * 40 * 0.112), abs=1e-12)
    assert row["cdVoxel"] == pytest.approx(2.5 / (0.5 * 1.204 * 40 * 40 * spec.voxelAreaM2), abs=1e-12)
    assert (row["fx"], row["fy"], row["fz"], row["steps"], row["msPerStep"]) == (2.5, 0, 0.25, 12, 2.5)


def test_runner_fluctuation_drift_and_history():
    row = V.row_from_run(short_spec(flowThroughSteps=6), RUN, CONTEXT)
    assert row["relStdPct"] == pytest.approx(50, abs=1e-12) and row["driftPct"] == pytest.approx(75, abs=1e-12)
    assert row["forceHistory"] == [[8, 1, 1, 0.25]]


def test_runner_pass_band_and_rejections():
    assert V.row_from_run(short_spec(), RUN, CONTEXT)["pass"]
    failed = V.row_from_run(short_spec(referenceCd=0.02), RUN, CONTEXT)
    assert not failed["pass"] and "exceeds 10%" in " ".join(failed["reasons"])
    with pytest.raises(ValueError, match="at least 12"):
        V.row_from_run(short_spec(), V.SolverRun(RUN.forces, 11, 2.5), CONTEXT)
    with pytest.raises(ValueError, match="averaging window"):
        V.row_from_run(short_spec(), V.SolverRun(RUN.forces[:1], 12, 2.5), CONTEXT)
    with pytest.raises(ValueError, match="non-finite"):
        V.row_from_run(short_spec(), V.SolverRun([V.ForceSample(12, math.nan, 0, 0)], 12, 2.5), CONTEXT)


def test_runner_maskell_for_cube_rows():
    from dataclasses import replace
    spec = replace(V.shape_case("cube", "S"), steps=12, flowThroughSteps=2, averageSteps=4)
    row = V.row_from_run(spec, RUN, CONTEXT)
    raw = 2.5 / (0.5 * 1.204 * 1600 * spec.voxelAreaM2)
    assert row["cdRaw"] == pytest.approx(raw, abs=1e-12)
    assert row["cdMaskell"] == pytest.approx(raw / (1 + 2.5 * raw * 0.04), abs=1e-12) and row["cdMaskell"] < raw


# ------------------------------------------------------------------------- scoring

def test_scores_the_foreign_fixture():
    out, err, status = V.score_results(json.load(open(os.path.join(GOLDEN, "foreign-results.json"))))
    assert status == 0
    assert out == ["ahmed-25, 0.3000, 0.285, 5.26", "sphere-re-100, 1.0000, 1.09, -8.26", "cube, 1.0500, -, -",
                   "unmatched: 1"]
    assert any("cylinder-re-100" in e for e in err)


def test_scores_reference_values_and_default_rows_only():
    out, _, _ = V.score_results({"solver": "fixture 1", "rows": [{"case": "ahmed-25", "cd": 0.2850001},
                                                                  {"case": "sphere-re-100", "cd": 1.09}]})
    assert out == ["ahmed-25, 0.2850, 0.285, 0.00", "sphere-re-100, 1.0900, 1.09, 0.00", "unmatched: 0"]
    doc = {"schemaVersion": 1,
           "laminar": [{"case": "sphere-re-100", "cd": 1.2, "isDefault": True},
                       {"case": "sphere-re-100", "cd": 9, "isDefault": False}],
           "validation": [{"case": "ahmed-25", "cd": 0.57, "isDefault": True}, {"case": "ahmed-25", "cd": 7, "isDefault": False},
                          {"case": "cube", "cd": 1.1, "isDefault": True}, {"case": "sphere", "cd": 0.5, "isDefault": True}],
           "confinement": [{"case": "cube-confinement", "cd": 1, "isDefault": True}],
           "sensitivity": [{"case": "ahmed-25", "cd": 8, "isDefault": False}]}
    out, _, status = V.score_results(doc)
    assert status == 0 and out == ["sphere-re-100, 1.2000, 1.09, 10.09", "ahmed-25, 0.5700, 0.285, 100.00",
                                   "cube, 1.1000, -, -", "sphere, 0.5000, -, -", "unmatched: 0"]
    assert V.score_results({"schemaVersion": 1, "laminar": [], "validation": []}) == (["unmatched: 0"], [], 0)


@pytest.mark.parametrize("doc", [{"rows": [{"case": "cube", "cd": 1}]}, {"solver": "s 1", "rows": [{"cd": 1}]},
                                 {"solver": "s 1", "rows": [{"case": "cube", "cd": "1"}]}, {"solver": "s 1", "rows": [3]},
                                 {"cases": []}, []])
def test_scoring_rejects_malformed_documents(doc):
    out, err, status = V.score_results(doc)
    assert status == 1 and out == [] and err and all(e.startswith("score-results:") for e in err)


def test_scores_the_reference_results_file_and_the_cli(tmp_path):
    out, _, status = V.score_results({"schemaVersion": 1, **{k: REF[k] for k in ("laminar", "validation")}})
    assert status == 0 and out[-1] == "unmatched: 0"
    assert "sphere-re-100, 1.1853, 1.09, 8.75" in out  # the L TRT row, Cd 1.1853357205575648
    bad = tmp_path / "bad.json"
    bad.write_text("{")
    run = lambda *a: subprocess.run([sys.executable, "-m", "pocketwind_amd.validation", *a], cwd=ROOT,
                                    capture_output=True, text=True)
    r = run("score", str(bad))
    assert r.returncode == 1 and "cannot read" in r.stderr
    assert run().returncode == 2
    r = run("score", os.path.join(GOLDEN, "foreign-results.json"))
    assert r.returncode == 0 and r.stdout.splitlines()[-1] == "unmatched: 1"


@pytest.mark.parametrize("x,digits,text", [(0.125, 2, "0.13"), (-0.125, 2, "-0.13"), (1.005, 2, "1.00"),
                                           (-0.0, 2, "0.00"), (-0.001, 2, "-0.00"), (5.2631578947, 2, "5.26")])
def test_to_fixed_follows_number_to_fixed(x, digits, text):
    assert V.to_fixed(x, digits) == text


def test_body_drag_binning_follows_the_flow_through_rule():
    """runBodyDrag's bins (cases.ts:293-321) over a sample every 8 steps of four flow-throughs
    of 1280 steps: 159, 160, 160, 160 samples ending at steps 1272, 2552, 3832, 5112 (the step
    5120 sample falls past the end), as the oracle fixture's bins record."""
    G = json.load(open(os.path.join(GOLDEN, "body_drag_oracle.json")))
    samples = [(s, 2.0 if s < 1280 else (1.0 if s % 16 else 3.0)) for s in range(8, 5121, 8)]
    bins = V.drag_flow_throughs(samples, 1280, 4, 0.5)
    oracle_bins = G["operators"]["trt"]["flowThroughs"]
    assert [(b["step"], b["samples"]) for b in bins] == [(b["step"], b["samples"]) for b in oracle_bins]
    assert bins[0]["meanCd"] == 1.0 and bins[0]["relStdPct"] == 0.0
    # 80 samples of 0.5 and 80 of 1.5 about the mean 1: std = sqrt(160 * 0.25 / 159).
    assert bins[1]["meanCd"] == 1.0 and bins[1]["relStdPct"] == pytest.approx(100 * math.sqrt(40 / 159), rel=1e-12)
    assert V.drag_flow_throughs([(8, 1.0)], 1280, 2, 1.0)[1] == {"flowThrough": 2, "step": 0, "samples": 0,
                                                                 "meanCd": None, "relStdPct": None}
