"""The validation harness end to end on the device: the reference's cases run on this solver
(forces in newtons every 8 steps, measured frontal area and blockage), turned into result rows by
the reference's rules and checked against the reference's stored rows (public/validation/
results.json).

Laminar sphere at Re 100, tier S: a steady flow, so Cd within 0.5 % of the stored row; the rows
are then written as a foreign results document and scored.

Every stored validation, confinement and sensitivity row: the same voxel area and blockage and
the same pass verdict. The laminar sphere in the XL domain is steady: Cd within 0.5 %. The
turbulent rows (Ahmed body at Re 2.78e6 with floor and stilts, cube face-on, sphere at Re 20 000;
REG + Smagorinsky LES with the wall law; tiers S to XL, f32 and the f16 row; the domain, body
length, inlet speed and stilts sensitivity rows): Cd within 3 % of the stored row. Both runs use the same
scheme from the same initial state, but the reference sums in a different float32 order, so the
chaotic wakes decorrelate within a flow-through and the two Cd values are two samples of a window
mean of a signal whose relative standard deviation is 1.2-4 % (the rows' relStdPct); over the
window's two flow-throughs that mean carries about 1 % sampling error, so two runs differ by
about 1.4 % rms and 3 % is about two of those. Measured on the validation rows: at most 1.84 %
(Ahmed, tier S), every M, L and XL row within 0.7 %."""
from __future__ import annotations

import json
import os

import pytest

from pocketwind_amd import validation as V

pytestmark = pytest.mark.gpu

REF = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_results.json")))
STORED = {r["id"]: r for r in REF["laminar"]}
OTHER = {r["id"]: r for s in ("validation", "confinement", "sensitivity") for r in REF[s]}
CASES = {c.id: c for c in V.build_cases() + [V.smoke_case()]}


def test_laminar_sphere_rows_match_the_reference_rows():
    rows = []
    for operator in ("trt", "reg"):
        spec = V.laminar_sphere_case("S", operator)
        run = V.run_laminar_sphere(spec)
        row = V.row_from_run(spec, run, V.RunContext("test", "2026-01-01T00:00:00.000Z"))
        ref = STORED[spec.id]
        print(f"{spec.id}: Cd {row['cd']:.10f} (reference {ref['cd']:.10f}), relStd {row['relStdPct']:.4f} %, "
              f"drift {row['driftPct']:.4f} %, pass {row['pass']} {row['reasons']}")
        assert (row["voxelAreaM2"], row["blockage"]) == (ref["voxelAreaM2"], ref["blockage"])
        assert abs(row["cd"] / ref["cd"] - 1) < 0.005
        # The S TRT row sits at the 10 % band edge (reference error 10.03 %), so the
        # verdict follows this run's own error.
        expected = [] if abs(row["errorPct"]) <= 10 else [f"reference error {V.to_fixed(row['errorPct'], 1)}% exceeds 10%"]
        assert (row["pass"], row["reasons"]) == (not expected, expected)
        assert len(row["forceHistory"]) == spec.steps // 8 and row["steps"] == spec.steps
        rows.append(row)
    out, err, status = V.score_results(V.foreign_results("pocketwind_amd", rows))
    assert status == 0 and out[-1] == "unmatched: 0" and len(out) == 3


@pytest.mark.parametrize("case_id", list(OTHER))
def test_validation_confinement_and_sensitivity_rows_match_the_reference_rows(case_id):
    spec, ref = CASES[case_id], OTHER[case_id]
    row = V.row_from_run(spec, V.run_case(spec), V.RunContext("test", "2026-01-01T00:00:00.000Z"))
    print(f"{case_id}: Cd {row['cd']:.6f} (reference {ref['cd']:.6f}, {100 * (row['cd'] / ref['cd'] - 1):+.2f} %), "
          f"relStd {row['relStdPct']:.3f} % (reference {ref['relStdPct']:.3f} %), "
          f"{row['msPerStep']:.3f} ms/step (reference {ref['msPerStep']:.3f})")
    assert (row["voxelAreaM2"], row["blockage"]) == (ref["voxelAreaM2"], ref["blockage"])
    assert abs(row["cd"] / ref["cd"] - 1) < (0.03 if spec.les else 0.005)
    # Ahmed rows fail against the experiment's 0.285 in both; the reason quotes each run's own error.
    assert row["pass"] == ref["pass"] and len(row["reasons"]) == len(ref["reasons"])
    assert len(row["forceHistory"]) == spec.steps // 8 and row["steps"] == spec.steps


SKIPPED = [r for r in REF["skipped"] if "operator" in r]


@pytest.mark.parametrize("row", SKIPPED, ids=[f"{r['case']}-{r['tier']}-{r['operator']}" for r in SKIPPED])
def test_the_reference_s_unstable_trt_runs_fail_here_too(row):
    """The reference skips two TRT runs whose force turned non-finite (a cube in the XL domain at
    step 616, the Ahmed body at tier L at step 552); TRT at tau0 = 1/2 + 3 nu with nu near 1e-6 has
    no stabilising viscosity. Here the same cases end with the same error, at a step within 15 % of
    the reference's (the blow-up grows from float32 rounding, whose order differs between the two
    implementations). Measured: steps 672 and 576."""
    import re as regex
    spec = next(c for c in V.build_cases() if c.case == row["case"] and c.tier == row["tier"]
                and c.operator == row["operator"] and c.group == row["group"])
    with pytest.raises(ValueError, match=r"The force was not finite at step \d+ of \d+\.") as err:
        V.run_case(spec)
    step = int(regex.search(r"step (\d+)", str(err.value)).group(1))
    ref_step = int(regex.search(r"step (\d+)", row["reason"]).group(1))
    print(f"{spec.id}: {err.value} (reference: {row['reason']})")
    assert str(err.value).endswith(f"of {spec.steps}.") and abs(step / ref_step - 1) < 0.15


@pytest.mark.parametrize("check", REF["gridCheck"], ids=[c["case"] for c in REF["gridCheck"]])
def test_grid_checks_match_the_reference_s(check):
    """The app's grid check from tier S at its defaults (floor on, REG + LES, 16-bit storage, the
    interactive viscosity floor fixing Re 320 for the Ahmed body and 90 for the sphere): the same
    body lengths, Cd per tier within 0.1 % of the stored ones (steady flows at these Re; measured
    within 0.01 %) and the same band verdict."""
    out = V.run_grid_check(check["case"])
    print(f"{check['case']}: {out['kind']} {out.get('reason')}, "
          + ", ".join(f"{r['tier']} {r['cd']:.6f} (reference {q['cd']:.6f})" for r, q in zip(out["rows"], check["rows"])))
    assert out["re"] == pytest.approx(check["re"], rel=1e-14)
    assert [r["lengthCells"] for r in out["rows"]] == [r["lengthCells"] for r in check["rows"]]
    for r, q in zip(out["rows"], check["rows"]):
        assert abs(r["cd"] / q["cd"] - 1) < 1e-3
    assert (out["kind"], out.get("reason")) == (check["kind"], check.get("reason"))


BODY_DRAG = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "body_drag_oracle.json")))


@pytest.mark.parametrize("operator", ["trt", "reg"])
def test_body_drag_per_flow_through_against_the_oracle(operator):
    """The case scripts/drag-cpu.ts runs (Ahmed body with stilts on the floor, 24 cells in
    64 x 32 x 32, LES, tau0 0.53, four flow-throughs), on the device in f32 against the FP64
    oracle (tests/golden/make_body_drag.py): the same frontal cells and bins; the first
    flow-through's mean Cd (the start-up transient, relative spread 60 %) within 0.1 % and the
    settled ones within 1e-4 (measured: 2.4e-4 and 1.3e-4 for the first, at most 7e-6 after)."""
    out = V.run_body_drag(operator)
    ref = BODY_DRAG["operators"][operator]
    print(f"{operator}: " + ", ".join(f"{b['meanCd']:.6f} (oracle {r['meanCd']:.6f})"
                                      for b, r in zip(out["flowThroughs"], ref["flowThroughs"])))
    assert out["frontalCells"] == BODY_DRAG["frontal_cells"] and out["finite"] and ref["finite"]
    assert out["tau0"] == pytest.approx(BODY_DRAG["tau0"], rel=1e-15)
    for k, (b, r) in enumerate(zip(out["flowThroughs"], ref["flowThroughs"])):
        assert (b["step"], b["samples"]) == (r["step"], r["samples"])
        assert abs(b["meanCd"] / r["meanCd"] - 1) < (1e-3 if k == 0 else 1e-4)
