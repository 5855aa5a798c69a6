"""Pins the oracle against the reference's own saved verification values
(public/validation/results.json `verification[]`, copied into tests/golden/reference_results.json
by tests/golden/make_reference_results.py). Those rows come from the reference's float64 CPU
solver (src/lbm/reference/cases.ts:73-211), so the FP64 restatement must reproduce them:
the Poiseuille step count to convergence exactly, the error percentages and the Taylor-Green
errors to near machine precision."""
from __future__ import annotations

import json
import os

import pytest

import oracle as O
from test_oracle_solver import run_poiseuille, run_taylor_green

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_results.json")))
ROWS = {r["case"]: r for r in GOLDEN["verification"]}


@pytest.mark.parametrize("case", ["poiseuille-tau0-0.6", "poiseuille-tau0-0.8", "poiseuille-tau0-1.5",
                                  "poiseuille-tau0-3", "poiseuille-bgk-tau0-3"])
def test_poiseuille_rows(case):
    row = ROWS[case]
    r = run_poiseuille(row["tau0"], O.BGK if "bgk" in case else O.TRT)
    assert r["steps"] == row["steps"]
    assert abs(r["max_err"] - row["errorPct"]) < 1e-9 * max(1, abs(row["errorPct"]))
    assert abs(r["profile_err"] - row["profileErrorPct"]) < 1e-6 * abs(row["profileErrorPct"]) + 1e-12
    if row.get("forceError"):
        assert abs((r["force"][0] - r["expected"]) - row["forceError"][0]) < 1e-14
        assert abs(r["force"][1]) < 1e-14 and abs(r["force"][2]) < 1e-14


@pytest.mark.parametrize("n,case", [(16, "taylor-green-16"), (32, "taylor-green-32")])
def test_taylor_green_rows(n, case):
    err, _, _ = run_taylor_green(n)
    assert abs(err / ROWS[case]["error"] - 1) < 1e-9


def test_sphere_rows_are_present():
    laminar = {r["id"]: r for r in GOLDEN["laminar"]}
    assert abs(laminar["sphere-re-100-S-trt"]["cd"] - 1.199324666867017) < 1e-15
    assert laminar["sphere-re-100-S-trt"]["steps"] == 15360
