"""Writes body_drag_oracle.json: the FP64 oracle on the case scripts/drag-cpu.ts runs.

Case (drag-cpu.ts:11-15, :71-110 with runBodyDrag, src/lbm/reference/cases.ts:264-321): the
Ahmed 25 degree preset with its four stilts as supports, placed at 24 cells on the floor of a
64 x 32 x 32 grid; LES on; tau0 = 3 nu + 1/2 with nu = max(1e-2, 0.05 * 24 / Re) at 40 m/s on
the 1.044 m body (the floor applies: tau0 0.53); uniform inflow state at u = 0.05; four
flow-throughs of ceil(64 / 0.05) = 1280 steps; a momentum-exchange sample every 8 steps goes to
flow-through floor(step / 1280) (the last step's sample falls past the end); per flow-through:
last step, sample count, mean Cd on the voxel frontal cells, relative standard deviation (n - 1,
percent). Operators TRT, REG and BGK (the last has no device counterpart; recorded for the
drag-cpu restatement only).

Run from the repository root: python tests/golden/make_body_drag.py  (about a minute, 8 threads)
"""
from __future__ import annotations

import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(os.path.dirname(HERE)), os.path.dirname(HERE)]

import oracle  # noqa: E402
from pocketwind_amd import bodies  # noqa: E402

LAYOUT = (64, 32, 32)
LENGTH = 24
U = 0.05
FLOW_THROUGHS = 4


def bins(cds_by_step, per_flow):
    out = []
    for k in range(FLOW_THROUGHS):
        rows = [(s, c) for s, c in cds_by_step if s // per_flow == k]
        values = [c for _, c in rows]
        mean = sum(values) / len(values)
        std = math.sqrt(sum((v - mean) ** 2 for v in values) / (len(values) - 1))
        out.append({"flowThrough": k + 1, "step": rows[-1][0], "samples": len(values), "meanCd": mean,
                    "relStdPct": 100 * std / abs(mean)})
    return out


def main(threads: int = 8) -> dict:
    body, sup = bodies.place_mesh(bodies.ahmed_mesh(), bodies.stilts_mesh(), LENGTH, LAYOUT, True)
    g = oracle.voxelize(body, sup, LAYOUT, True, link_fractions=False)
    nx, ny, nz = LAYOUT
    re_phys = 40 * 1.044 / 1.5e-5
    nu = max(1e-2, U * LENGTH / re_phys)
    tau0 = 3 * nu + 0.5
    per_flow = math.ceil(nx / U)
    cd_per_force = 1 / (0.5 * U * U * g.frontal_cells)
    result = {"layout": list(LAYOUT), "body_length_cells": LENGTH, "tau0": tau0, "frontal_cells": int(g.frontal_cells),
              "solid_cells": int(g.solid_cells), "steps_per_flow_through": per_flow, "operators": {}}
    for name, op in (("trt", oracle.TRT), ("reg", oracle.REG), ("bgk", oracle.BGK)):
        start = time.time()
        cpu = oracle.Solver(LAYOUT, tau0, op, True, oracle.BB, flags=g.geometry[:nx * ny * nz],
                            equilibrium_state=(1.0, (U, 0.0, 0.0)), threads=threads)
        samples = []
        for step in range(8, FLOW_THROUGHS * per_flow + 1, 8):
            cpu.step(8)
            samples.append((step, float(cpu.momentum_exchange()[0]) * cd_per_force))
        result["operators"][name] = {"flowThroughs": bins(samples, per_flow),
                                     "finite": bool(np.isfinite([c for _, c in samples]).all()),
                                     "oracle_seconds": round(time.time() - start, 1)}
    with open(os.path.join(HERE, "body_drag_oracle.json"), "w") as f:
        json.dump(result, f, indent=1)
    return result


if __name__ == "__main__":
    print(json.dumps(main(), indent=1))
