"""Writes c1_sphere_oracle.json: the FP64 oracle's drag on the small laminar sphere case.

Case: 96 x 48 x 48 lattice; analytic sphere of diameter 12 cells centred at (32, 24, 24)
(solid where the cell centre lies strictly inside); x = 0 EQ inlet at u = 0.05; zero-gradient
outlet; free-slip side walls; TRT with halfway bounce-back; Re 100 on the diameter, so
nu = 0.05 * 12 / 100 and tau0 = 3 nu + 1/2. 20 000 steps from the uniform inflow state, the
momentum-exchange force sampled every 8 steps, Cd = mean fx over the last 5 000 steps /
(1/2 u^2 pi (D/2)^2). Run with TRT, as the case names it, and with REG: at this viscosity
(tau0 = 0.518) TRT's force turns non-finite near the inlet axis after about 10 000 steps, with
the reference's boundaries (periodic sides, equilibrium outlet) as with these; the file records
the first non-finite sample step for TRT and the drag for REG.

Run from the repository root: python tests/golden/make_c1_sphere.py  (about 5 min, 8 threads)
"""
from __future__ import annotations

import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(os.path.dirname(HERE)), os.path.dirname(HERE)]

import oracle  # noqa: E402
from grids import grid_from_kinds, sphere_kinds  # noqa: E402

LAYOUT = (96, 48, 48)
CENTRE = (32, 24, 24)
DIAMETER = 12
U_IN = 0.05
REYNOLDS = 100
STEPS = 20_000
AVERAGE = 5_000
EVERY = 8


def run(operator, threads: int) -> dict:
    grid = grid_from_kinds(LAYOUT, sphere_kinds(LAYOUT, CENTRE, DIAMETER))
    nu = U_IN * DIAMETER / REYNOLDS
    cpu = oracle.Solver(LAYOUT, 3 * nu + 0.5, operator, False, oracle.BB,
                        flags=grid.geometry[:grid.cells], equilibrium_state=(1.0, (U_IN, 0.0, 0.0)),
                        threads=threads, sides="free-slip", outlet="zero-gradient")
    start = time.time()
    window, first_nonfinite = [], None
    for step in range(EVERY, STEPS + 1, EVERY):
        cpu.step(EVERY)
        force = cpu.momentum_exchange()
        if not all(math.isfinite(v) for v in force):
            first_nonfinite = step
            break
        if step > STEPS - AVERAGE:
            window.append(force)
    out = {"first_nonfinite_sample": first_nonfinite, "oracle_seconds": round(time.time() - start, 1)}
    if first_nonfinite is None:
        fx, fy, fz = (sum(f[k] for f in window) / len(window) for k in range(3))
        out.update(samples=len(window), force=[fx, fy, fz],
                   cd=fx / (0.5 * U_IN ** 2 * math.pi * (DIAMETER / 2) ** 2))
    return out


def main(threads: int = 8) -> dict:
    grid = grid_from_kinds(LAYOUT, sphere_kinds(LAYOUT, CENTRE, DIAMETER))
    nu = U_IN * DIAMETER / REYNOLDS
    result = {
        "layout": list(LAYOUT), "centre": list(CENTRE), "diameter": DIAMETER, "u_in": U_IN,
        "reynolds": REYNOLDS, "tau0": 3 * nu + 0.5, "steps": STEPS, "average_steps": AVERAGE,
        "sample_every": EVERY, "solid_cells": grid.solid_cells,
        "boundary_cells": int(grid.boundary_cells.size), "threads": threads,
        "trt": run(oracle.TRT, threads), "reg": run(oracle.REG, threads),
    }
    with open(os.path.join(HERE, "c1_sphere_oracle.json"), "w") as f:
        json.dump(result, f, indent=1)
    return result


if __name__ == "__main__":
    print(json.dumps(main(), indent=1))
