"""Writes bluff_subcase_oracle.json: the FP64 oracle on the 80 x 40 x 40 bluff-body sub-case.

Case: the slanted bluff body of the 256 x 128 x 128 benchmark scaled to this grid (length 20
cells, nose at 0.25 nx, centred; voxelised by the oracle voxelizer, which the device voxelizer
matches bit for bit); x = 0 EQ inlet at u = 0.05, equilibrium outlet, periodic sides (the
reference's boundaries); REG + Smagorinsky LES with the wall law, halfway bounce-back; Re 320
on the body length, so nu = 0.05 * 20 / 320 and tau0 = 3 nu + 1/2. 30 000 steps from the
uniform inflow state. Recorded: Cd (mean fx over the last 5 000 steps, samples every 8 steps,
on the voxel frontal area), the x-velocity along the line y = ny / 2, z = nz / 2 averaged over
the macro fields at steps 26 000, 27 000, ..., 30 000, and the total density of the non-wall
cells at steps 0 and 30 000.

Run from the repository root: python tests/golden/make_bluff_subcase.py  (about 5 min, 8 threads)
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

LAYOUT = (80, 40, 40)
LENGTH = 20
U_IN = 0.05
REYNOLDS = 320
STEPS = 30_000
AVERAGE = 5_000
EVERY = 8
PROFILE_EVERY = 1000


def grid():
    body, _ = bodies.place_mesh(bodies.bluff_body(), None, LENGTH, LAYOUT, False)
    return oracle.voxelize(body, None, LAYOUT, False)


def main(threads: int = 8) -> dict:
    g = grid()
    nx, ny, nz = LAYOUT
    cells = nx * ny * nz
    flags = g.geometry[:cells]
    kinds = flags & 7
    open_cells = (kinds == 0) | (kinds == 4)
    nu = U_IN * LENGTH / REYNOLDS
    cpu = oracle.Solver(LAYOUT, 3 * nu + 0.5, oracle.REG, True, oracle.BB, flags=flags,
                        equilibrium_state=(1.0, (U_IN, 0.0, 0.0)), threads=threads)
    mass0 = float(cpu.macro_all()[open_cells, 0].sum())
    line = [x + nx * (ny // 2 + ny * (nz // 2)) for x in range(nx)]
    start = time.time()
    forces, profiles = [], []
    for step in range(EVERY, STEPS + 1, EVERY):
        cpu.step(EVERY)
        if step > STEPS - AVERAGE:
            forces.append(cpu.momentum_exchange())
            if step % PROFILE_EVERY == 0:
                profiles.append(cpu.macro_all()[line, 1])
    macro = cpu.macro_all()
    fx, fy, fz = np.mean(forces, axis=0)
    result = {
        "layout": list(LAYOUT), "length_cells": LENGTH, "u_in": U_IN, "reynolds": REYNOLDS,
        "tau0": 3 * nu + 0.5, "steps": STEPS, "average_steps": AVERAGE, "sample_every": EVERY,
        "solid_cells": int(g.solid_cells), "frontal_cells": int(g.frontal_cells),
        "force": [float(fx), float(fy), float(fz)],
        "cd": float(fx / (0.5 * U_IN ** 2 * g.frontal_cells)),
        "centreline_ux": np.mean(profiles, axis=0).tolist(), "profile_samples": len(profiles),
        "profile_every": PROFILE_EVERY,
        "mass_start": mass0, "mass_end": float(macro[open_cells, 0].sum()),
        "oracle_seconds": round(time.time() - start, 1), "threads": threads,
    }
    result["mass_drift"] = result["mass_end"] / result["mass_start"] - 1
    with open(os.path.join(HERE, "bluff_subcase_oracle.json"), "w") as f:
        json.dump(result, f, indent=1)
    return result


if __name__ == "__main__":
    r = main()
    print(json.dumps({k: v for k, v in r.items() if k != "centreline_ux"}, indent=1))
