"""A seeded set of small independent tunnels, stepped in one batch and one by one.

Each tunnel draws its grid from {32^3, 64x32x32, 96x48x48, 128x64x64} and one body: a sphere,
an ellipsoid or a rounded box whose axes are 0.5-1.5 times a base of 8 cells, turned by a yaw
of 0-30 degrees; Re (on the body's largest extent) uniform in 20-200, inflow 0.03-0.06; TRT,
halfway bounce-back, f32 storage (``operator="reg"`` runs the same set with the regularized
collision). ``run`` steps every tunnel in one ``TunnelBatch`` (one launch
per step, graph-replayed step pairs) and then each tunnel alone in a ``GpuSolver`` with the
batch's per-cell step variant, and reports per-tunnel drag from both (momentum-exchange
samples every 8 steps, mean over the second half of the run, on the tunnel's voxel frontal
area) and the aggregate throughput of either way. The reference runs one tunnel at a time
(src/app/solver-slot.ts:14-44); batching is this library's own.

    python -m pocketwind_amd.tunnel_set --out tunnels.json
"""
from __future__ import annotations

import json
import math
import time
from typing import Callable

import numpy as np

from . import bodies
from .batch import TunnelBatch
from .gpu_solver import FORCE_EVERY_STEPS, GpuSolver, SolverParams
from .sweep import yawed
from .voxelize import voxelize

LAYOUTS = ((32, 32, 32), (64, 32, 32), (96, 48, 48), (128, 64, 64))
KINDS = ("sphere", "ellipsoid", "rounded-box")
BASE_CELLS = 8.0


def random_cases(count: int = 64, seed: int = 2024) -> list[dict]:
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(count):
        layout = LAYOUTS[int(rng.integers(len(LAYOUTS)))]
        kind = KINDS[int(rng.integers(len(KINDS)))]
        axes = BASE_CELLS * rng.uniform(0.5, 1.5, 3)
        if kind == "sphere":
            axes[:] = axes[0]
        cases.append({"id": k, "layout": list(layout), "body": kind, "axes_cells": axes.round(6).tolist(),
                      "yaw_deg": float(rng.uniform(0, 30)), "re": float(rng.uniform(20, 200)),
                      "u_in": float(rng.uniform(0.03, 0.06))})
    return cases


def body_mesh(case: dict) -> np.ndarray:
    """The case's body in lattice cells, nose at 0.25 nx, centred, turned by its yaw."""
    ax = np.asarray(case["axes_cells"], np.float64)
    if case["body"] == "rounded-box":
        unit = bodies.bluff_body(length=1.0, width=1.0, height=1.0, radius=0.2, slant_length=0.05,
                                 slant_deg=0.0)
    else:
        unit = bodies.sphere_mesh()  # radius 50 about the origin
    v = unit.reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    v = (v - lo) / (hi - lo) * ax  # box [0, axes]
    placed, _ = bodies.place_mesh(v.astype(np.float32).reshape(-1), None, float(ax.max()),
                                  tuple(case["layout"]), False)
    return yawed(placed, case["yaw_deg"])


def case_grid(case: dict):
    return voxelize(body_mesh(case), None, tuple(case["layout"]), False, link_fractions=False)


def tau_for(case: dict) -> float:
    nu = case["u_in"] * max(case["axes_cells"]) / case["re"]
    return 3 * nu + 0.5


def drag_coefficient(samples: np.ndarray, steps: int, u_in: float, area: int) -> float:
    window = samples[samples[:, 0] > steps // 2]
    return float(window[:, 1].mean() / (0.5 * u_in ** 2 * area))


def run(count: int = 64, steps: int = 10_000, seed: int = 2024, timing_steps: int = 1000,
        operator: str = "trt", log: Callable[[str], None] = print) -> dict:
    cases = random_cases(count, seed)
    t0 = time.time()
    grids = [case_grid(c) for c in cases]
    log(f"voxelised {count} tunnels in {time.time() - t0:.1f} s; "
        f"{sum(g.cells for g in grids)} cells in all")
    shared = SolverParams(0.6, 0.05, False, operator=operator, precision="f32")
    taus = [tau_for(c) for c in cases]
    us = [c["u_in"] for c in cases]
    batch = TunnelBatch.create(grids, taus, us, shared)
    try:
        batch.step(10)
        batch_ms = min(batch.time_steps(timing_steps) for _ in range(3))
        batch.destroy()
        batch = TunnelBatch.create(grids, taus, us, shared)
        t1 = time.time()
        samples = batch.step_sampled(steps, FORCE_EVERY_STEPS)
        batch_run_s = time.time() - t1
    finally:
        batch.destroy()
    cells = sum(g.cells for g in grids)
    rows, single_ms, single_run_s = [], 0.0, 0.0
    for k, (c, g) in enumerate(zip(cases, grids)):
        s = GpuSolver.create(g.layout, g, SolverParams(taus[k], us[k], False, operator=operator, precision="f32"))
        try:
            s.set_kernel("cell-nt")  # the batch's per-cell step variant
            s.step(10, False)
            single_ms += min(s.time_steps(timing_steps) for _ in range(3))
        finally:
            s.destroy()
        s = GpuSolver.create(g.layout, g, SolverParams(taus[k], us[k], False, operator=operator, precision="f32"))
        try:
            s.set_kernel("cell-nt")
            t1 = time.time()
            single = s.step_sampled(steps, FORCE_EVERY_STEPS)
            single_run_s += time.time() - t1
        finally:
            s.destroy()
        cd_batch = drag_coefficient(samples[k], steps, c["u_in"], g.frontal_cells)
        cd_single = drag_coefficient(single, steps, c["u_in"], g.frontal_cells)
        rows.append({**c, "tau0": taus[k], "frontal_cells": g.frontal_cells, "solid_cells": g.solid_cells,
                     "cd_batch": cd_batch, "cd_single": cd_single,
                     "identical_samples": bool(np.array_equal(samples[k], single, equal_nan=True))})
    identical = sum(r["identical_samples"] for r in rows)
    log(f"{identical} of {count} tunnels give bit-identical force samples batched and alone")
    return {
        "tunnels": count, "seed": seed, "steps": steps, "cells": int(cells), "operator": operator,
        "batch_step_ms": batch_ms / timing_steps, "single_steps_ms": single_ms / timing_steps,
        "batch_mlups": cells * timing_steps / (batch_ms * 1e3),
        "single_mlups": cells * timing_steps / (single_ms * 1e3),
        "batch_run_s": round(batch_run_s, 2), "single_run_s": round(single_run_s, 2),
        "identical_tunnels": identical,
        "finite_tunnels": sum(bool(np.isfinite(r["cd_batch"])) for r in rows),
        "max_cd_difference": max((abs(r["cd_batch"] - r["cd_single"]) for r in rows if np.isfinite(r["cd_batch"])), default=0.0),
        "rows": rows,
    }


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out")
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--steps", type=int, default=10_000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--operator", choices=["trt", "reg"], default="trt")
    a = ap.parse_args(argv)
    rec = run(a.count, a.steps, a.seed, operator=a.operator, log=lambda m: print(m, flush=True))
    summary = {k: v for k, v in rec.items() if k != "rows"}
    print(json.dumps(summary, indent=1))
    for r in rec["rows"]:
        print(f"  {r['id']:>2} {r['body']:<11} {str(tuple(r['layout'])):<15} yaw {r['yaw_deg']:5.1f} "
              f"Re {r['re']:6.1f}  Cd batch {r['cd_batch']:.6f} single {r['cd_single']:.6f} "
              f"{'identical' if r['identical_samples'] else 'DIFFERENT'}")
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rec, f, indent=1)


if __name__ == "__main__":
    main()
