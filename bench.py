#!/usr/bin/env python3
"""Throughput of the D3Q19 step on one MI355X: MLUPS and % of HBM roofline.

A "step" is one fused pull-stream + collide pass over the whole 256x128x128 lattice with the
populations already resident in HBM. Timing: W untimed warm-up steps, then K steps between
two HIP events recorded on the solver's own stream (the stream the step kernel runs on);
every step is one launch of the step kernel, so the per-launch duration is the timed span
divided by K. Algorithmic bytes per cell update follow the reference's BYTES_PER_CELL_STEP
(src/app/bench-page.ts:44-45): 19 populations read + 19 written + one 4-byte flag, i.e.
156 B at f32 and 80 B at f16/bf16 storage.

The CPU baseline is the FP64 C restatement of the reference CPU solver (the oracle) on a
bounded sample of the same workload, single-threaded like the reference's JS path.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline, D3Q19 256x128x128; learned+fp16 vs plain fp32 at Cd±1%"
HBM_PEAK_GBPS = 8000.0
LAYOUT = (256, 128, 128)


def bytes_per_cell(precision: str) -> int:
    return 156 if precision == "f32" else 80


def empty_grid(layout):
    from pocketwind_amd.lattice import NO_Q, VoxelGrid
    cells = layout[0] * layout[1] * layout[2]
    geometry = np.zeros(2 * cells, np.uint32)
    geometry[cells:] = NO_Q
    return VoxelGrid(*layout, geometry)


def time_solver(grid, params, steps, warmup, repeats=3):
    from pocketwind_amd.gpu_solver import GpuSolver
    solver = GpuSolver.create(grid.layout, grid, params)
    if not isinstance(solver, GpuSolver):
        raise MemoryError(solver.message)
    try:
        solver.step(warmup, False)
        solver.synchronize()
        runs = []
        for _ in range(repeats):
            wall0 = time.perf_counter()
            ms = solver.time_steps(steps)
            wall = (time.perf_counter() - wall0) * 1e3
            runs.append((ms, wall))
        macro = solver.read_macro()
        finite = bool(np.all(np.isfinite(macro)))
    finally:
        solver.destroy()
    runs.sort()
    ms, wall = runs[len(runs) // 2]
    return ms, wall, finite


def cpu_baseline(seconds_target: float = 15.0):
    """Oracle (FP64, one thread) on a 64x32x32 sample of the same periodic REG+LES tunnel."""
    import oracle
    layout = (64, 32, 32)
    cells = layout[0] * layout[1] * layout[2]
    flags = np.zeros(cells, np.uint32)
    cpu = oracle.Solver(layout, 0.53, oracle.REG, True, flags=flags,
                        equilibrium_state=(1.0, (0.05, 0.0, 0.0)), threads=1)
    cpu.step(2)
    steps, elapsed = 0, 0.0
    t0 = time.perf_counter()
    while elapsed < seconds_target and steps < 100000:
        cpu.step(4)
        steps += 4
        elapsed = time.perf_counter() - t0
    mlups = cells * steps / elapsed / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"FP64 C oracle, REG+LES periodic {layout[0]}x{layout[1]}x{layout[2]}, "
                      f"{steps} steps in {elapsed:.1f} s, one thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one tunnel shard per GPU and launches "
              "no multi-GPU work (--gpus must be 1).", file=sys.stderr)
        return 2

    from pocketwind_amd import _native
    from pocketwind_amd.gpu_solver import SolverParams, copy_bandwidth
    if _native.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    grid = empty_grid(LAYOUT)
    cells = grid.cells
    results = {}
    for precision in ("f32", "f16"):
        params = SolverParams(0.53, 0.05, True, operator="reg", precision=precision)
        ms, wall, finite = time_solver(grid, params, args.steps, args.warmup)
        launch_ms = ms / args.steps
        mlups = cells * args.steps / (ms * 1e3)
        gbps = bytes_per_cell(precision) * cells / (launch_ms * 1e6)
        results[precision] = dict(ms=ms, wall=wall, mlups=mlups, gbps=gbps, launch_ms=launch_ms,
                                  finite=finite)
        print(f"# {precision}: {mlups:.1f} MLUPS, {launch_ms * 1e3:.2f} us/step, {gbps:.0f} GB/s, "
              f"wall {wall / args.steps * 1e3:.2f} us/step, finite={finite}", file=sys.stderr, flush=True)
    copy_gbps = copy_bandwidth(256 << 20, 40)
    head = results["f16"]
    line = {
        "metric": METRIC,
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["launch_ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 arithmetic, f16 storage",
        "data": "synthetic",
        "config": {"workload": "empty periodic D3Q19 tunnel 256x128x128 (the reference bench grid), "
                               "REG + Smagorinsky Cs 0.1, tau0 0.53, u_in 0.05",
                   "precision": "f16 deviation storage", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(head["gbps"], 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(head["gbps"] / HBM_PEAK_GBPS, 4), "traffic": None},
        "fp32": {"value": round(results["f32"]["mlups"], 1), "ms_per_step": round(results["f32"]["launch_ms"], 5),
                 "achieved_gbps": round(results["f32"]["gbps"], 1),
                 "frac": round(results["f32"]["gbps"] / HBM_PEAK_GBPS, 4)},
        "speedup_f16_vs_f32": round(head["mlups"] / results["f32"]["mlups"], 3),
        "copy_gbps": round(copy_gbps, 1),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
