#!/usr/bin/env python3
"""Throughput of the D3Q19 step on one MI355X: MLUPS and % of HBM roofline, with Cd.

Workload (the metric's configuration): the bluff body of 1.0 x 0.37 x 0.28 L (front edges
rounded at 0.1 L, 25 degree rear slant over the last 0.2 L), L = 64 cells, nose at 0.25 nx,
centred, in a 256x128x128 D3Q19 tunnel with EQ inlet, pressure outlet, periodic y/z,
halfway bounce-back, regularised collision + Smagorinsky (Cs 0.1) with the wall-law
boundary collision, Re 320 on L (nu 0.01, tau0 0.53), u_in 0.05. The grid is voxelised on
the device. The reference's own throughput grid (all fluid, fully periodic,
src/app/bench-page.ts:245-260) is timed beside it.

A "step" is one fused pull-stream + collide pass over the whole lattice with the populations
resident in HBM: exactly one launch of the step kernel. Timing: W untimed warm-up steps,
then K steps between two HIP events recorded on the solver's own stream (the stream the step
kernel runs on), median of three repeats; the per-launch duration is the timed span / K.
Algorithmic bytes per cell update follow the reference's BYTES_PER_CELL_STEP
(src/app/bench-page.ts:44-45): 19 populations read + 19 written + one 4-byte flag, 156 B at
f32 and 80 B at f16/bf16 storage, over every cell of the lattice.

Headline: f16 deviation storage with the learned wall closure (the network shipped in
pocketwind_amd/closure_c2.npz, trained by `python -m pocketwind_amd.closure` on coarse/fine
pairs of other bodies at this resolution), against plain f32 storage with the wall law. Every
8 steps the closure is evaluated over the wall cells by the first blocks of that step's own
launch, on the state the step reads, and its stress serves the 8 steps after it; the same with
an evaluation in every step's launch is reported beside it. The roofline is the step kernel's,
timed in the f16 run without the closure.

Cd: six flow-throughs (6 ceil(nx / u_in) steps) with an FP64 force sample every 8 steps,
mean over the last two (src/validation/cases.ts:294-320), Cd = fx / (1/2 u^2 A) with A the
voxel frontal area; reported for each configuration, and for plain f32 with the wall law on
the 2x finer grid (512x256x256, L = 128, the same Re), the metric's Cd reference.

The CPU baseline is the FP64 C restatement of the reference CPU solver (the oracle) stepping
the same body grid single-threaded, like the reference's JS path, for a bounded number of
steps. Beside it: scripts/drag-cpu.ts's small Ahmed case and the laminar sphere in 96x48x48
(free-slip sides, zero-gradient outlet), each timed on the device and on the oracle.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline, D3Q19 256x128x128; learned+fp16 vs plain fp32 at Cd±1%"
HBM_PEAK_GBPS = 8000.0
LAYOUT = (256, 128, 128)
BODY_LENGTH = 64
U_IN = 0.05
TAU0 = 3 * (U_IN * BODY_LENGTH / 320) + 0.5  # Re 320 on L: nu = 0.01, tau0 = 0.53


def bytes_per_cell(precision: str) -> int:
    """The reference's convention (src/app/bench-page.ts:44-45): 19 populations read and
    written plus a 4-byte flag per cell."""
    return 156 if precision == "f32" else 80


def algorithmic_bytes(grid, precision: str) -> int:
    """Bytes one step has to move on this grid (pocketwind_amd.lattice.step_bytes)."""
    from pocketwind_amd.lattice import step_bytes
    return step_bytes(grid, precision)


def empty_grid(layout):
    from pocketwind_amd.lattice import NO_Q, VoxelGrid
    cells = layout[0] * layout[1] * layout[2]
    geometry = np.zeros(2 * cells, np.uint32)
    geometry[cells:] = NO_Q
    return VoxelGrid(*layout, geometry)


CLOSURE = os.path.join(ROOT, "pocketwind_amd", "closure_c2.npz")
CLOSURE_EVERY = 8  # steps between closure evaluations (the wall stress is reused in between)
# The closure runs in the first blocks of an evaluating step's own launch, on the state that step
# reads; its output serves the next CLOSURE_EVERY steps (GpuSolver.set_closure_placement).
CLOSURE_PLACEMENT = "alongside"


def body_grid(layout=LAYOUT, length=BODY_LENGTH):
    """The body with its wall-link fractions (the closure's q feature; the walls stay halfway
    bounce-back, which does not read them)."""
    from pocketwind_amd import bodies
    from pocketwind_amd.voxelize import voxelize
    body, _ = bodies.place_mesh(bodies.bluff_body(), None, length, layout, False)
    return voxelize(body, None, layout, False, link_fractions=True)


def params_for(precision: str, learned: bool = False, tau0: float = TAU0):
    from pocketwind_amd.gpu_solver import SolverParams
    return SolverParams(tau0, U_IN, True, operator="reg", precision=precision,
                        wall_model="learned" if learned else "law")


def make_solver(grid, params, closure=None, every=CLOSURE_EVERY, placement=CLOSURE_PLACEMENT):
    from pocketwind_amd.gpu_solver import GpuSolver
    solver = GpuSolver.create(grid.layout, grid, params)
    if not isinstance(solver, GpuSolver):
        raise MemoryError(solver.message)
    if closure is not None:
        solver.set_closure(*closure)
        solver.set_closure_interval(every)
        solver.set_closure_placement(placement)
    return solver


def time_solver(grid, params, steps, warmup, repeats=3, closure=None, every=CLOSURE_EVERY,
                placement=CLOSURE_PLACEMENT):
    solver = make_solver(grid, params, closure, every, placement)
    try:
        solver.step(warmup, False)
        solver.synchronize()
        runs = []
        for _ in range(repeats):
            wall0 = time.perf_counter()
            ms = solver.time_steps(steps)
            wall = (time.perf_counter() - wall0) * 1e3
            runs.append((ms, wall))
        solver.step(1, True)
        finite = bool(np.all(np.isfinite(solver.read_macro())))
        variant = f"{solver.kernel}/{solver.block_size}/xcd-chunks-{solver.xcd_chunks}"
    finally:
        solver.destroy()
    runs.sort()
    ms, wall = runs[len(runs) // 2]
    launch_ms = ms / steps
    cells = grid.cells
    return {"ms": ms, "wall_ms": wall, "launch_us": launch_ms * 1e3,
            "mlups": cells * steps / (ms * 1e3), "finite": finite, "variant": variant,
            "gbps": algorithmic_bytes(grid, params.precision) / (launch_ms * 1e6),
            "bytes": algorithmic_bytes(grid, params.precision),
            "gbps_reference_convention": bytes_per_cell(params.precision) * cells / (launch_ms * 1e6)}


def measured_traffic(precision: str, variant: str):
    """HBM bytes per launch that rocprof counters gave for this storage, cache policy and XCD
    chunking (profiles/step_traffic.json), or None when that variant was not profiled. The
    block size is left out of the key: it changes the traffic by well under a percent."""
    path = os.path.join(ROOT, "profiles", "step_traffic.json")
    try:
        table = json.load(open(path))["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None
    kernel, _, chunking = variant.split("/")
    return table.get(f"{precision}/{kernel}/{chunking}")


def body_cd(grid, params, closure=None, every=CLOSURE_EVERY):
    from pocketwind_amd import drag
    solver = make_solver(grid, params, closure, every)
    try:
        flow = drag.flow_through_steps(grid.nx, U_IN)
        samples = solver.step_sampled(6 * flow, 8)
        fx, _, _ = drag.window_mean_force(samples, 6 * flow, 2 * flow)
    finally:
        solver.destroy()
    return fx * drag.cd_per_lattice_force(grid.frontal_cells, U_IN)


def voxelize_timing(layout=LAYOUT, length=BODY_LENGTH, repeats=5):
    """Device voxelisation of the 217 088-triangle noisy body (pw_voxelize, host copies in and
    out included, median of `repeats`) beside the oracle's restatement of voxelize.ts on one
    thread; the reference's own budget is 400 ms for 200k triangles at this grid
    (tests/unit/voxelize/voxelize.bench.ts:7-16)."""
    import oracle
    from pocketwind_amd import bodies
    from pocketwind_amd.voxelize import voxelize
    body, _ = bodies.place_mesh(bodies.noisy_bluff_body(), None, length, layout, False)
    voxelize(body, None, layout, False, link_fractions=True)  # warm-up (module load, allocations)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        grid = voxelize(body, None, layout, False, link_fractions=True)
        times.append((time.perf_counter() - t0) * 1e3)
    t0 = time.perf_counter()
    ref = oracle.voxelize(body, None, layout, False, link_fractions=True)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    return {"triangles": int(body.size // 9), "layout": list(layout), "link_fractions": True,
            "device_ms": round(sorted(times)[len(times) // 2], 2),
            "bit_exact": bool(np.array_equal(grid.geometry, ref.geometry)
                              and np.array_equal(grid.boundary_cells, ref.boundary_cells)),
            "cpu_baseline": {"value_ms": round(cpu_ms, 1), "cores": 1, "kind": "port",
                             "sample": "FP64 C oracle (restated voxelize.ts and wall-links.ts), the same mesh"}}


def drag_cpu_case(oracle_steps: int = 480):
    """The case scripts/drag-cpu.ts runs on the reference's CPU solver (the Ahmed body with its
    stilts on the floor, 24 cells in 64 x 32 x 32, LES, TRT, four flow-throughs of 1280 steps,
    drag-cpu.ts:11-15): the device's whole run (voxelisation and set-up included) with its Cd
    per flow-through, beside the oracle stepping the same grid on one thread."""
    import oracle
    from pocketwind_amd import bodies
    from pocketwind_amd import validation as V
    t0 = time.perf_counter()
    out = V.run_body_drag("trt")
    device_s = time.perf_counter() - t0
    layout = V.DRAG_LAYOUT
    cells = layout[0] * layout[1] * layout[2]
    steps = V.DRAG_FLOW_THROUGHS * out["stepsPerFlowThrough"]
    body, sup = bodies.place_mesh(bodies.ahmed_mesh(), bodies.stilts_mesh(), V.DRAG_BODY_LENGTH_CELLS, layout, True)
    g = oracle.voxelize(body, sup, layout, True, link_fractions=False)
    cpu = oracle.Solver(layout, out["tau0"], oracle.TRT, True, flags=g.geometry[:cells],
                        equilibrium_state=(1.0, (U_IN, 0.0, 0.0)), threads=1)
    t0 = time.perf_counter()
    cpu.step(oracle_steps)
    cpu_s = time.perf_counter() - t0
    return {"layout": list(layout), "operator": "trt", "steps": steps, "device_s_with_setup": round(device_s, 3),
            "device_mlups": round(cells * steps / device_s / 1e6, 1),
            "cd_per_flow_through": [round(b["meanCd"], 6) for b in out["flowThroughs"]],
            "cpu_baseline": {"value": round(cells * oracle_steps / cpu_s / 1e6, 3), "unit": "MLUPS", "cores": 1,
                             "kind": "port", "sample": f"FP64 C oracle (restated CpuSolver), the same grid and "
                                                       f"operator, {oracle_steps} steps on one thread"}}


C1_LAYOUT, C1_CENTRE, C1_DIAMETER, C1_REYNOLDS = (96, 48, 48), (32, 24, 24), 12, 100


def sphere_case(steps: int = 4000, oracle_seconds: float = 5.0):
    """The small laminar sphere (D = 12 at (32, 24, 24) in 96x48x48, Re 100, REG, f32, free-slip
    sides, zero-gradient outlet): device MLUPS over `steps` steps timed with HIP events on the
    solver's stream (median of three), beside the oracle stepping the same grid on one thread for
    about `oracle_seconds`. At 221 184 cells a step is a few microseconds, so this line shows the
    launch-bound end of the device, not its bandwidth."""
    import oracle
    from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
    from pocketwind_amd.lattice import grid_from_kinds, sphere_kinds
    grid = grid_from_kinds(C1_LAYOUT, sphere_kinds(C1_LAYOUT, C1_CENTRE, C1_DIAMETER))
    tau0 = 3 * (U_IN * C1_DIAMETER / C1_REYNOLDS) + 0.5
    solver = GpuSolver.create(C1_LAYOUT, grid, SolverParams(tau0, U_IN, False, operator="reg",
                                                            sides="free-slip", outlet="zero-gradient"))
    try:
        solver.step(200, False)
        ms = sorted(solver.time_steps(steps) for _ in range(3))[1]
        solver.step(1, True)
        finite = bool(np.all(np.isfinite(solver.read_macro())))
        variant = f"{solver.kernel}/{solver.block_size}/xcd-chunks-{solver.xcd_chunks}"
    finally:
        solver.destroy()
    cpu = oracle.Solver(C1_LAYOUT, tau0, oracle.REG, False, flags=grid.flags,
                        equilibrium_state=(1.0, (U_IN, 0.0, 0.0)), threads=1,
                        sides="free-slip", outlet="zero-gradient")
    cpu.step(1)
    n, elapsed, t0 = 0, 0.0, time.perf_counter()
    while elapsed < oracle_seconds:
        cpu.step(4)
        n += 4
        elapsed = time.perf_counter() - t0
    return {"layout": list(C1_LAYOUT), "operator": "reg", "storage": "f32", "sides": "free-slip",
            "outlet": "zero-gradient", "device_mlups": round(grid.cells * steps / (ms * 1e-3) / 1e6, 1),
            "us_per_step": round(ms * 1e3 / steps, 2), "variant": variant, "finite": finite,
            "cpu_baseline": {"value": round(grid.cells * n / elapsed / 1e6, 3), "unit": "MLUPS", "cores": 1,
                             "kind": "port", "sample": f"FP64 C oracle, the same grid and boundaries, {n} steps "
                                                       f"in {elapsed:.1f} s on one thread"}}


def _oracle_mlups(grid, threads: int, seconds_target: float):
    import oracle
    cpu = oracle.Solver(grid.layout, TAU0, oracle.REG, True, flags=grid.flags,
                        equilibrium_state=(1.0, (U_IN, 0.0, 0.0)), threads=threads)
    cpu.step(1)
    steps, elapsed = 0, 0.0
    t0 = time.perf_counter()
    while elapsed < seconds_target:
        cpu.step(1)
        steps += 1
        elapsed = time.perf_counter() - t0
    return grid.cells * steps / elapsed / 1e6, steps, elapsed


def cpu_baseline(grid, seconds_target: float = 15.0, parallel_seconds: float = 8.0):
    """Oracle (FP64) stepping the same body grid for about `seconds_target` on one thread (the
    reference's CPU path is single-threaded JavaScript), and for about `parallel_seconds` with
    OpenMP over the host cores this process may use (at most 16, one GPU's share)."""
    mlups, steps, elapsed = _oracle_mlups(grid, 1, seconds_target)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    par, par_steps, par_elapsed = _oracle_mlups(grid, threads, parallel_seconds)
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"FP64 C oracle (restated CpuSolver), the same 256x128x128 bluff-body REG+LES "
                      f"tunnel, {steps} steps in {elapsed:.1f} s on one thread",
            "parallel": {"value": round(par, 3), "unit": "MLUPS", "cores": threads,
                         "sample": f"the same, OpenMP over {threads} threads, {par_steps} steps in "
                                   f"{par_elapsed:.1f} s"}}


def time_placements(steps: int, warmup: int) -> int:
    """The f16 learned-closure step on the body grid under each closure placement, beside the f16
    wall-law step (one line each on stderr)."""
    from pocketwind_amd import closure as closure_mod
    grid = body_grid()
    net, norm, _ = closure_mod.load(CLOSURE)
    law = time_solver(grid, params_for("f16"), steps, warmup)
    print(f"# f16 wall law: {law['launch_us']:.2f} us/step, {law['variant']}", file=sys.stderr, flush=True)
    for every in (CLOSURE_EVERY, 1):
        for placement in ("kernel", "overlapped", "alongside"):
            r = time_solver(grid, params_for("f16", learned=True), steps, warmup, closure=(net, norm), every=every,
                            placement=placement)
            print(f"# f16 learned, {placement}, every {every}: {r['launch_us']:.2f} us/step "
                  f"(+{r['launch_us'] - law['launch_us']:.2f}), {r['variant']}, finite={r['finite']}",
                  file=sys.stderr, flush=True)
    net.destroy()
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--only", default=None,
                    help="comma-separated storage precisions (f16, f32, bf16) to time on the body "
                         "grid, then one short copy probe, and nothing else (profiling runs)")
    ap.add_argument("--grid", choices=["body", "empty"], default="body",
                    help="grid of an --only run: the bluff body or the reference's all-fluid grid")
    ap.add_argument("--placements", action="store_true",
                    help="time the f16 learned-closure step on the body grid under each closure placement "
                         "(kernel, overlapped, alongside) at intervals 8 and 1, and nothing else (profiling runs)")
    ap.add_argument("--no-cd", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one tunnel per GPU and launches no "
              "multi-GPU work (--gpus must be 1).", file=sys.stderr)
        return 2

    from pocketwind_amd import _native
    from pocketwind_amd.gpu_solver import copy_bandwidth
    if _native.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    if args.placements:
        return time_placements(args.steps, args.warmup)
    grid = empty_grid(LAYOUT) if (args.only and args.grid == "empty") else body_grid()
    precisions = args.only.split(",") if args.only else ["f16", "f32", "bf16"]
    if any(p not in ("f16", "f32", "bf16") for p in precisions):
        print("--only takes f16, f32 and bf16", file=sys.stderr)
        return 2
    body = {}
    for precision in precisions:
        r = time_solver(grid, params_for(precision), args.steps, args.warmup)
        body[precision] = r
        print(f"# {args.grid if args.only else 'body'} {precision}: {r['mlups']:.1f} MLUPS, {r['launch_us']:.2f} us/launch, "
              f"{r['gbps']:.0f} GB/s, wall {r['wall_ms'] / args.steps * 1e3:.2f} us/step, "
              f"{r['variant']}, finite={r['finite']}", file=sys.stderr, flush=True)
    if args.only:
        print(f"# copy probe: {copy_bandwidth(512 << 20, 4):.0f} GB/s (512 MiB, non-temporal float4)",
              file=sys.stderr, flush=True)
        return 0
    from pocketwind_amd import closure as closure_mod
    net, norm, meta = closure_mod.load(CLOSURE)
    learned = time_solver(grid, params_for("f16", learned=True), args.steps, args.warmup, closure=(net, norm))
    print(f"# body f16 + learned closure: {learned['mlups']:.1f} MLUPS, {learned['launch_us']:.2f} us/step "
          f"(closure + step kernel), {learned['variant']}, finite={learned['finite']}", file=sys.stderr, flush=True)
    learned1 = time_solver(grid, params_for("f16", learned=True), args.steps, args.warmup, closure=(net, norm), every=1)
    print(f"# body f16 + learned closure every step: {learned1['mlups']:.1f} MLUPS, {learned1['launch_us']:.2f} us/step",
          file=sys.stderr, flush=True)
    empty = {}
    for precision in ("f16", "f32"):
        r = time_solver(empty_grid(LAYOUT), params_for(precision), args.steps, args.warmup)
        empty[precision] = r
        print(f"# empty {precision}: {r['mlups']:.1f} MLUPS, {r['launch_us']:.2f} us/launch, "
              f"{r['gbps']:.0f} GB/s, {r['variant']}", file=sys.stderr, flush=True)
    copy_gbps = copy_bandwidth(512 << 20, 40)
    step16 = body["f16"]
    line = {
        "metric": METRIC,
        "value": round(learned["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(learned["launch_us"] / 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 arithmetic, f16 storage",
        "data": "synthetic",
        "config": {"workload": "bluff body L=64 (1.0x0.37x0.28 L, rounded nose, 25 deg slant) in a "
                               "256x128x128 D3Q19 tunnel, REG + Smagorinsky Cs 0.1, halfway bounce-back, "
                               "Re 320 (tau0 0.53), u_in 0.05, device-voxelised",
                   "precision": "f16 deviation storage", "wall_closure": f"learned (bf16 MFMA MLP, evaluated every {CLOSURE_EVERY} steps in the step launch), wall-law fallback",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(step16["gbps"], 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(step16["gbps"] / HBM_PEAK_GBPS, 4),
                     "traffic": measured_traffic("f16", step16["variant"]),
                     "algorithmic_bytes": step16["bytes"],
                     "frac_at_80_bytes_per_cell": round(step16["gbps_reference_convention"] / HBM_PEAK_GBPS, 4),
                     "kernel": "stream_collide_kernel, f16 storage", "kernel_us": round(step16["launch_us"], 2)},
        "f32": {"value": round(body["f32"]["mlups"], 1), "ms_per_step": round(body["f32"]["launch_us"] / 1e3, 5),
                "achieved_gbps": round(body["f32"]["gbps"], 1),
                "frac": round(body["f32"]["gbps"] / HBM_PEAK_GBPS, 4), "wall_closure": "wall law"},
        "f16_wall_law": {"value": round(step16["mlups"], 1), "achieved_gbps": round(step16["gbps"], 1)},
        "bf16_wall_law": {"value": round(body["bf16"]["mlups"], 1), "achieved_gbps": round(body["bf16"]["gbps"], 1)},
        "speedup_vs_f32": round(learned["mlups"] / body["f32"]["mlups"], 3),
        "f16_learned_every_step": {"value": round(learned1["mlups"], 1),
                                   "ms_per_step": round(learned1["launch_us"] / 1e3, 5),
                                   "speedup_vs_f32": round(learned1["mlups"] / body["f32"]["mlups"], 3)},
        "empty_grid": {p: {"value": round(r["mlups"], 1), "achieved_gbps": round(r["gbps"], 1)}
                       for p, r in empty.items()},
        "variants": {p: r["variant"] for p, r in body.items()},
        "copy_gbps": round(copy_gbps, 1),
    }
    if not args.no_cd:
        cds = {"f16_learned": body_cd(grid, params_for("f16", learned=True), closure=(net, norm)),
               "f16_learned_every_step": body_cd(grid, params_for("f16", learned=True), closure=(net, norm), every=1),
               "f16_law": body_cd(grid, params_for("f16")),
               "f32_law": body_cd(grid, params_for("f32")),
               "bf16_law": body_cd(grid, params_for("bf16"))}
        for k, v in cds.items():
            print(f"# Cd {k}: {v:.6f}", file=sys.stderr, flush=True)
        fine_layout = tuple(2 * v for v in LAYOUT)
        fine = body_grid(fine_layout, 2 * BODY_LENGTH)
        cds["f32_law_2x_grid"] = body_cd(fine, params_for("f32", tau0=3 * (2 * (TAU0 - 0.5) / 3) + 0.5))
        print(f"# Cd f32_law_2x_grid: {cds['f32_law_2x_grid']:.6f}", file=sys.stderr, flush=True)
        ref = cds["f32_law_2x_grid"]
        line["cd"] = {k: round(v, 6) for k, v in cds.items()}
        for k in ("f16_learned", "f16_learned_every_step", "f16_law", "f32_law", "bf16_law"):
            line["cd"][f"{k}_vs_2x_grid_pct"] = round(100 * (cds[k] / ref - 1), 3)
        line["cd"]["f16_law_vs_f32_law_pct"] = round(100 * (cds["f16_law"] / cds["f32_law"] - 1), 4)
    net.destroy()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(grid)
        line["voxelize"] = voxelize_timing()
        line["drag_cpu_case"] = drag_cpu_case()
        line["sphere_case"] = sphere_case()
        print(f"# sphere case: {line['sphere_case']}", file=sys.stderr, flush=True)
        print(f"# voxelize: {line['voxelize']}", file=sys.stderr, flush=True)
        print(f"# drag-cpu case: {line['drag_cpu_case']}", file=sys.stderr, flush=True)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
