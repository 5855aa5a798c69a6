"""Yaw sweep of one body in batched tunnels: the full pipeline in one run.

A seeded procedural mesh (the slanted bluff body, subdivided, vertex noise:
``bodies.noisy_bluff_body``, 217 088 triangles by default) is placed as ``placeMesh`` places it
(transform.ts:118-143), turned about the vertical axis through its centre by each yaw angle,
and voxelised on the device (``voxelize``, bit-exact against the reference voxelizer). The
tunnels of all yaws step together in one launch per step (``TunnelBatch``: f16 deviation
storage, REG + Smagorinsky LES, the learned wall closure with the wall law as its fallback),
while streakline particles (particles_update.wgsl:1-60) sample each tunnel's velocity after
every step. Drag and side force per yaw are window means of momentum-exchange samples taken
every 8 steps (runner.ts:234-244), as coefficients on the zero-yaw body's voxel frontal area.

The reference runs one tunnel at a time and has no yaw sweep; the sweep composes its parts.

    python -m pocketwind_amd.sweep --out sweep.json            # 512x256x256, 8 yaws, 1M particles
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Callable, Sequence

import numpy as np

from . import bodies
from .batch import TunnelBatch
from .gpu_solver import FORCE_EVERY_STEPS, SolverParams
from .lattice import step_bytes
from .particles import Particles
from .voxelize import voxelize

DEFAULT_YAWS = (0, 2, 4, 6, 8, 10, 12, 14)  # degrees
DEFAULT_CLOSURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closure_c2.npz")


def yawed(body: np.ndarray, yaw_deg: float) -> np.ndarray:
    """A placed body turned by `yaw_deg` about the vertical axis through the centre of its
    horizontal bounding box (positive yaw turns the nose towards +y)."""
    v = np.asarray(body, np.float32).reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    cx, cy = (lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2
    a = math.radians(yaw_deg)
    c, s = math.cos(a), math.sin(a)
    out = v.copy()
    dx, dy = v[:, 0] - cx, v[:, 1] - cy
    # The nose (smallest x) swings to +y for positive yaw: rotate by -a in the x-y plane.
    out[:, 0] = cx + c * dx + s * dy
    out[:, 1] = cy - s * dx + c * dy
    return out.astype(np.float32).reshape(-1)


def sweep_grids(mesh, layout, length_cells: float, yaws: Sequence[float]):
    """(grid, body) per yaw: the body placed once, then turned, then voxelised on the device."""
    placed, _ = bodies.place_mesh(mesh, None, length_cells, layout, False)
    out = []
    for yaw in yaws:
        body = yawed(placed, yaw)
        out.append((voxelize(body, None, layout, False, link_fractions=False), body))
    return out


def coefficients(force_mean: np.ndarray, u_in: float, area_cells: float):
    """(Cd, Cs): wind-axis drag and side force over 1/2 rho u^2 A with rho = 1."""
    q = 0.5 * u_in ** 2 * area_cells
    return float(force_mean[0] / q), float(force_mean[1] / q)


def run(layout=(512, 256, 256), yaws: Sequence[float] = DEFAULT_YAWS, length_cells: float = 128,
        tau0: float = 0.53, u_in: float = 0.05, flow_throughs: float = 4.0,
        average_flow_throughs: float = 2.0, particles: int = 1_000_000, particle_seed: int = 1,
        closure_file: str | None = DEFAULT_CLOSURE, closure_every: int = 8, mesh_levels: int = 5,
        log: Callable[[str], None] = print) -> dict:
    """Run the sweep and return its record (see the module docstring)."""
    from . import closure as closure_mod
    t_start = time.time()
    mesh = bodies.noisy_bluff_body(levels=mesh_levels)
    t0 = time.time()
    pairs = sweep_grids(mesh, layout, length_cells, yaws)
    t_vox = time.time() - t0
    grids = [g for g, _ in pairs]
    log(f"voxelised {len(yaws)} yaws of a {mesh.size // 9}-triangle mesh in {t_vox:.1f} s "
        f"(solid cells {[g.solid_cells for g in grids]})")
    learned = closure_file is not None
    shared = SolverParams(tau0, u_in, True, operator="reg", precision="f16",
                          wall_model="learned" if learned else "law")
    batch = TunnelBatch.create(grids, [tau0] * len(grids), [u_in] * len(grids), shared)
    parts = []
    try:
        if learned:
            net, norm, _ = closure_mod.load(closure_file)
            batch.set_closure(net, norm)
            batch.set_closure_interval(closure_every)
        per_tunnel = [particles // len(grids) + (k < particles % len(grids)) for k in range(len(grids))]
        for k, (g, body) in enumerate(pairs):
            lo, hi = bodies.mesh_bounds(body)
            if per_tunnel[k]:
                parts.append(Particles.create_in_batch(batch, k, (lo, hi), per_tunnel[k], particle_seed + k,
                                                       max_count=1 << 24))
        cells = batch.cells
        step_bytes_all = sum(step_bytes(g, "f16") for g in grids)
        # Step throughput of the batched launch alone (graph-replayed pairs, no macro writes).
        batch.step(20)
        step_ms = min(batch.time_steps(100) for _ in range(3)) / 100
        total = int(round(flow_throughs * layout[0] / u_in))
        window = int(round(average_flow_throughs * layout[0] / u_in))
        log(f"{len(grids)} tunnels, {cells} cells: {step_ms * 1e3:.1f} us per batched step "
            f"({cells / step_ms / 1e3:.0f} MLUPS); running {total} steps")
        # The pipeline: every step writes the macros, the particles advect on them, the closure
        # runs every `closure_every` steps and the force is sampled every 8.
        batch.force()  # drains the queue
        samples = []
        t0 = time.time()
        done = batch.step_count
        for k in range(1, total + 1):
            batch.step(1, True)
            for p in parts:
                p.update(1.0)
            if (done + k) % FORCE_EVERY_STEPS == 0 and k > total - window:
                samples.append(batch.force())
            if k % 4096 == 0:
                log(f"  step {k} of {total}, {time.time() - t0:.0f} s")
        batch.force()
        loop_s = time.time() - t0
        # Particle advection throughput: every particle set advected 100 times, bracketed by
        # the batch queue's drains.
        batch.force()
        tp = time.time()
        for _ in range(100):
            for p in parts:
                p.update(1.0)
        batch.force()
        particle_s = time.time() - tp
        mean = np.mean(np.stack(samples), axis=0)  # (tunnels, 3)
        area = grids[0].frontal_cells
        rows = []
        for k, yaw in enumerate(yaws):
            cd, cs = coefficients(mean[k], u_in, area)
            rows.append({"yaw_deg": yaw, "cd": cd, "cs": cs, "force": mean[k].tolist(),
                         "solid_cells": grids[k].solid_cells, "frontal_cells": grids[k].frontal_cells})
            log(f"  yaw {yaw:>4} deg: Cd {cd:.4f}  Cs {cs:+.4f}")
        positions = np.concatenate([p.read()[:, :3] for p in parts]) if parts else np.zeros((0, 3))
        mlups_step = cells / step_ms / 1e3
        record = {
            "layout": list(layout), "tunnels": len(grids), "cells": int(cells), "yaws_deg": list(yaws),
            "triangles": int(mesh.size // 9), "length_cells": length_cells, "tau0": tau0, "u_in": u_in,
            "reynolds_on_length": u_in * length_cells / ((tau0 - 0.5) / 3),
            "storage": "f16", "collision": "REG + Smagorinsky LES",
            "wall_closure": f"learned every {closure_every} steps" if learned else "wall law",
            "steps": total, "average_steps": window, "force_samples": len(samples),
            "voxelise_s": round(t_vox, 2),
            "batched_step_us": round(step_ms * 1e3, 2), "batched_step_mlups": round(mlups_step, 1),
            "batched_step_variant": batch.variant,
            "batched_step_bytes": step_bytes_all,
            "batched_step_gbps": round(step_bytes_all / step_ms / 1e6, 1),
            "hbm_roofline_frac": round(step_bytes_all / step_ms / 1e6 / 8000.0, 4),
            "pipeline_s": round(loop_s, 2), "pipeline_mlups": round(cells * total / loop_s / 1e6, 1),
            "particles": int(sum(per_tunnel)),
            "particle_updates_per_s": round(sum(per_tunnel) * 100 / particle_s, 1),
            "particles_finite": bool(np.all(np.isfinite(positions))),
            "rows": rows, "wall_s": round(time.time() - t_start, 1),
        }
        return record
    finally:
        for p in parts:
            p.destroy()
        batch.destroy()


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out")
    ap.add_argument("--layout", type=int, nargs=3, default=[512, 256, 256])
    ap.add_argument("--yaws", type=float, nargs="+", default=list(DEFAULT_YAWS))
    ap.add_argument("--length", type=float, default=128)
    ap.add_argument("--tau0", type=float, default=0.53)
    ap.add_argument("--flow-throughs", type=float, default=4.0)
    ap.add_argument("--average", type=float, default=2.0)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--closure", default=DEFAULT_CLOSURE, help="closure file, or 'none' for the wall law")
    ap.add_argument("--mesh-levels", type=int, default=5)
    a = ap.parse_args(argv)
    record = run(tuple(a.layout), a.yaws, a.length, a.tau0, 0.05, a.flow_throughs, a.average, a.particles,
                 closure_file=None if a.closure == "none" else a.closure, mesh_levels=a.mesh_levels,
                 log=lambda m: print(m, flush=True))
    text = json.dumps(record, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
