"""The build-only configurations at their stated workloads.

C3: the seeded set of 64 tunnels (grids 32^3 to 128x64x64, sphere / ellipsoid / rounded box,
yaw 0-30 degrees, Re 20-200, u_in 0.03-0.06, f32), 10 000 steps with a force sample every 8,
TRT and REG: every tunnel's samples from the batched launch equal those of the same tunnel
stepped alone, bit for bit (NaN where TRT diverges, in the 40 tunnels with tau0 <= 0.5175).

C4: the closure trained at 128x64x64 on 200 coarse/fine pairs (pocketwind_amd/closure_c4.npz)
evaluated on the slanted bluff body at 256x128x128, Re 2000, REG + LES, f16: nearly every wall
cell is inside the network's trusted range, and the drag is reported beside the wall law's and
the wall law's on the twice finer grid. The run is two flow-throughs with the mean over the
second (the full six take about a minute of the three solvers).

C5: 8 yaws (0-14 degrees) of the 217 088-triangle noisy slanted body, each in a 512x256x256
tunnel (268 M cells), f16 REG + LES with the learned closure every 8 steps, in one batched
launch per step, 1 M particles advected every step; a short window of the pipeline (0.1
flow-throughs, forces over the last 0.05), and the zero-yaw tunnel stepped alone with the same
closure gives the batched tunnel's populations bit for bit.
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from pocketwind_amd import closure, sweep, tunnel_set
from pocketwind_amd.batch import TunnelBatch
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
C4_CLOSURE = os.path.join(HERE, "..", "pocketwind_amd", "closure_c4.npz")


@pytest.mark.parametrize("operator,finite", [("trt", 24), ("reg", 64)])
def test_c3_sixty_four_tunnels_batched_equal_singles(operator, finite):
    rec = tunnel_set.run(count=64, steps=10_000, timing_steps=200, operator=operator,
                         log=lambda m: print(m, flush=True))
    print({k: rec[k] for k in ("batch_mlups", "single_mlups", "batch_run_s", "single_run_s", "finite_tunnels")})
    assert rec["cells"] == 12509184
    assert rec["identical_tunnels"] == 64
    assert rec["finite_tunnels"] == finite
    assert rec["max_cd_difference"] == 0.0
    assert rec["batch_mlups"] > rec["single_mlups"]


def test_c4_closure_on_the_bluff_body_at_re_2000():
    rec = closure.evaluate_cd(2000.0, (256, 128, 128), 64, C4_CLOSURE, flow_throughs=2, average=1,
                              log=lambda m: print(m, flush=True))
    print(rec)
    assert rec["wall_cells_on_network"] >= 0.95 * rec["wall_cells"]
    for key in ("cd_f16_learned", "cd_f32_law", "cd_f32_law_2x_grid"):
        assert math.isfinite(rec[key]) and 0.2 < rec[key] < 3.0


def test_c5_eight_yaws_at_512x256x256_with_a_million_particles():
    rec = sweep.run(layout=(512, 256, 256), length_cells=128, flow_throughs=0.1, average_flow_throughs=0.05,
                    particles=1_000_000, log=lambda m: print(m, flush=True))
    print({k: v for k, v in rec.items() if k != "rows"})
    assert rec["tunnels"] == 8 and rec["cells"] == 8 * 512 * 256 * 256
    assert rec["triangles"] == 217088
    assert rec["particles"] == 1_000_000 and rec["particles_finite"]
    assert rec["force_samples"] == rec["average_steps"] // 8
    assert [r["yaw_deg"] for r in rec["rows"]] == list(sweep.DEFAULT_YAWS)
    for r in rec["rows"]:
        assert math.isfinite(r["cd"]) and math.isfinite(r["cs"]) and 0.1 < r["cd"] < 5
    assert rec["batched_step_mlups"] > 0 and rec["particle_updates_per_s"] > 0


def test_c5_zero_yaw_tunnel_equals_its_single_run():
    layout, length, tau0, u_in = (512, 256, 256), 128, 0.53, 0.05
    mesh = sweep.bodies.noisy_bluff_body(levels=5)
    grids = [g for g, _ in sweep.sweep_grids(mesh, layout, length, sweep.DEFAULT_YAWS)]
    shared = SolverParams(tau0, u_in, True, operator="reg", precision="f16", wall_model="learned")
    net, norm, _ = closure.load(sweep.DEFAULT_CLOSURE)
    steps = 40
    batch = TunnelBatch.create(grids, [tau0] * len(grids), [u_in] * len(grids), shared)
    try:
        batch.set_closure(net, norm)
        batch.set_closure_interval(8)
        batch.step(steps, True)
        batched = batch.read_populations(0)
    finally:
        batch.destroy()
    single = GpuSolver.create(layout, grids[0], shared)
    try:
        single.set_closure(net, norm)
        single.set_closure_interval(8)
        single.step(steps, True)
        alone = single.read_populations()
        on_network = int(np.sum(single.closure_output()[:, 3] > 0))
    finally:
        single.destroy()
        net.destroy()
    print("wall cells on the network:", on_network, "differing words:",
          int(np.count_nonzero(batched.view(np.uint16) != alone.view(np.uint16))))
    assert on_network > 0
    assert np.array_equal(batched.view(np.uint16), alone.view(np.uint16))
