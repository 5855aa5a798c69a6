"""The learned wall closure inside the solver.

* The per-wall-cell forces are the terms of the FP64 force sample (their sum equals it).
* The features are well formed (unit normals, link fractions, viscosity).
* With every cell outside the network's trusted range the learned solver falls back to the
  wall law bit for bit.
* Trained on coarse/fine pairs that this solver generates (coarse 48x24x24, body 16 cells,
  fine 96x48x48, Re 100-400, LES, halfway bounce-back with the voxelizer's link fractions as
  the q feature), the network predicts the fine-grid wall shear stress of held-out bodies from
  their coarse-grid features better than the wall law does; a coarse run with the closure
  attached stays finite with the network in use, and evaluating it every 8 steps instead of
  every step moves the mean drag by less than 0.5 % (about 0.2 % measured).
"""
from __future__ import annotations

import numpy as np
import pytest

from pocketwind_amd import bodies, closure, drag
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu

U_IN = 0.05


def body_solver(layout=(48, 24, 24), length=16, wall_model="law", seed=3, nu=0.004):
    return closure._tunnel(closure.random_body(seed), layout, length, nu, U_IN, wall_model=wall_model)


def test_wall_forces_sum_to_the_force_sample():
    solver, grid = body_solver()
    try:
        solver.step(120, False)
        solver.encode_force()
        sample = solver.read_force()
        per_cell = solver.wall_forces()
        cells = solver.wall_cells()
        assert per_cell.shape == (cells.size, 3) and cells.size > 100
        total = per_cell.sum(0)
        for a, b in zip(total, (sample.fx, sample.fy, sample.fz)):
            assert abs(a - b) <= 1e-10 * max(1e-12, abs(sample.fx)) + 1e-15
    finally:
        solver.destroy()


def test_features_are_well_formed():
    solver, grid = body_solver()
    try:
        solver.step(200, False)
        rec = solver.closure_features(U_IN)
        cells = solver.wall_cells()
    finally:
        solver.destroy()
    flags = grid.flags[cells]
    assert np.all((flags & 7) == 0) and np.all((flags >> 3) != 0)
    n = rec[:, 8:11]
    norm = np.linalg.norm(n, axis=1)
    assert np.all((np.abs(norm - 1) < 1e-5) | (norm == 0))
    assert np.all((rec[:, 6] > 0) & (rec[:, 6] <= 1))
    assert np.allclose(rec[:, 7], np.log10(np.float32((np.float32(3 * 0.004 + 0.5) - 0.5) / 3)), atol=1e-5)
    assert np.all(rec[:, 11] > 0.16) and np.all(np.isfinite(rec))


def test_out_of_range_cells_fall_back_to_the_wall_law_bit_for_bit():
    net = closure.Mlp(closure.init_params(1))
    norm = np.concatenate([np.zeros(12), np.ones(12), [1e-4, 1e-6, U_IN]]).astype(np.float32)
    law, _ = body_solver()
    learned, _ = body_solver(wall_model="learned")
    try:
        learned.set_closure(net, norm)
        law.step(60, True)
        learned.step(60, True)
        assert np.array_equal(law.read_populations(), learned.read_populations())
        assert np.all(learned.closure_output()[:, 3] == 0)
    finally:
        law.destroy()
        learned.destroy()
        net.destroy()


def test_learned_closure_beats_the_wall_law_on_held_out_bodies():
    train = [closure.coarse_fine_pair(seed) for seed in range(10)]
    held = [closure.coarse_fine_pair(seed) for seed in (101, 102, 103)]
    x = np.concatenate([p["features"] for p in train])
    t = np.concatenate([p["target"] for p in train])
    net, norm, history = closure.train_closure(x, t, steps=3000, batch=512, lr=2e-3, seed=5)
    try:
        assert history[-100:].mean() < 0.5 * history[:20].mean()
        xh = np.concatenate([p["features"] for p in held])
        th = np.concatenate([p["target"] for p in held])
        law = np.concatenate([p["law"] for p in held])
        coarse = np.concatenate([p["coarse_stress"] for p in held])
        pred = closure.predict(net, norm, xh)
        rms = lambda a: float(np.sqrt(np.mean(np.sum(a ** 2, axis=1))))
        e_net, e_law, e_coarse = rms(pred - th), rms(law - th), rms(coarse - th)
        print(f"held-out wall shear RMS error: network {e_net:.3e}, wall law {e_law:.3e}, "
              f"coarse wall-law run {e_coarse:.3e}, target RMS {rms(th):.3e}, samples {len(th)}")
        assert e_net < 0.8 * e_law
        # The closure in coarse runs of a held-out body, evaluated every step and every 8 steps.
        flow = drag.flow_through_steps(48, U_IN)
        fx = {}
        for every in (1, 8):
            solver, grid = body_solver(wall_model="learned", seed=101, nu=0.004)
            try:
                solver.set_closure(net, norm)
                solver.set_closure_interval(every)
                samples = solver.step_sampled(4 * flow, 8)
                used = solver.closure_output()[:, 3]
                print(f"closure every {every} steps: used on {used.mean():.1%} of {used.size} wall cells")
                assert used.mean() > 0.8
                assert np.all(np.isfinite(samples)) and np.all(np.isfinite(solver.read_macro()))
                fx[every] = drag.window_mean_force(samples, 4 * flow, flow)[0]
            finally:
                solver.destroy()
        print(f"mean drag force: every step {fx[1]:.6e}, every 8 steps {fx[8]:.6e}")
        # The interval's effect is about 0.2 % and moves with rounding-level changes of the run;
        # the bar is half of the +-1 % Cd budget the interval spends from.
        assert abs(fx[8] / fx[1] - 1) < 5e-3
    finally:
        net.destroy()
