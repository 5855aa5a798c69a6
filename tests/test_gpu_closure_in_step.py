"""The learned closure evaluated inside the step's wall wavefronts (placement "step") against
its own kernel over the wall-cell list (placement "kernel", every step). Both evaluate the same
features with the same network on the state the step reads, and a product's rows do not mix,
so they must give the same stresses and the same states bit for bit: for a single solver in
every storage format, and for a batch of a row tunnel (in-step) and a per-cell tunnel (which
keeps the kernel), whose row tunnel also equals the same tunnel stepped alone in-step."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from pocketwind_amd import _native as N
from pocketwind_amd import bodies, closure
from pocketwind_amd.batch import TunnelBatch
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu

NET = os.path.join(os.path.dirname(closure.__file__), "closure_c2.npz")
STEPS = 60


def body_grid(layout, kind="bluff"):
    mesh = {"sphere": bodies.sphere_mesh, "bluff": bodies.bluff_body}[kind]()
    body, _ = bodies.place_mesh(mesh, None, layout[0] / 4, layout, False)
    return voxelize(body, None, layout, False, link_fractions=True)


def params(precision):
    return SolverParams(0.53, 0.05, True, operator="reg", precision=precision, wall_model="learned")


def words(a: np.ndarray) -> np.ndarray:
    return a.view(np.uint16 if a.itemsize == 2 else np.uint32)


@pytest.mark.parametrize("precision", ["f32", "f16", "bf16"])
def test_in_step_closure_equals_the_closure_kernel(precision):
    net, norm, _ = closure.load(NET)
    grid = body_grid((128, 64, 64))
    runs = {}
    try:
        for placement in ("step", "kernel"):
            s = GpuSolver.create(grid.layout, grid, params(precision))
            try:
                s.set_closure(net, norm)
                s.set_closure_interval(1)
                s.set_closure_placement(placement)
                samples = s.step_sampled(STEPS, 8)
                s.step(1, True)
                runs[placement] = (samples, words(s.read_populations()).copy(), s.read_macro(),
                                   int(np.sum(s.closure_output()[:, 3] > 0)), s.wall_cells().size)
            finally:
                s.destroy()
    finally:
        net.destroy()
    step, kernel = runs["step"], runs["kernel"]
    differ = int(np.count_nonzero(step[1] != kernel[1]))
    print(f"{precision}: {step[3]} of {step[4]} wall cells on the network; differing population words {differ}")
    assert step[3] > 0.9 * step[4]
    assert differ == 0
    assert np.array_equal(step[0], kernel[0])
    assert np.array_equal(step[2].view(np.uint32), kernel[2].view(np.uint32))


def test_batch_in_step_closure_equals_kernel_and_single():
    net, norm, _ = closure.load(NET)
    grids = [body_grid((128, 64, 64)), body_grid((96, 48, 48), "sphere")]
    shared = params("f16")
    out = {}
    try:
        for placement in ("step", "kernel"):
            b = TunnelBatch.create(grids, [0.53] * 2, [0.05] * 2, shared)
            try:
                b.set_closure(net, norm)
                b.set_closure_interval(1)
                b.set_closure_placement(placement)
                samples = b.step_sampled(STEPS, 8)
                b.step(1, True)
                out[placement] = (samples, [words(b.read_populations(k)).copy() for k in range(2)])
            finally:
                b.destroy()
        s = GpuSolver.create(grids[0].layout, grids[0], shared)
        try:
            s.set_closure(net, norm)
            s.set_closure_interval(1)
            s.set_closure_placement("step")
            single = s.step_sampled(STEPS, 8)
            s.step(1, True)
            alone = words(s.read_populations()).copy()
        finally:
            s.destroy()
    finally:
        net.destroy()
    for k in range(2):
        differ = int(np.count_nonzero(out["step"][1][k] != out["kernel"][1][k]))
        print(f"tunnel {k}: differing population words {differ}")
        assert differ == 0
    assert np.array_equal(out["step"][0], out["kernel"][0])
    assert np.array_equal(out["step"][1][0], alone)
    assert np.array_equal(out["step"][0][0], single)


def test_placement_is_validated():
    grid = body_grid((64, 32, 32))
    s = GpuSolver.create(grid.layout, grid, params("f16"))
    try:
        with pytest.raises(ValueError):
            s.set_closure_placement("wall")
        assert N.lib().pw_solver_set_closure_placement(s._live(), 4) == N.PW_EINVAL
    finally:
        s.destroy()
