"""The learned closure evaluated beside the step in the step's own launch (placement "alongside")
against the overlapped placement, whose schedule it shares and which
test_gpu_closure_overlapped.py pins to its definition: the closure of step k (k a multiple of the
interval) reads state k in the first blocks of step k's launch, step k uses the previous output
and steps k + 1 .. k + E the new one. Populations, macros, closure outputs and force samples must
be bit-identical, on the row path and the per-cell path, at every block size the step launches
with (the closure's tiles map to the blocks' wavefronts) and across step-call boundaries."""
from __future__ import annotations

import os

import numpy as np
import pytest

from pocketwind_amd import bodies, closure
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu

NET = os.path.join(os.path.dirname(closure.__file__), "closure_c2.npz")
CALLS = (5, 1, 8, 3, 7)


def body_grid(layout, kind):
    mesh = {"sphere": bodies.sphere_mesh, "bluff": bodies.bluff_body}[kind]()
    body, _ = bodies.place_mesh(mesh, None, layout[0] / 4, layout, False)
    return voxelize(body, None, layout, False, link_fractions=True)


def params(precision):
    return SolverParams(0.53, 0.05, True, operator="reg", precision=precision, wall_model="learned")


def words(a: np.ndarray) -> np.ndarray:
    return a.view(np.uint16 if a.itemsize == 2 else np.uint32)


def run(grid, precision, placement, every, net, norm, variant=None):
    s = GpuSolver.create(grid.layout, grid, params(precision))
    try:
        if variant:
            s.set_variant(*variant)
        s.set_closure(net, norm)
        s.set_closure_interval(every)
        s.set_closure_placement(placement)
        for count in CALLS:
            s.step(count, False)
        s.step(1, True)
        return (words(s.read_populations()).copy(), s.read_macro().view(np.uint32).copy(),
                s.closure_output().copy(), int(s.wall_cells().size))
    finally:
        s.destroy()


@pytest.mark.parametrize("precision,layout,kind,every,variant", [
    ("f16", (128, 64, 64), "bluff", 4, None),
    ("f16", (128, 64, 64), "bluff", 1, ("cell-nt-stores", 64, 16)),
    ("f32", (128, 64, 64), "bluff", 3, ("cell-nt", 128, 1)),
    ("bf16", (128, 64, 64), "sphere", 2, ("cell", 256, 64)),
    ("f16", (96, 48, 48), "sphere", 4, None),  # nx = 96: the per-cell step
])
def test_alongside_closure_equals_overlapped(precision, layout, kind, every, variant):
    net, norm, _ = closure.load(NET)
    try:
        grid = body_grid(layout, kind)
        along = run(grid, precision, "alongside", every, net, norm, variant)
        over = run(grid, precision, "overlapped", every, net, norm, variant)
    finally:
        net.destroy()
    on_network = int(np.sum(along[2][:, 3] > 0))
    differ = int(np.count_nonzero(along[0] != over[0]))
    print(f"{precision} {kind} {layout} every {every} {variant}: {on_network} of {along[3]} wall cells on the "
          f"network; differing population words {differ}")
    assert on_network > 0.25 * along[3]
    assert differ == 0
    assert np.array_equal(along[1], over[1])
    assert np.array_equal(along[2], over[2])


def test_alongside_closure_sampled_forces_equal_across_call_boundaries():
    """Force samples every 8 steps through step_sampled equal those of single-step calls."""
    net, norm, _ = closure.load(NET)
    grid = body_grid((128, 64, 64), "bluff")
    runs = []
    try:
        for split in (False, True):
            s = GpuSolver.create(grid.layout, grid, params("f16"))
            try:
                s.set_closure(net, norm)
                s.set_closure_interval(8)
                s.set_closure_placement("alongside")
                if split:
                    samples = []
                    for _ in range(5):
                        for _ in range(7):
                            s.step(1, False)
                        samples.append(s.step_sampled(1, 1))
                    samples = np.concatenate(samples)
                else:
                    samples = s.step_sampled(40, 8)
                runs.append((samples, words(s.read_populations()).copy()))
            finally:
                s.destroy()
    finally:
        net.destroy()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
