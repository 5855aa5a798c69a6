"""The learned closure overlapped with the step (placement "overlapped") against its definition
restated through the C-ABI: the closure kernel of step k (k a multiple of the interval E) reads
state k on a second stream beside step k, step k still uses the previous output, and steps
k + 1 .. k + E use the new one; the first step evaluates in line and its output also serves
step E.

The restatement steps a second solver with the wall stresses prescribed from the host
(pw_debug_set_wall_stress, the same path through the wall collision as a network's output) and
takes each output from a third solver loaded with the state in question (its closure kernel
evaluates the network on the state its next step reads). Populations, macros and force samples
must be bit-identical, whatever the step calls' boundaries."""
from __future__ import annotations

import os

import numpy as np
import pytest

from pocketwind_amd import bodies, closure
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu

NET = os.path.join(os.path.dirname(closure.__file__), "closure_c2.npz")
EVERY = 4
CALLS = (5, 1, 8, 3)  # 17 steps: evaluations at 0 (in line), 4, 8, 12 and 16


def body_grid(layout, kind):
    mesh = {"sphere": bodies.sphere_mesh, "bluff": bodies.bluff_body}[kind]()
    body, _ = bodies.place_mesh(mesh, None, layout[0] / 4, layout, False)
    return voxelize(body, None, layout, False, link_fractions=True)


def params(precision):
    return SolverParams(0.53, 0.05, True, operator="reg", precision=precision, wall_model="learned")


def words(a: np.ndarray) -> np.ndarray:
    return a.view(np.uint16 if a.itemsize == 2 else np.uint32)


def evaluate(evaluator: GpuSolver, state: np.ndarray) -> np.ndarray:
    """The network's output on `state`: the evaluator's closure kernel runs before its step."""
    evaluator.write_populations(state)
    evaluator.step(1, False)
    return evaluator.closure_output()


@pytest.mark.parametrize("precision,layout,kind", [
    ("f32", (128, 64, 64), "bluff"),
    ("f16", (128, 64, 64), "bluff"),
    ("bf16", (96, 48, 48), "sphere"),
])
def test_overlapped_closure_matches_its_definition(precision, layout, kind):
    net, norm, _ = closure.load(NET)
    grid = body_grid(layout, kind)
    total = sum(CALLS)
    overlapped = GpuSolver.create(grid.layout, grid, params(precision))
    restated = GpuSolver.create(grid.layout, grid, params(precision))
    evaluator = GpuSolver.create(grid.layout, grid, params(precision))
    try:
        overlapped.set_closure(net, norm)
        overlapped.set_closure_interval(EVERY)
        overlapped.set_closure_placement("overlapped")
        evaluator.set_closure(net, norm)
        evaluator.set_closure_interval(1)
        for count in CALLS:
            overlapped.step(count, False)
        overlapped.step(1, True)
        used_on_network = []
        # The restatement: output T_k of state k serves steps k + 1 .. k + E (T_0 also step 0).
        out = evaluate(evaluator, restated.read_populations())
        used_on_network.append(int(np.sum(out[:, 3] > 0)))
        restated.prescribe_wall_stress(out[:, :3], out[:, 3] > 0)
        while restated.step_count < total + 1:
            k = restated.step_count
            if k > 0 and k % EVERY == 0:
                fresh = evaluate(evaluator, restated.read_populations())
                restated.step(1, k + 1 == total + 1)  # step k keeps the previous output
                out = fresh
                used_on_network.append(int(np.sum(out[:, 3] > 0)))
                restated.prescribe_wall_stress(out[:, :3], out[:, 3] > 0)
            else:
                restated.step(1, k + 1 == total + 1)
        a, b = words(overlapped.read_populations()), words(restated.read_populations())
        differ = int(np.count_nonzero(a != b))
        print(f"{precision} {kind}: wall cells on the network per evaluation {used_on_network} "
              f"of {overlapped.wall_cells().size}; differing population words {differ}")
        # The uniform start leaves most features out of range; a few steps in, most cells are on.
        assert used_on_network[-1] > 0.25 * overlapped.wall_cells().size
        assert differ == 0
        assert np.array_equal(overlapped.read_macro().view(np.uint32), restated.read_macro().view(np.uint32))
        assert np.array_equal(overlapped.closure_output(), out)
    finally:
        for s in (overlapped, restated, evaluator):
            s.destroy()
        net.destroy()


def test_overlapped_closure_sampled_forces_equal_across_call_boundaries():
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
                s.set_closure_placement("overlapped")
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
