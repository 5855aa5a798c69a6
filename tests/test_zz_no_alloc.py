"""pw_solver_step_sampled_noalloc (the sampled stepper whose force history lives in a fixed
device array of the library instead of a per-call allocation) against
pw_solver_step_sampled, called through the C-ABI: two calls in a row, the second asking for
fewer samples, bit-identical records (the same kernels in the same order); a call beyond the
4096-sample limit is rejected before it steps."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from pocketwind_amd import _native as N
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from grids import blob_kinds, grid_from_kinds

pytestmark = pytest.mark.gpu


def sampled(fn, solver, count, every):
    first = (solver.step_count // every + 1) * every
    samples = max(0, (solver.step_count + count - first) // every + 1)
    out = np.full(4 * samples, np.nan)
    status = getattr(N.lib(), fn)(solver._live(), count, every, out.ctypes.data, out.size)
    assert status == N.PW_OK, N.last_error()
    return out.reshape(-1, 4)


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_noalloc_sampled_steps_equal_the_allocating_entry_point(precision):
    layout = (64, 32, 32)
    grid = grid_from_kinds(layout, blob_kinds(layout))
    params = SolverParams(0.8, 0.06, True, operator="reg", precision=precision)
    a = GpuSolver.create(layout, grid, params)
    b = GpuSolver.create(layout, grid, params)
    try:
        for s in (a, b):
            s.set_variant("cell-nt-stores", 128, 16)
        for count, every in ((200, 8), (37, 5)):  # the second call holds fewer samples
            ref = sampled("pw_solver_step_sampled", a, count, every)
            got = sampled("pw_solver_step_sampled_noalloc", b, count, every)
            print(precision, count, every, ref.shape[0], "samples, last", ref[-1])
            assert ref.shape == got.shape and ref.shape[0] > 0
            assert np.array_equal(ref.view(np.uint64), got.view(np.uint64))
            assert a.step_count == b.step_count
        assert np.array_equal(a.read_populations(), b.read_populations())
    finally:
        a.destroy()
        b.destroy()


def test_noalloc_rejects_more_samples_than_its_array_holds():
    layout = (64, 32, 32)
    grid = grid_from_kinds(layout, blob_kinds(layout))
    s = GpuSolver.create(layout, grid, SolverParams(0.8, 0.06, False))
    try:
        out = np.zeros(4 * 4097)
        status = N.lib().pw_solver_step_sampled_noalloc(s._live(), 4097, 1, out.ctypes.data, out.size)
        assert status == N.PW_EINVAL and "4096" in N.last_error()
        assert s.step_count == 0
        status = N.lib().pw_solver_step_sampled_noalloc(s._live(), 4096, 1, out.ctypes.data, out.size)
        assert status == N.PW_OK and s.step_count == 4096
        assert np.array_equal(out[:4 * 4096:4][:3], [1, 2, 3]) and np.all(np.isfinite(out[:4 * 4096]))
    finally:
        s.destroy()
