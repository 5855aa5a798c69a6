"""The C-ABI library loads, exports every symbol include/pw.h declares, and rejects invalid
input on the host before any device work (gpu-solver.ts:164-181, geometry-contract.ts:39-66,
gpu-force.ts:76-81). No kernel is launched here."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from grids import blob_grid, grid_from_kinds
from pocketwind_amd import _native as N
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams, _pack
from pocketwind_amd.lattice import NO_Q, SOLID, VoxelGrid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pw.h")).read()
    return sorted(set(re.findall(r"\b(pw_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(N.SIGNATURES), set(names) - set(N.SIGNATURES)
    assert lib.pw_abi_version() == 8


def _create(grid, params, boundary=None):
    geo = np.ascontiguousarray(grid.geometry, np.uint32)
    bc = np.ascontiguousarray(grid.boundary_cells if boundary is None else boundary, np.uint32)
    out = C.c_void_p()
    status = N.lib().pw_solver_create(C.byref(N.Layout(grid.nx, grid.ny, grid.nz)),
                                      geo.ctypes.data_as(N._u32p), geo.size,
                                      bc.ctypes.data_as(N._u32p), bc.size, C.byref(params),
                                      C.byref(out))
    return status, N.last_error()


def _params(**kw):
    base = dict(tau0=0.8, u_in=0.06, les=0, op=0, walls=0, storage=0, ramp_steps=0, wall_model=0)
    base.update(kw)
    return N.Params(**base)


def test_rejects_tau_at_or_below_half_after_f32_rounding():
    grid = blob_grid("bb")
    for tau in (0.5, 0.5 + 1e-9, float("nan"), 0.2):
        status, msg = _create(grid, _params(tau0=tau))
        assert status == N.PW_EINVAL and "Relaxation time" in msg


def test_rejects_inconsistent_wall_mask():
    grid = blob_grid("bb")
    geo = grid.geometry.copy()
    fluid = int(np.flatnonzero(geo[: grid.cells] == 0)[0])
    geo[fluid] |= 1 << 4
    bad = VoxelGrid(grid.nx, grid.ny, grid.nz, geo, grid.boundary_cells)
    status, msg = _create(bad, _params())
    assert status == N.PW_EINVAL and "Inconsistent wall mask" in msg


def test_rejects_bad_q_offsets_and_length():
    grid = blob_grid("bouzidi")
    geo = grid.geometry.copy()
    cells = grid.cells
    with_q = int(np.flatnonzero(geo[cells:2 * cells] != NO_Q)[1])
    geo[cells + with_q] += 1
    status, msg = _create(VoxelGrid(32, 32, 32, geo, grid.boundary_cells), _params(walls=1))
    assert status == N.PW_EINVAL and "q offset" in msg
    short = VoxelGrid(32, 32, 32, grid.geometry[:-1].copy(), grid.boundary_cells)
    status, msg = _create(short, _params(walls=1))
    assert status == N.PW_EINVAL and "length" in msg


def test_rejects_boundary_cells_without_wall_links():
    grid = blob_grid("bb")
    status, msg = _create(grid, _params(), boundary=np.array([0], np.uint32))
    assert status == N.PW_EINVAL and "Boundary cells" in msg
    status, msg = _create(grid, _params(), boundary=np.array([grid.cells], np.uint32))
    assert status == N.PW_EINVAL


def test_rejects_unknown_options_and_kinds():
    grid = blob_grid("bb")
    assert _create(grid, _params(storage=7))[0] == N.PW_EINVAL
    geo = grid.geometry.copy()
    geo[0] = 5
    assert _create(VoxelGrid(32, 32, 32, geo, grid.boundary_cells), _params())[0] == N.PW_EINVAL


def test_host_mirror_raises_value_error_like_range_error():
    grid = blob_grid("bb")
    with pytest.raises(ValueError, match="Relaxation time"):
        GpuSolver.create(grid.layout, grid, SolverParams(0.5, 0.06, False))
    with pytest.raises(ValueError):
        GpuSolver.create(grid.layout, grid, SolverParams(0.8, 0.06, False, operator="bgk"))
    with pytest.raises(ValueError):
        GpuSolver.create((16, 32, 32), grid, SolverParams(0.8, 0.06, False))


def test_boundary_options_are_checked_on_the_host():
    """Free-slip sides need walls two cells clear of the sides; the learned closure and
    batched tunnels take the reference's boundaries only; a zero-gradient outlet needs nx >= 2."""
    blob = blob_grid("bb")  # its SUPPORT/FLOOR pair and blob sit well inside the 32^3 box
    assert _create(blob, _params(sides=2))[0] == N.PW_EINVAL
    assert _create(blob, _params(outlet=2))[0] == N.PW_EINVAL
    status, msg = _create(blob, _params(les=1, wall_model=1, sides=1))
    assert status == N.PW_EINVAL and "periodic sides" in msg
    kinds = np.zeros((8, 8, 8), np.uint32)
    kinds[1, 4, 4] = SOLID  # one cell from the z = 0 side
    near = grid_from_kinds((8, 8, 8), kinds.reshape(-1))
    status, msg = _create(near, _params(sides=1))
    assert status == N.PW_EINVAL and "two cells from the sides" in msg
    line = grid_from_kinds((1, 4, 4), np.zeros(16, np.uint32))
    status, msg = _create(line, _params(outlet=1))
    assert status == N.PW_EINVAL and "nx >= 2" in msg
    with pytest.raises(ValueError):
        _pack(SolverParams(0.8, 0.06, False, sides="closed"))


def test_batched_tunnels_take_the_reference_boundaries_only():
    from pocketwind_amd.batch import TunnelBatch
    grid = blob_grid("bb")
    with pytest.raises(ValueError, match="periodic"):
        TunnelBatch.create([grid], [0.8], [0.06], SolverParams(0.8, 0.06, False, sides="free-slip"))


def test_batch_closure_particles_and_force_entry_points_reject_null_handles():
    lib = N.lib()
    norm = np.zeros(27, np.float32)
    out = C.c_void_p()
    lo = (C.c_double * 3)(1, 1, 1)
    hi = (C.c_double * 3)(2, 2, 2)
    assert lib.pw_batch_set_closure(None, None, norm.ctypes.data, 27) == N.PW_EINVAL
    assert lib.pw_batch_set_closure_interval(None, 8) == N.PW_EINVAL
    assert lib.pw_particles_create_batch(None, 0, 16, 1, lo, hi, C.byref(out)) == N.PW_EINVAL
    assert out.value is None
    buf = np.zeros(3)
    assert lib.pw_batch_force(None, buf.ctypes.data, 3) == N.PW_EINVAL


def test_unknown_or_destroyed_handles_are_rejected_without_being_read():
    """A handle that pw_solver_create did not return (or that was destroyed) is looked up in
    the library's set of live handles, never dereferenced: every call returns PW_ESTATE and
    the getters return 0."""
    lib = N.lib()
    stale = C.c_void_p(0x1000)  # never a live handle; dereferencing it would fault the process
    assert lib.pw_solver_step(stale, 1, 0) == N.PW_ESTATE
    assert "destroyed" in N.last_error()
    assert lib.pw_solver_reset(stale) == N.PW_ESTATE
    assert lib.pw_solver_encode_force(stale) == N.PW_ESTATE
    assert lib.pw_solver_set_step_variant(stale, 1, 256, 1) == N.PW_ESTATE
    assert lib.pw_solver_step_count(stale) == 0
    assert lib.pw_solver_population_bytes(stale) == 0
    assert lib.pw_solver_wall_cell_count(stale) == 0
    lib.pw_solver_destroy(stale)  # ignored
    out = C.c_void_p()
    assert lib.pw_particles_create(stale, 1, 1, (C.c_double * 3)(0, 0, 0), (C.c_double * 3)(1, 1, 1),
                                   C.byref(out)) == N.PW_EINVAL


def test_grid_from_kinds_matches_the_per_cell_wall_mask():
    """The packed flags of grid_from_kinds (rolled wall fields) equal force.spec.ts's per-cell
    loop over the 19 pulls with periodic wrap (grids.pull_wall_mask) on the seeded blob, whose
    SUPPORT and FLOOR cells count as walls, and the analytic C1 sphere has the solid and boundary
    counts the oracle's golden run recorded."""
    import json

    from grids import blob_kinds, pull_wall_mask, sphere_kinds
    from pocketwind_amd.lattice import FLUID
    layout = (32, 32, 32)
    kinds = blob_kinds(layout)
    g = grid_from_kinds(layout, kinds)
    for cell in range(kinds.size):
        expected = pull_wall_mask(kinds, layout, cell) if kinds[cell] == FLUID else int(kinds[cell])
        assert int(g.geometry[cell]) == expected, cell
    assert np.all(g.geometry[kinds.size:] == NO_Q)
    fluid_walled = [c for c in range(kinds.size) if kinds[c] == FLUID and g.geometry[c] >> 3]
    assert g.boundary_cells.tolist() == fluid_walled
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_sphere_oracle.json")))
    sphere = grid_from_kinds(golden["layout"], sphere_kinds(golden["layout"], golden["centre"], golden["diameter"]))
    assert sphere.solid_cells == golden["solid_cells"]
    assert sphere.boundary_cells.size == golden["boundary_cells"]
