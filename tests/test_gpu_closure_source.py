"""The wall collision's momentum-source mode (pw_debug_set_wall_stress with w = 2): after the wall
law's collision the wall cell's populations gain 3 w_i c_i . C, so the cell gains the momentum C
and no mass. One step from the same state with and without the sources must differ by exactly
that at the wall cells (to f32 rounding of the sum) and not at all elsewhere; a source of zero
must leave the step bit-identical to the wall law's."""
from __future__ import annotations

import numpy as np
import pytest

from pocketwind_amd import bodies
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.lattice import DIRECTIONS, WEIGHTS
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu

LAYOUT = (64, 32, 32)


def grid():
    body, _ = bodies.place_mesh(bodies.sphere_mesh(), None, 12, LAYOUT, False)
    return voxelize(body, None, LAYOUT, False, link_fractions=True)


def solver(g):
    return GpuSolver.create(g.layout, g, SolverParams(0.53, 0.05, True, operator="reg", wall_model="learned"))


def test_momentum_sources_add_their_momentum_and_no_mass():
    g = grid()
    plain, sourced, zero = solver(g), solver(g), solver(g)
    try:
        for s in (plain, sourced, zero):
            s.step(30, False)  # a developed state, the same in all three
        cells = plain.wall_cells()
        rng = np.random.default_rng(5)
        c = rng.normal(0.0, 1e-4, (cells.size, 3)).astype(np.float32)
        plain.prescribe_wall_stress(np.zeros_like(c), np.zeros(cells.size, bool))
        sourced.prescribe_wall_stress(c, np.ones(cells.size, bool), source=True)
        zero.prescribe_wall_stress(np.zeros_like(c), np.ones(cells.size, bool), source=True)
        for s in (plain, sourced, zero):
            s.step(1, False)
        a = plain.read_populations().astype(np.float64)
        b = sourced.read_populations().astype(np.float64)
        assert np.array_equal(zero.read_populations(), plain.read_populations())
        diff = b - a
        others = np.ones(plain.cells, bool)
        others[cells] = False
        assert np.all(diff[:, others] == 0)
        ci = np.asarray(DIRECTIONS, np.float64)  # (19, 3)
        w = np.asarray(WEIGHTS, np.float64)
        expected = 3.0 * w[:, None] * (ci @ c.astype(np.float64).T)  # (19, n_wall)
        got = diff[:, cells]
        assert np.abs(got - expected).max() < 1e-7
        # momentum and mass of the added part
        assert np.allclose(ci.T @ got, c.T, atol=1e-7)
        assert np.abs(got.sum(0)).max() < 1e-7
    finally:
        for s in (plain, sourced, zero):
            s.destroy()


def test_momentum_sources_need_f32_storage():
    g = grid()
    s = GpuSolver.create(g.layout, g, SolverParams(0.53, 0.05, True, operator="reg", precision="f16",
                                                   wall_model="learned"))
    try:
        n = s.wall_cells().size
        with pytest.raises(ValueError):
            s.prescribe_wall_stress(np.zeros((n, 3)), np.ones(n, bool), source=True)
        s.prescribe_wall_stress(np.zeros((n, 3)), np.ones(n, bool))  # shear vectors are fine
    finally:
        s.destroy()
