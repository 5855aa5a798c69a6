"""The oracle's free-slip sides and zero-gradient outlet (boundary options with no reference
counterpart; the reference is periodic in y and z with an equilibrium outlet,
src/lbm/shaders/common.wgsl:62-68, stream_collide.wgsl:82-107).

What pins them, without restating the source rule:
* free-slip (halfway specular reflection) is the restriction of a periodic lattice twice as
  wide in y and z whose second halves are the mirror images of the first (geometry mirrored,
  the mirrored velocity component negated). The doubled run uses the oracle's periodic sides,
  which the reference's own golden tests pin;
* the zero-gradient outlet makes the outlet plane's unknown populations (c_x = -1) equal the
  upstream neighbour's after streaming, and every other pull equal that of a plain fluid
  plane: checked on the pull table against the same lattice with a fluid outlet plane;
* uniform inflow through a channel with free-slip sides and a zero-gradient outlet is steady.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from grids import grid_from_kinds
from pocketwind_amd.lattice import DIRECTIONS, EQ, FLUID, SOLID

TAU0 = 0.56
U_IN = 0.05


def channel_kinds(layout, block) -> np.ndarray:
    """EQ planes x = 0 and x = nx - 1, a solid box (x0, x1, y0, y1, z0, z1) inclusive."""
    nx, ny, nz = layout
    kinds = np.zeros((nz, ny, nx), np.uint32)
    x0, x1, y0, y1, z0, z1 = block
    kinds[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1] = SOLID
    kinds[:, :, 0] = EQ
    kinds[:, :, nx - 1] = EQ
    return kinds


def random_state(layout, seed) -> np.ndarray:
    nx, ny, nz = layout
    rng = np.random.default_rng(seed)
    state = np.empty((nz, ny, nx, 4))
    state[..., 0] = 1 + 0.02 * rng.standard_normal((nz, ny, nx))
    state[..., 1:] = 0.03 * rng.standard_normal((nz, ny, nx, 3))
    return state


def mirrored(field: np.ndarray, flip_components: bool) -> np.ndarray:
    """(nz, ny, nx, ...) -> (2 nz, 2 ny, nx, ...): y' >= ny holds y = 2 ny - 1 - y', likewise
    z; with flip_components the y (z) velocity of a y- (z-) mirrored copy is negated."""
    y_img = field[:, ::-1].copy()
    if flip_components:
        y_img[..., 2] *= -1
    wide = np.concatenate([field, y_img], axis=1)
    z_img = wide[::-1].copy()
    if flip_components:
        z_img[..., 3] *= -1
    return np.concatenate([wide, z_img], axis=0)


def run(layout, kinds, state, operator, les, steps, sides="periodic", outlet="equilibrium"):
    cells = int(np.prod(layout))
    flags = grid_from_kinds(layout, kinds.reshape(-1)).geometry[:cells]
    cpu = oracle.Solver(layout, TAU0, operator, les, oracle.BB, flags=flags,
                        equilibrium_state=(1.0, (U_IN, 0.0, 0.0)),
                        initial=state.reshape(-1, 4), threads=4, sides=sides, outlet=outlet)
    cpu.step(steps)
    nx, ny, nz = layout
    return cpu.populations.copy().reshape(19, nz, ny, nx)


@pytest.mark.parametrize("outlet", ["equilibrium", "zero-gradient"])
@pytest.mark.parametrize("operator,les", [(oracle.TRT, False), (oracle.REG, True)])
def test_free_slip_is_the_mirror_image_of_a_doubled_periodic_lattice(operator, les, outlet):
    layout = (12, 6, 5)
    nx, ny, nz = layout
    kinds = channel_kinds(layout, (4, 6, 2, 3, 2, 2))
    state = random_state(layout, 7)
    slip = run(layout, kinds, state, operator, les, 12, sides="free-slip", outlet=outlet)

    wide_layout = (nx, 2 * ny, 2 * nz)
    wide_kinds = mirrored(kinds[..., None], False)[..., 0]
    wide_state = mirrored(state, True)
    wide = run(wide_layout, wide_kinds, wide_state, operator, les, 12, outlet=outlet)
    # The first octant of the doubled lattice is the free-slip lattice, direction by direction.
    np.testing.assert_allclose(slip, wide[:, :nz, :ny, :], rtol=0, atol=1e-13)
    # ... and the doubled run really is mirror symmetric (its y image holds mirrored directions).
    dirs = [tuple(int(v) for v in c) for c in DIRECTIONS]
    image = [dirs.index((c[0], -c[1], c[2])) for c in dirs]
    np.testing.assert_allclose(wide[image][:, :, ::-1][:, :nz, :ny], slip, rtol=0, atol=1e-13)
    # The flow crosses the sides' rows: a periodic run of the same lattice differs there.
    periodic = run(layout, kinds, state, operator, les, 12, outlet=outlet)
    assert np.abs(periodic - slip).max() > 1e-4


@pytest.mark.parametrize("sides", ["periodic", "free-slip"])
def test_zero_gradient_outlet_copies_the_upstream_unknown_populations(sides):
    layout = (8, 6, 5)
    nx, ny, nz = layout
    kinds = channel_kinds(layout, (3, 4, 2, 3, 2, 2))
    flags = grid_from_kinds(layout, kinds.reshape(-1)).geometry[:nx * ny * nz]
    zg = oracle.Solver(layout, TAU0, flags=flags, sides=sides, outlet="zero-gradient")
    fluid_outlet = kinds.copy()
    fluid_outlet[:, :, nx - 1] = FLUID
    plain_flags = grid_from_kinds(layout, fluid_outlet.reshape(-1)).geometry[:nx * ny * nz]
    plain = oracle.Solver(layout, TAU0, flags=plain_flags, sides=sides)
    n = nx * ny * nz
    for cell in range(n):
        x = cell % nx
        for i, c in enumerate(DIRECTIONS):
            if x == nx - 1 and c[0] == -1:
                # after streaming, outlet and upstream neighbour hold the same unknown value
                assert zg.source(cell, i) == zg.source(cell - 1, i) == plain.source(cell - 1, i)
                assert zg.source(cell, i) != plain.source(cell, i)
            else:
                assert zg.source(cell, i) == plain.source(cell, i), (cell, i)
    assert all(zg.source(nx - 1 + nx * row, 0) >= 0 for row in range(ny * nz))  # outlet cells are fluid


def test_uniform_inflow_is_steady_with_free_slip_sides_and_zero_gradient_outlet():
    layout = (10, 6, 5)
    kinds = channel_kinds(layout, (0, -1, 0, -1, 0, -1))  # empty box: no solid
    state = np.zeros(layout[::-1] + (4,))
    state[..., 0] = 1
    state[..., 1] = U_IN
    start = run(layout, kinds, state, oracle.TRT, False, 0, "free-slip", "zero-gradient")
    end = run(layout, kinds, state, oracle.REG, True, 60, "free-slip", "zero-gradient")
    np.testing.assert_allclose(end, start, rtol=0, atol=1e-15)


def test_options_are_validated():
    with pytest.raises(ValueError):
        oracle.Solver((1, 4, 4), TAU0, outlet="zero-gradient")
