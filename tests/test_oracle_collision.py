"""Pins the oracle's collision math (restates tests/unit/reference/collision.test.ts).

The checks are independent of the formulas they test: Smagorinsky tau against a bisection
on nu_eff = nu0 + Cs^2 |S| with Pi = -2 rho cs^2 tau S; REG equilibrium as a fixed point
and its mass/momentum conservation; the wall-law branches, their continuity; wall-cell
collision keeps moments and gives the wall the wall-law shear along the slip.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O

CS2 = 1 / 3
CS = 0.1


def expected_relaxation(stress, rho, tau0):
    low, high = tau0, tau0 + 1
    for _ in range(100):
        tau = (low + high) / 2
        strain = [-c / (2 * rho * CS2 * tau) for row in stress for c in row]
        mag = math.sqrt(2 * sum(c * c for c in strain))
        nu = (tau - 0.5) * CS2
        target = (tau0 - 0.5) * CS2 + CS ** 2 * mag
        if nu > target:
            high = tau
        else:
            low = tau
    return (low + high) / 2


STRESSES = {
    "normal": [[0.02, 0, 0], [0, 0, 0], [0, 0, 0]],
    "xy": [[0, 0.02, 0], [0.02, 0, 0], [0, 0, 0]],
    "xz": [[0, 0, -0.02], [0, 0, 0], [-0.02, 0, 0]],
    "yz": [[0, 0, 0], [0, 0, 0.02], [0, 0.02, 0]],
    "mixed": [[0.02, -0.007, 0.011], [-0.007, -0.013, 0.005], [0.011, 0.005, 0.009]],
}


@pytest.mark.parametrize("tau0", [0.6, 0.8, 1.5])
@pytest.mark.parametrize("name", list(STRESSES))
def test_smagorinsky_relaxation_matches_bisection(name, tau0):
    tensor = STRESSES[name]
    rho = 1.02
    geq = O.equilibrium(rho, (0, 0, 0))
    pert = np.array([
        O.W19[i] * sum((O.C19[i][a] * O.C19[i][b] - (CS2 if a == b else 0)) * tensor[a][b]
                       for a in range(3) for b in range(3)) / (2 * CS2 ** 2)
        for i in range(19)])
    g = geq + pert
    tau = expected_relaxation(tensor, rho, tau0)
    out = O.collide(g, geq, rho, tau0, O.TRT, True)
    np.testing.assert_allclose(out, geq + pert * (1 - 1 / tau), rtol=0, atol=5e-16)
    assert tau > tau0


@pytest.mark.parametrize("les", [False, True])
def test_reg_leaves_equilibrium_unchanged(les):
    geq = O.equilibrium(1.04, (0.03, -0.02, 0.01))
    out = O.collide(geq, geq, 1.04, 0.8, O.REG, les)
    assert np.array_equal(out, geq)


def test_reg_conserves_mass_and_momentum():
    rho = 1.02
    geq = O.equilibrium(rho, (0.05, -0.03, 0.02))
    c = O.C19
    shape = (c[:, 0] * c[:, 1] + 0.5 * (c[:, 0] ** 2 - c[:, 1] ** 2) + 0.5 * c[:, 0] * c[:, 2]) * 0.02
    g = geq + O.W19 * shape
    out = O.collide(g, geq, rho, 0.8, O.REG, False)
    assert abs(out.sum() - g.sum()) < 5e-16
    np.testing.assert_allclose(c.T @ out, c.T @ g, atol=5e-16)
    assert np.any(np.abs(out - g) > 1e-9)


def test_wall_law_branches():
    viscous = O.friction_velocity(1e-4, 1e-3)
    assert abs(viscous - math.sqrt(1e-3 * 1e-4 / 0.5)) < 5e-16
    assert 0.5 * viscous / 1e-3 < 11.81
    turbulent = O.friction_velocity(0.05, 1e-6)
    y_plus = 0.5 * turbulent / 1e-6
    assert y_plus > 11.81
    assert abs(0.05 / turbulent - 8.3 * y_plus ** (1 / 7)) < 5e-10


def test_wall_law_is_continuous_at_the_junction():
    nu = 1e-4
    speed = 11.81 ** 2 * nu / 0.5
    below = O.friction_velocity(speed * (1 - 1e-9), nu)
    above = O.friction_velocity(speed * (1 + 1e-9), nu)
    assert abs(above / below - 1) < 5e-5


FROM_FLOOR = [5, 9, 11, 16, 18]
STATE_RHO, STATE_U = 1.002, (0.04, -0.01, 0.002)


def wall_cell(u=STATE_U, incoming=FROM_FLOOR):
    geq = O.equilibrium(STATE_RHO, u)
    g = geq + 1e-4 * np.sin(np.arange(19))
    applied, out = O.collide_wall(g, geq, STATE_RHO, u, incoming, 1e-6)
    return g, geq, applied, out


def test_wall_cell_keeps_mass_and_momentum():
    _, _, applied, out = wall_cell()
    assert applied
    rho, u = O.moments(out)
    assert abs(rho - STATE_RHO) < 5e-15
    np.testing.assert_allclose(u, STATE_U, atol=5e-15)


def test_wall_cell_gives_wall_law_shear_and_keeps_normal_flux():
    _, geq, _, out = wall_cell()

    def flux(p):
        s = np.zeros(3)
        for i in FROM_FLOOR:
            j = i + 1 if i % 2 == 1 else i - 1
            s += O.C19[j] * 2 * (p[j] + O.W19[j])
        return s

    ux, uy, _ = STATE_U
    speed = math.hypot(ux, uy)
    shear = STATE_RHO * O.friction_velocity(speed, 1e-6) ** 2
    fx, fy, fz = flux(out)
    assert abs(fx - shear * ux / speed) < 5e-15
    assert abs(fy - shear * uy / speed) < 5e-15
    assert abs(fz - flux(geq)[2]) < 5e-15


@pytest.mark.parametrize("u,incoming", [((0, 0, 0.01), FROM_FLOOR), (STATE_U, [3, 5])])
def test_wall_cell_falls_back_without_slip_or_flux(u, incoming):
    g, _, applied, out = wall_cell(u, incoming)
    assert not applied
    assert np.array_equal(out, g)


def test_wall_cell_falls_back_on_inadmissible_pore():
    geq = O.equilibrium(1.0, (0.05, 0, 0))
    applied, out = O.collide_wall(geq, geq, 1.0, (0.05, 0, 0), [1, 2, 4, 8], 0.01)
    assert not applied and np.array_equal(out, geq)


def test_equilibrium_moments_round_trip():
    for rho, u in ((1.0, (0, 0, 0)), (1.03, (0.02, -0.01, 0.005)), (0.97, (-0.05, 0.04, 0.01))):
        g = O.equilibrium(rho, u)
        r, v = O.moments(g)
        assert abs(r - rho) < 1e-15
        np.testing.assert_allclose(v, u, atol=1e-16)


def test_opposite_pairs():
    for i in range(19):
        j = O.opp(i)
        assert np.array_equal(O.C19[j], -O.C19[i])
        assert O.W19[i] == O.W19[j]
