"""Pins the oracle's solver: restates tests/unit/reference/cpu-solver.test.ts and the
analytic-flow cases of tests/unit/reference/cases.test.ts (Taylor-Green decay and order,
Poiseuille TRT/REG/BGK, Bouzidi second order and q = 1/2), each against closed forms."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O

WALLS = [O.SOLID, O.SUPPORT, O.FLOOR]


@pytest.mark.parametrize("kind", WALLS)
def test_walls_hold_rest_in_both_buffers(kind):
    flags = np.array([O.FLUID, kind, O.EQ], np.uint32)
    initial = np.tile([1.03, 0.02, -0.01, 0.005], 3)
    s = O.Solver((3, 1, 1), 0.8, flags=flags, equilibrium_state=(1.02, (0.05, 0, 0)),
                 initial=initial)
    for _ in range(5):
        rho, u = s.macroscopic(1)
        assert rho == 1 and np.all(u == 0)
        assert np.all(s.populations.reshape(19, 3)[:, 1] == 0)
        s.step()


@pytest.mark.parametrize("rho", [0, -1, float("nan"), float("inf")])
def test_rejects_invalid_density(rho):
    with pytest.raises(ValueError, match="density"):
        O.Solver((1, 1, 1), 0.8, initial=np.array([rho, 0, 0, 0.0]))
    with pytest.raises(ValueError, match="density"):
        O.Solver((1, 1, 1), 0.8, equilibrium_state=(rho, (0, 0, 0)))


def test_periodic_rest_stays_exactly_zero():
    s = O.Solver((4, 4, 4), 0.8)
    s.step()
    assert np.all(s.populations == 0)


@pytest.mark.parametrize("op", [O.TRT, O.BGK])
def test_uniform_moving_equilibrium_is_preserved_with_les(op):
    state = np.array([1.02, 0.02, -0.01, 0.005])
    s = O.Solver((4, 3, 2), 0.8, op, True, initial=np.tile(state, 24))
    s.step(10)
    m = s.macro_all()
    np.testing.assert_allclose(m, np.tile(state, (24, 1)), rtol=0, atol=5e-15)


def test_inlet_and_pressure_outlet():
    flags = np.array([O.EQ, O.FLUID, O.EQ], np.uint32)
    s = O.Solver((3, 1, 1), 0.8, flags=flags, equilibrium_state=(1, (0.05, 0, 0)),
                 initial=np.zeros(12) + np.tile([1, 0, 0, 0], 3), body_force=(0.001, 0, 0))
    for _ in range(5):
        _, upstream = s.macroscopic(1)
        s.step()
        rho0, u0 = s.macroscopic(0)
        assert abs(u0[0] - 0.05) < 5e-15 and abs(rho0 - 1) < 5e-15
        rho2, u2 = s.macroscopic(2)
        assert abs(rho2 - 1) < 5e-15
        np.testing.assert_allclose(u2, upstream, atol=5e-15)
    assert s.macroscopic(1)[1][0] > 0


def test_body_force_impulse_after_collision():
    a = np.array([1e-5, -2e-5, 3e-5])
    s = O.Solver((2, 2, 2), 0.8, body_force=a)
    s.step()
    pops = s.populations.reshape(19, 8)
    np.testing.assert_allclose(pops[:, 0], 3 * O.W19 * (O.C19 @ a), rtol=0, atol=5e-19)
    rho, u = s.macroscopic(0)
    assert abs(rho - 1) < 5e-15
    np.testing.assert_allclose(u, a, atol=5e-17)


@pytest.mark.parametrize("i", range(1, 19))
def test_pulls_each_direction_across_periodic_boundaries(i):
    s = O.Solver((5, 5, 5), 0.8)
    s.populations[i * s.cells + 0] = 0.001
    s.step()
    cx, cy, cz = O.C19[i]
    dest = (cx % 5) + 5 * ((cy % 5) + 5 * (cz % 5))
    m = s.macro_all()
    for cell in range(125):
        excess = 0.001 if cell == dest else 0.0
        assert abs(m[cell, 0] - (1 + excess)) < 5e-15
        np.testing.assert_allclose(m[cell, 1:], np.array([cx, cy, cz]) * excess / (1 + excess),
                                   atol=5e-15)


@pytest.mark.parametrize("kind", WALLS)
def test_momentum_exchange_known_value(kind):
    flags = np.zeros(125, np.uint32)
    solid = 2 + 5 * (2 + 5 * 2)
    flags[solid] = kind
    s = O.Solver((5, 5, 5), 0.8, flags=flags)
    assert np.linalg.norm(s.momentum_exchange()) < 1e-15
    s.populations[s.cells + solid - 1] = 0.01
    f = s.momentum_exchange()
    assert abs(f[0] - (0.01 if kind == O.SOLID else 0)) < 5e-16
    assert abs(f[1]) < 1e-15 and abs(f[2]) < 1e-15
    s.step()
    rho, u = s.macroscopic(solid - 1)
    assert abs(rho - 1.01) < 5e-15 and abs(u[0] + 0.01 / 1.01) < 5e-15


def test_paired_exchange_cancels_period_two_oscillation():
    flags = np.zeros(125, np.uint32)
    solid = 62
    flags[solid] = O.SOLID
    s = O.Solver((5, 5, 5), 0.8, flags=flags)
    s.populations[s.cells + solid - 1] = 0.01
    s.step()
    s.populations[s.cells + solid - 1] = -0.01
    assert abs(s.momentum_exchange()[0]) < 1e-15


def test_no_reference_pressure_on_a_body_cut_by_its_support():
    flags = np.zeros(125, np.uint32)
    flags[62] = O.SOLID
    flags[62 - 25] = O.SUPPORT
    s = O.Solver((5, 5, 5), 0.8, flags=flags)
    s.step(3)
    assert np.linalg.norm(s.momentum_exchange()) < 1e-15


def test_wall_function_gives_each_face_the_wall_law_shear():
    flags = np.zeros(64, np.uint32)
    flags[:16] = O.SOLID
    tau0 = 0.5 + 3e-5
    s = O.Solver((4, 4, 4), tau0, les=True, flags=flags,
                 initial=np.tile([1, 0.05, 0.01, 0], 64))

    def expected_shear():
        total = 0.0
        for z in (1, 3):
            for cell in range(16 * z, 16 * (z + 1)):
                rho, u = s.macroscopic(cell)
                speed = math.hypot(u[0], u[1])
                total += rho * O.friction_velocity(speed, 1e-5) ** 2 * u[0] / speed
        return total

    s.step()
    first = expected_shear()
    s.step()
    second = expected_shear()
    assert abs(s.momentum_exchange()[0] - (first + second) / 2) < 5e-13
    assert first > 0


def packed(solids, fractions, cells=125, fluid=61):
    geo = np.zeros(2 * cells + len(fractions), np.uint32)
    for c in solids:
        geo[c] = O.SOLID
    geo[cells:2 * cells] = O.NO_Q
    if fractions:
        geo[cells + fluid] = 0
    geo[2 * cells:] = np.asarray(fractions, np.float32).view(np.uint32)
    return geo


def bouzidi(geo):
    return O.Solver((5, 5, 5), 0.8, walls=O.BOUZIDI, flags=geo)


def test_bouzidi_short_link_uses_second_node():
    s = bouzidi(packed([62], [0.25]))
    s.populations[125 + 61] = 0.01
    s.populations[125 + 60] = 0.004
    assert abs(s.momentum_exchange()[0] - (0.01 + 0.007) / 2) < 5e-16


def test_bouzidi_long_link_uses_leaving_population():
    s = bouzidi(packed([62], [0.75]))
    s.populations[125 + 61] = 0.01
    s.populations[250 + 61] = 0.004
    assert abs(s.momentum_exchange()[0] - (0.01 + 0.008) / 2) < 5e-16
    s.step()
    rho, u = s.macroscopic(61)
    assert abs(rho - 1.008) < 5e-15 and abs(u[0] + 0.008 / 1.008) < 5e-15


@pytest.mark.parametrize("fractions", [[0.5], []])
def test_bouzidi_half_reduces_to_halfway(fractions):
    s = bouzidi(packed([62], fractions))
    s.populations[125 + 61] = 0.01
    assert abs(s.momentum_exchange()[0] - 0.01) < 5e-16


def test_bouzidi_falls_back_when_second_node_is_not_fluid():
    s = bouzidi(packed([62, 60], [0.25, 0.25]))
    s.populations[125 + 61] = 0.01
    assert abs(s.momentum_exchange()[0] - 0.01) < 5e-16


def test_bouzidi_rejects_bad_input():
    flags = np.zeros(125, np.uint32)
    flags[62] = O.SOLID
    with pytest.raises(ValueError, match="packed geometry"):
        O.Solver((5, 5, 5), 0.8, walls=O.BOUZIDI, flags=flags)
    for q in (0, -0.25, 1.5, float("nan")):
        with pytest.raises(ValueError, match="fraction"):
            bouzidi(packed([62], [q]))
    with pytest.raises(ValueError, match="missing"):
        bouzidi(packed([62], [0.25])[:250])


# ---- analytic flows (src/lbm/reference/cases.ts) --------------------------------------
def taylor_green_state(n, x, y, step=0):
    k = 2 * math.pi / n
    nu = (0.8 - 0.5) / 3
    amp = (0.01 * 32 / n) * math.exp(-2 * nu * k * k * step)
    return (1 - 0.75 * amp * amp * (math.cos(2 * k * x) + math.cos(2 * k * y)),
            -amp * math.cos(k * x) * math.sin(k * y), amp * math.sin(k * x) * math.cos(k * y), 0.0)


def run_taylor_green(n, decay=False):
    cells = n * n * 4
    init = np.array([taylor_green_state(n, c % n, (c // n) % n) for c in range(cells)])
    s = O.Solver((n, n, 4), 0.8, initial=init, threads=4)
    steps = int(200 * (n / 32) ** 2)
    xs = np.arange(cells) % n
    ys = (np.arange(cells) // n) % n
    basis = np.array([taylor_green_state(n, x, y)[1:] for x, y in zip(xs, ys)]) / (0.01 * 32 / n)

    def metrics(step):
        u = s.macro_all()[:, 1:]
        exp = np.array([taylor_green_state(n, x, y, step)[1:] for x, y in zip(xs, ys)])
        amp = float((u * basis).sum() / (basis * basis).sum())
        err = math.sqrt(((u - exp) ** 2).sum() / cells) / (0.01 * 32 / n)
        return amp, err

    samples = []
    for step in range(1, steps + 1):
        s.step()
        if decay and step >= 20:
            samples.append((step, metrics(step)[0]))
    amp, err = metrics(steps)
    rate = None
    if decay:
        st = np.array([a for a, _ in samples], float)
        la = np.log([b for _, b in samples])
        rate = -float(((st - st.mean()) * (la - la.mean())).sum() / ((st - st.mean()) ** 2).sum())
    return err, rate, 2 * ((0.8 - 0.5) / 3) * (2 * math.pi / n) ** 2


def taylor_green_error(n):
    return run_taylor_green(n)[0]


@pytest.mark.slow
def test_taylor_green_decay_and_second_order():
    coarse, _, _ = run_taylor_green(16)
    medium, rate, expected = run_taylor_green(32, True)
    fine, _, _ = run_taylor_green(64)
    assert abs(rate / expected - 1) < 0.01
    for order in (math.log2(coarse / medium), math.log2(medium / fine)):
        assert 1.8 < order < 2.2


def run_poiseuille(tau0, op=O.TRT, fluid=16, q=None):
    layout = (2, 2, fluid + 2)
    plane, cells = 4, 4 * (fluid + 2)
    if q is None:
        geo = np.zeros(cells, np.uint32)
        geo[:plane] = O.SOLID
        geo[cells - plane:] = O.SOLID
    else:
        links = 5
        geo = np.zeros(2 * cells + 2 * plane * links, np.uint32)
        geo[:plane] = O.SOLID
        geo[cells - plane:cells] = O.SOLID
        geo[cells:2 * cells] = O.NO_Q
        for c in range(plane):
            geo[cells + plane + c] = c * links
            geo[2 * cells - 2 * plane + c] = (plane + c) * links
        geo[2 * cells:] = np.full(2 * plane * links, q, np.float32).view(np.uint32)
    wall = 0.5 if q is None else float(np.float32(q))
    height = fluid - 1 + 2 * wall
    nu = (tau0 - 0.5) / 3
    accel = 8 * nu * 0.01 / height ** 2
    s = O.Solver(layout, tau0, op, walls=O.BB if q is None else O.BOUZIDI, flags=geo,
                 body_force=(accel, 0, 0))
    prev = None
    for steps in range(1, 100001):
        s.step()
        u = s.macro_all()[plane:cells - plane, 1:]
        if prev is not None and np.abs(u - prev).max() < 1e-12:
            break
        prev = u
    profile = np.array([s.macroscopic((i + 1) * plane)[1][0] for i in range(fluid)])
    d = np.arange(fluid) + wall
    analytic = accel * d * (height - d) / (2 * nu)
    centered = profile - accel / 2
    return dict(steps=steps, max_err=100 * (profile.max() - 0.01) / 0.01,
                profile_err=100 * np.abs(centered - analytic).max() / 0.01,
                force=s.momentum_exchange(), expected=accel * plane * fluid)


@pytest.mark.parametrize("tau0", [0.6, 0.8, 1.5, 3.0])
def test_poiseuille_trt(tau0):
    r = run_poiseuille(tau0)
    assert abs(r["max_err"]) < 1 and r["profile_err"] < 1
    assert abs(r["force"][0] - r["expected"]) < 5e-9
    assert abs(r["force"][1]) < 1e-14 and abs(r["force"][2]) < 1e-14


def test_poiseuille_bgk_wall_location_error_at_tau3():
    assert abs(run_poiseuille(3.0, O.BGK)["max_err"]) > 5


@pytest.mark.parametrize("tau0", [0.6, 0.8, 1.5])
def test_poiseuille_reg(tau0):
    r = run_poiseuille(tau0, O.REG)
    assert abs(r["max_err"]) < 1 and r["profile_err"] < 1
    assert abs(r["force"][0] - r["expected"]) < 5e-9


@pytest.mark.parametrize("tau0", [0.6, 1.5])
def test_poiseuille_bouzidi_q03(tau0):
    r = run_poiseuille(tau0, q=0.3)
    assert r["profile_err"] < 1
    assert abs(r["force"][0] / r["expected"] - 1) < 1e-6


@pytest.mark.parametrize("q", [0.25, 0.75])
def test_bouzidi_second_order(q):
    runs = [(n - 1 + 2 * q, run_poiseuille(0.8, q=q, fluid=n)["profile_err"]) for n in (8, 16, 32)]
    for (hc, ec), (hf, ef) in zip(runs, runs[1:]):
        order = math.log(ec / ef) / math.log(hf / hc)
        assert 1.8 < order < 2.2


def test_bouzidi_half_matches_analytic_profile():
    for n in (8, 16, 32):
        assert run_poiseuille(0.8, q=0.5, fluid=n)["profile_err"] < 1e-4
