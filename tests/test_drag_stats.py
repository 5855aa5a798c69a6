"""Host FP64 drag statistics (restates tests/unit/forces/drag-stats.test.ts), plus the Cd
conversion and the validation window mean against closed forms."""
from __future__ import annotations

import math

import pytest

from pocketwind_amd import drag as D

FLOW_THROUGH = 100


def mulberry32(seed):
    state = seed & 0xFFFFFFFF

    def imul(a, b):
        return (a * b) & 0xFFFFFFFF

    def rnd():
        nonlocal state
        state = (state + 0x6D2B79F5) & 0xFFFFFFFF
        t = state
        t = imul(t ^ (t >> 15), t | 1)
        t ^= (t + imul(t ^ (t >> 7), t | 61)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296
    return rnd


def normal(rnd):
    return math.sqrt(-2 * math.log(1 - rnd())) * math.cos(2 * math.pi * rnd())


def naive_sem(values):
    m = sum(values) / len(values)
    var = sum((v - m) ** 2 for v in values) / (len(values) - 1)
    return 100 * math.sqrt(var / len(values)) / abs(m)


def test_constant_force_is_steady():
    stats = D.create_drag_stats(400)
    for step in range(0, 2401, 10):
        stats.push(step, 12)
    s = stats.summary(None)
    assert s.status == {"kind": "steady", "relStdPct": 0, "driftPct": 0}
    assert s.cd == 12 and s.samples == 40 and s.sem_pct == 0


def test_ten_percent_sine_is_unsteady():
    stats = D.create_drag_stats(FLOW_THROUGH)
    for step in range(0, 1001):
        stats.push(step, 10 + math.sin(2 * math.pi * step / 200))
    assert stats.summary(None).status["kind"] == "unsteady"


def test_settling_then_steady():
    stats = D.create_drag_stats(FLOW_THROUGH)
    for step in range(0, 101, 10):
        stats.push(step, step / 20)
    assert stats.summary(None).status == {"kind": "settling", "etaS": None}
    assert stats.summary(50).status == {"kind": "settling", "etaS": 2}
    for step in range(110, 1201, 10):
        stats.push(step, 10)
    assert stats.summary(None).status["kind"] == "steady"


def test_drift_is_unsteady():
    stats = D.create_drag_stats(800)
    for step in range(0, 1601, 10):
        stats.push(step, 10 * (1 + 0.0000375 * (step - 800)))
    s = stats.summary(None)
    assert s.status["kind"] == "unsteady"
    assert s.status["driftPct"] > 1 and s.status["relStdPct"] < 1


def test_sparse_and_dense_means_agree():
    dense, sparse = D.create_drag_stats(1000), D.create_drag_stats(1000)
    for step in range(0, 3151):
        dense.push(step, 10 + math.sin(2 * math.pi * step / 50))
    for step in range(0, 3151, 50):
        sparse.push(step, 10 + math.sin(2 * math.pi * step / 50))
    assert abs(dense.summary(None).cd - sparse.summary(None).cd) < 1e-12


def test_white_noise_sem():
    rnd = mulberry32(42)
    values = [1 + 0.02 * normal(rnd) for _ in range(512)]
    est = D.sem_pct(values)
    assert est is not None and abs(est / (2 / math.sqrt(512)) - 1) < 0.25


def test_autocorrelated_sem_is_inflated():
    rnd = mulberry32(7)
    innov = 0.02 * math.sqrt(1 - 0.81)
    state = 0.0
    for _ in range(200):
        state = 0.9 * state + innov * normal(rnd)
    values = []
    for _ in range(512):
        state = 0.9 * state + innov * normal(rnd)
        values.append(1 + state)
    assert D.sem_pct(values) >= 2 * naive_sem(values)


def test_withholds_cd_with_too_few_batches():
    stats = D.create_drag_stats(250)
    for step in range(0, 2001):
        stats.push(step, math.sin(2 * math.pi * step / 1000))
    s = stats.summary(None)
    assert s.cd is None and s.sem_pct is None and s.reason == D.TOO_FEW_INDEPENDENT_SAMPLES


def test_batch_length_covers_a_third_of_a_sine_period():
    values = [math.sin(2 * math.pi * step / 2000) for step in range(10241)]
    assert D.batch_length(values) >= 0.3 * 2000


def test_white_noise_integral_time_is_one_interval():
    assert D.tau_int([1, -1, 1, -1, 1, -1, 1, -1], 8) == 8


def test_rejects_invalid_input():
    with pytest.raises(ValueError):
        D.create_drag_stats(0)
    with pytest.raises(ValueError):
        D.create_drag_stats(FLOW_THROUGH, float("nan"))
    with pytest.raises(ValueError):
        D.tau_int([1, 2, 3], 0)
    stats = D.create_drag_stats(FLOW_THROUGH)
    stats.push(10, 1)
    with pytest.raises(ValueError):
        stats.push(10, 2)
    with pytest.raises(ValueError):
        stats.push(11, float("nan"))


def test_cd_conversion_and_window_mean():
    # Cd = F / (1/2 rho u^2 A) with rho = 1 in lattice units.
    assert D.cd_per_lattice_force(193) == 1 / (0.5 * 0.05 * 0.05 * 193)
    assert D.flow_through_steps(128) == 2560
    samples = [(s, float(s), 0.0, 1.0) for s in range(8, 801, 8)]
    fx, fy, fz = D.window_mean_force(samples, 800, 200)
    assert fx == sum(range(608, 801, 8)) / len(range(608, 801, 8)) and fz == 1.0
