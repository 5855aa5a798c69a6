"""The particle oracle (oracle/particles.c) against the reference's own known answers.

tests/unit/render/particles.test.ts:13-62 and :98-118: seed region, reseed rule, the lowbias32
seed hash values, the packed seed records, the sub-cell cross-section and the f32 rounding
into a blocked neighbour. tests/browser/particles.spec.ts:19-127: one particle over a 16^3
macro field: the trilinear interpolation of the velocity channels, the one-cell displacement
cap, a stationary frame, the inclusive age limit and the three respawn causes (age, outlet,
next generation). The same spec values are asserted on the device in test_gpu_particles.py.
Also the product's host-side argument checks, which reject before any device call.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from pocketwind_amd import _native as N
from pocketwind_amd.particles import Particles

LAYOUT = (32, 20, 20)
BOUNDS = ((8, 5, 5), (20, 15, 15))


def test_seed_region_enlarges_the_cross_section_and_clips():
    # particles.test.ts:13-19
    assert list(O.seed_region(*BOUNDS, LAYOUT)) == [3.5, 3.5, 16.5, 16.5]
    assert list(O.seed_region((8, 0, 0), (20, 20, 20), LAYOUT)) == [1, 1, 18, 18]
    for bmin, bmax, layout in [(BOUNDS[0], BOUNDS[1], (4, 20, 20)),
                               ((0, 8, 8), (1, 7, 8), LAYOUT),
                               ((0, 30, 30), (1, 40, 40), LAYOUT)]:
        with pytest.raises(ValueError):
            O.seed_region(bmin, bmax, layout)


def test_reseed_rule_keeps_wall_link_fluid_and_the_inclusive_age_limit():
    # particles.test.ts:21-41
    assert not O.particle_needs_reseed((2, 10, 10), 4000, LAYOUT, 1 << 4)
    assert not O.particle_needs_reseed((29.99, 10, 10), 0, LAYOUT, 0)
    assert O.particle_needs_reseed((30, 10, 10), 0, LAYOUT, 0)
    assert O.particle_needs_reseed((2, 10, 10), 4001, LAYOUT, 0)
    for kind in (1, 2, 3, 4):
        assert O.particle_needs_reseed((2, 10, 10), 1, LAYOUT, kind | (1 << 5))
    for position in [(-1, 2, 2), (2, -1, 2), (2, 19, 2), (2, 2, -1), (2, 2, 19), (math.nan, 2, 2)]:
        assert O.particle_needs_reseed(position, 0, LAYOUT, 0)


def test_seed_hash_values():
    # particles.test.ts:43-48
    assert O.particle_hash(0) == 0
    assert O.particle_hash(1) == 1753845952
    assert O.particle_hash(2) == 3507691905
    assert O.particle_hash(0xFFFFFFFF) == 1734902346


def test_seed_records_pack_pos_and_prev_at_fluid_inlet_nodes():
    # particles.test.ts:50-62
    flags = np.ones(32 * 20 * 20, np.uint32)
    flags[2 + 32 * (4 + 20 * 4)] = 1 << 4
    flags[2 + 32 * (5 + 20 * 4)] = 0
    data, _, fallback = O.particle_seeds(flags, LAYOUT, *BOUNDS, 4, 0)
    assert list(fallback) == [2, 4, 4]
    assert data.nbytes == 128
    assert list(data) == [2, 4, 4, 0, 2, 4, 4, 0, 2, 4, 4, 0, 2, 4, 4, 0,
                          2, 5, 4, 0, 2, 5, 4, 0, 2, 5, 4, 0, 2, 5, 4, 0]
    again, _, _ = O.particle_seeds(flags, LAYOUT, *BOUNDS, 4, 0)
    assert np.array_equal(again, data)
    other, _, _ = O.particle_seeds(flags, LAYOUT, *BOUNDS, 4, 1)
    assert not np.array_equal(other, data)


def test_seeds_reject_a_region_without_fluid_inlet_nodes():
    # particles.test.ts:64-69 and :112-118 (f32 rounding of 5.49999999 lands on y = 6)
    with pytest.raises(ValueError, match="no fluid inlet"):
        O.particle_seeds(np.ones(12800, np.uint32), LAYOUT, *BOUNDS, 1, 1)
    flags = np.ones(12800, np.uint32)
    flags[2 + 32 * (5 + 20 * 6)] = 0
    with pytest.raises(ValueError, match="no fluid inlet"):
        O.particle_seeds(flags, LAYOUT, (8, 5.49999999, 6.1), (9, 5.5, 6.2), 1, 0)


def test_sub_cell_cross_section_seeds_without_a_node_inside():
    # particles.test.ts:98-109
    data, _, _ = O.particle_seeds(np.zeros(12800, np.uint32), LAYOUT, (8, 5.1, 6.1), (9, 5.2, 6.2), 1, 0)
    assert data[0] == 2
    assert data[1] == pytest.approx(5.085, abs=5e-6)
    assert data[2] == pytest.approx(6.085, abs=5e-6)
    assert list(data[4:7]) == list(data[0:3])


# tests/browser/particles.spec.ts:19-127: one particle, seed 7, 16^3 all-fluid grid, body
# bounds (6, 4, 4)-(10, 10, 10), macro (2, 0.01 x, 0.02 y, 0.03 z) and then (20, 3, 4, 0).
SPEC_LAYOUT = (16, 16, 16)
SPEC_BOUNDS = ((6, 4, 4), (10, 10, 10))
SPEC_SEED = 7


def spec_macros():
    cell = np.arange(4096)
    graded = np.stack([np.full(4096, 2.0), (cell % 16) * 0.01, ((cell // 16) % 16) * 0.02,
                       (cell // 256) * 0.03], axis=1).astype(np.float32)
    uniform = np.tile(np.array([20, 3, 4, 0], np.float32), (4096, 1))
    return graded, uniform


def spec_record(age=10.0, x=4.25, generation=0.0):
    return np.array([x, 5.5, 6.75, age, x, 5.5, 6.75, generation], np.float32)


def check_spec_results(r):
    """The spec's expectations on the seven records it reads back (particles.spec.ts:100-125)."""
    for got, expected in ((r["interpolated"], [4.335, 5.72, 7.155, 11, 4.25, 5.5, 6.75, 0]),
                          (r["capped"], [4.85, 6.3, 6.75, 11, 4.25, 5.5, 6.75, 0])):
        assert np.allclose(got, expected, rtol=0, atol=5e-6), (got, expected)
    assert list(r["stationary"]) == [4.25, 5.5, 6.75, 11, 4.25, 5.5, 6.75, 0]
    assert list(r["age_limit"]) == [4.25, 5.5, 6.75, 4000, 4.25, 5.5, 6.75, 0]
    for respawn in (r["expired"], r["outlet"], r["next_generation"]):
        assert respawn[0] == 2 and respawn[3] == 0
        assert list(respawn[4:7]) == list(respawn[0:3])
        assert np.all((respawn[1:3] >= 3.1) & (respawn[1:3] <= 10.9))
    assert list(r["expired"]) == list(r["outlet"])
    assert r["expired"][7] == 1 and r["next_generation"][7] == 2
    assert list(r["next_generation"][1:3]) != list(r["expired"][1:3])


def test_advection_known_values_of_the_browser_spec():
    graded, uniform = spec_macros()
    flags = np.zeros(4096, np.uint32)
    data, region, fallback = O.particle_seeds(flags, SPEC_LAYOUT, *SPEC_BOUNDS, 1, SPEC_SEED)
    assert list(region) == pytest.approx([3.1, 3.1, 10.9, 10.9])

    def record(macro, steps, **kw):
        return O.particles_update(spec_record(**kw), macro, flags, SPEC_LAYOUT, SPEC_SEED, steps,
                                  region.astype(np.float32), fallback[1:])

    check_spec_results({
        "interpolated": record(graded, 2),
        "capped": record(uniform, 9),
        "stationary": record(uniform, 0),
        "age_limit": record(uniform, 0, age=3999),
        "expired": record(uniform, 0, age=4000),
        "outlet": record(uniform, 0, x=14),
        "next_generation": record(uniform, 0, age=4000, generation=1),
    })


def test_planted_interpolation_mistake_fails_the_spec_values():
    """The spec's interpolated values tell apart a transposed velocity channel: the oracle run
    with the macro's y and z velocity columns swapped misses them."""
    graded, _ = spec_macros()
    swapped = graded[:, [0, 1, 3, 2]].copy()
    flags = np.zeros(4096, np.uint32)
    _, region, fallback = O.particle_seeds(flags, SPEC_LAYOUT, *SPEC_BOUNDS, 1, SPEC_SEED)
    got = O.particles_update(spec_record(), swapped, flags, SPEC_LAYOUT, SPEC_SEED, 2,
                             region.astype(np.float32), fallback[1:])
    assert not np.allclose(got[:3], [4.335, 5.72, 7.155], atol=5e-6)


# ---- the product's host-side checks (no device call is made before they reject) ----

def test_product_rejects_invalid_counts_and_seeds_before_device_work():
    # particles.test.ts:71-76 (count 0, 1.5, 65537; seed -1, 1.5, 2^32)
    geometry = np.zeros(12800, np.uint32)
    for count in (0, 1.5, 65537):
        with pytest.raises(ValueError):
            Particles.create_on_field(LAYOUT, geometry, 1, BOUNDS, count, 1)
    for seed in (-1, 1.5, 2 ** 32):
        with pytest.raises(ValueError):
            Particles.create_on_field(LAYOUT, geometry, 1, BOUNDS, 1, seed)


def test_product_rejects_short_geometry_and_empty_cross_sections_before_device_work():
    # particles.test.ts:67-69 (geometry shorter than the grid) and :16-18 (bad regions)
    with pytest.raises(ValueError, match="geometry buffer"):
        Particles.create_on_field(LAYOUT, np.zeros(0, np.uint32), 1, BOUNDS, 1, 1)
    for bmin, bmax, layout in [(BOUNDS[0], BOUNDS[1], (4, 20, 20)),
                               ((0, 8, 8), (1, 7, 8), LAYOUT),
                               ((0, 30, 30), (1, 40, 40), LAYOUT)]:
        geometry = np.zeros(layout[0] * layout[1] * layout[2], np.uint32)
        with pytest.raises(ValueError, match="positive cross-section"):
            Particles.create_on_field(layout, geometry, 1, (bmin, bmax), 1, 1)


def test_product_rejects_non_finite_or_negative_steps():
    # particles.test.ts:88-91: -1, NaN, Infinity, Number.MAX_VALUE (infinite in f32)
    dummy = Particles(None, None, 1)
    for steps in (-1, math.nan, math.inf, 1.7976931348623157e308):
        with pytest.raises(ValueError):
            dummy.update(steps)
    lib = N.lib()
    assert lib.pw_particles_update(None, C.c_float(1.0)) == N.PW_EINVAL
