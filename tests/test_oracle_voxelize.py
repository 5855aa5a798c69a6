"""Pins the oracle voxelizer and wall links (restates tests/unit/voxelize/voxelize.test.ts and
tests/unit/wall-links/wall-links.test.ts), checks it against the reference's saved sphere
frontal areas (public/validation/results.json: 193 / 437 / 793 cells at S / M / L), and checks
that the product's preset meshes and placement equal the oracle's bit for bit."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import oracle as O
from meshes import box, sphere
from pocketwind_amd import bodies

LAYOUT = (64, 64, 64)
CELLS = 64 ** 3


def idx(x, y, z):
    return x + 64 * (y + 64 * z)


def test_empty_tunnel_keeps_floor_and_inlet():
    for floor in (False, True):
        g = O.voxelize(np.zeros(0, np.float32), None, LAYOUT, floor, False)
        assert g.solid_cells == 0 and g.frontal_cells == 0 and g.boundary_cells.size == 0
        kinds = g.geometry[:CELLS] & 7
        assert np.any(kinds == O.EQ)
        assert np.any(kinds == O.FLOOR) == floor


def test_sphere_volume_and_area():
    g = O.voxelize(sphere((32, 32, 32), 20), None, LAYOUT, False)
    assert g.watertight
    assert abs(g.solid_cells / (4 / 3 * math.pi * 20 ** 3) - 1) < 0.02
    assert abs(g.frontal_cells / (math.pi * 400) - 1) < 0.03
    assert g.blockage == g.frontal_cells / (64 * 64)


def test_majority_vote_with_a_missing_face():
    closed = O.voxelize(box([20.2] * 3, [40.2] * 3), None, LAYOUT, False)
    opened = O.voxelize(box([20.2] * 3, [40.2] * 3, True), None, LAYOUT, False)
    assert closed.watertight and closed.solid_cells == 8000
    assert not opened.watertight
    assert abs(opened.solid_cells / closed.solid_cells - 1) < 0.05
    assert opened.missed_links > 0


def test_plate_thinner_than_a_cell_stays_fluid():
    g = O.voxelize(box([20.2, 20.2, 20.2], [20.5, 40.2, 40.2]), None, LAYOUT, False)
    assert g.solid_cells == 0 and g.frontal_cells == 0 and g.boundary_cells.size == 0 and g.watertight


def test_body_precedence_frontal_and_blockage():
    g = O.voxelize(box([20.2] * 3, [24.2] * 3), box([22.2, 22.2, 18.2], [26.2, 26.2, 22.2]), LAYOUT, False)
    assert g.solid_cells == 64 and g.frontal_cells == 16
    assert g.blockage == 28 / (64 * 64)
    assert g.geometry[idx(23, 23, 21)] & 7 == O.SOLID
    assert g.geometry[idx(25, 25, 19)] & 7 == O.SUPPORT


def test_equilibrium_faces_and_floor_corners():
    mesh = box([20.2] * 3, [24.2] * 3)
    g = O.voxelize(mesh, None, LAYOUT, True)
    assert g.geometry[idx(0, 4, 0)] == O.FLOOR and g.geometry[idx(63, 4, 0)] == O.FLOOR
    assert g.geometry[idx(0, 4, 4)] == O.EQ and g.geometry[idx(63, 4, 4)] == O.EQ
    assert g.geometry[idx(4, 4, 63)] == O.EQ
    p = O.voxelize(mesh, None, LAYOUT, False)
    assert p.geometry[idx(4, 4, 0)] == O.FLUID and p.geometry[idx(4, 4, 63)] == O.FLUID


def test_shared_diagonals_count_once_regardless_of_winding():
    mesh = box([20] * 3, [24] * 3)
    rev = mesh.reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1).copy()
    f = O.voxelize(mesh, None, LAYOUT, False)
    b = O.voxelize(rev, None, LAYOUT, False)
    assert f.solid_cells == 64 and f.watertight and b.watertight
    assert np.array_equal(f.geometry, b.geometry)


def test_open_support_mesh_is_not_watertight():
    g = O.voxelize(box([20] * 3, [24] * 3), box([30.2] * 3, [34.2] * 3, True), LAYOUT, False)
    assert not g.watertight


def test_rejects_bad_dimensions_and_coordinates():
    mesh = box([20] * 3, [24] * 3)
    with pytest.raises(ValueError):
        O.voxelize(mesh, None, (1, 64, 64), False)
    bad = mesh.copy()
    bad[0] = np.nan
    with pytest.raises(ValueError):
        O.voxelize(mesh, bad, LAYOUT, False)


# ---- wall links ----------------------------------------------------------------------
def analytic_q(origin, d, centre, radius):
    dd = [-v for v in d]
    rel = [o - c for o, c in zip(origin, centre)]
    a = sum(v * v for v in dd)
    b = 2 * sum(r * v for r, v in zip(rel, dd))
    c = sum(r * r for r in rel) - radius * radius
    return (-b - math.sqrt(b * b - 4 * a * c)) / (2 * a)


def test_sphere_links_match_analytic_ray_sphere():
    centre, radius = (31.37, 30.23, 32.41), 20
    mesh = sphere(centre, radius, 5)
    assert mesh.size // 9 == 20480
    g = O.voxelize(mesh, None, LAYOUT, False)
    q = g.fractions
    checked = analytic = 0
    max_radial = max_q = 0.0
    for cell in g.boundary_cells:
        flag = int(g.geometry[cell])
        base = int(g.geometry[CELLS + cell])
        assert base != O.NO_Q
        origin = (cell % 64, (cell // 64) % 64, cell // 4096)
        k = 0
        for i in range(19):
            if not flag & (1 << (3 + i)):
                continue
            d = O.C19[i]
            exp = analytic_q(origin, d, centre, radius)
            act = float(q[base + k])
            hit = [origin[a] - act * d[a] - centre[a] for a in range(3)]
            max_radial = max(max_radial, abs(math.sqrt(sum(h * h for h in hit)) - radius))
            ah = [origin[a] - exp * d[a] - centre[a] for a in range(3)]
            cosine = abs(sum(ah[a] * d[a] for a in range(3))) / (radius * math.sqrt(sum(v * v for v in d)))
            if 0 <= exp <= 1 and cosine >= 0.6:
                max_q = max(max_q, abs(act - exp))
                analytic += 1
            assert 0.001 - 1e-9 <= act <= 1
            k += 1
            checked += 1
    assert checked > 10000 and analytic > 0
    assert max_radial < 0.02 and max_q < 0.02
    assert g.missed_links == 0 and q.size == checked


def test_packing_masks_offsets_and_body_only_boundary():
    g = O.voxelize(box([20.25] * 3, [24.25] * 3), box([30.25] * 3, [34.25] * 3), LAYOUT, True)
    q = g.fractions
    boundary = set(int(c) for c in g.boundary_cells)
    kinds = g.geometry[:CELLS] & 7
    next_base = floor_links = support_links = 0
    for cell in range(CELLS):
        flag = int(g.geometry[cell])
        links = []
        if kinds[cell] == O.FLUID:
            x, y, z = cell % 64, (cell // 64) % 64, cell // 4096
            for i in range(1, 19):
                wx, wy, wz = x - O.C19[i][0], (y - O.C19[i][1]) % 64, z - O.C19[i][2]
                t = O.EQ if (wx < 0 or wx >= 64 or wz < 0 or wz >= 64) else int(kinds[wx + 64 * (wy + 64 * wz)])
                if O.SOLID <= t <= O.FLOOR:
                    links.append((i, t))
        mask = sum(1 << (3 + i) for i, _ in links)
        assert (flag & ~7) == mask
        assert (cell in boundary) == any(t == O.SOLID for _, t in links)
        assert g.geometry[CELLS + cell] == (O.NO_Q if not links else next_base)
        for k, (_, t) in enumerate(links):
            if t == O.FLOOR:
                assert q[next_base + k] == 0.5
                floor_links += 1
            if t == O.SUPPORT:
                support_links += 1
        next_base += len(links)
    assert q.size == next_base and floor_links > 0 and support_links > 0
    assert g.missed_links == 0


def test_halfway_fractions_when_not_measuring():
    mesh = box([20.25] * 3, [24.25] * 3)
    m = O.voxelize(mesh, None, LAYOUT, False)
    h = O.voxelize(mesh, None, LAYOUT, False, link_fractions=False)
    assert np.array_equal(h.geometry[:2 * CELLS], m.geometry[:2 * CELLS])
    assert h.geometry.size == m.geometry.size
    assert np.array_equal(h.boundary_cells, m.boundary_cells)
    assert h.fractions.size > 0 and np.all(h.fractions == 0.5) and h.missed_links == 0


def test_nearest_hit_known_values():
    mesh = box([4, 4, 4], [4.75, 8, 8])
    lay = (16, 16, 16)
    assert abs(O.nearest_link_hit(mesh, lay, (3.5, 6, 6), (1, 0, 0)) - 0.5) < 5e-3
    assert abs(O.nearest_link_hit(mesh, lay, (5, 6, 6), (-1, 0, 0)) - 0.25) < 5e-3
    assert abs(O.nearest_link_hit(mesh, lay, (3, 6, 6), (1, 0, 0)) - 1) < 5e-3
    assert O.nearest_link_hit(mesh, lay, (3, 6, 6), (-1, 0, 0)) is None


@pytest.mark.parametrize("axis", [1, 2])
def test_periodic_image_link(axis):
    lo, hi = [20.2] * 3, [24.2] * 3
    lo[axis], hi[axis] = 60.2, 63.2
    g = O.voxelize(box(lo, hi), None, LAYOUT, False)
    pt = [22, 22, 22]
    pt[axis] = 0
    cell = pt[0] + 64 * (pt[1] + 64 * pt[2])
    off = int(g.geometry[CELLS + cell])
    checked = 0
    for i in range(19):
        if not g.geometry[cell] & (1 << (3 + i)):
            continue
        if O.C19[i][axis] == 1:
            assert abs(g.fractions[off] - 0.8) < 5e-5
            checked += 1
        off += 1
    assert checked == 5 and g.missed_links == 0


def test_clamps_a_wall_close_to_a_fluid_centre():
    g = O.voxelize(box([20.0001, 20.2, 20.2], [24.2] * 3), None, LAYOUT, False)
    cell = 20 + 64 * (22 + 64 * 22)
    assert abs(g.fractions[int(g.geometry[CELLS + cell])] - 0.001) < 5e-7


def test_missed_intersections_use_mid_link():
    flags = np.zeros(CELLS, np.uint32)
    flags[32 + 64 * (32 + 64 * 32)] = O.SOLID
    g = O.build_wall_links(flags, box([10] * 3, [12] * 3), None, LAYOUT, False)
    assert g.boundary_cells.size == 18 and g.missed_links == 18
    assert np.all(g.fractions == 0.5) and g.fractions.size == 18


# ---- presets, placement and the reference's saved sphere voxels -------------------------
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_results.json")))


@pytest.mark.parametrize("name", ["sphere", "ahmed-25", "cube", "ahmed-stilts"])
def test_product_presets_equal_oracle_presets(name):
    product = bodies.stilts_mesh() if name == "ahmed-stilts" else bodies.preset_mesh(name)
    assert np.array_equal(product.view(np.uint32), O.preset_mesh(name).view(np.uint32))


@pytest.mark.parametrize("floor", [False, True])
def test_product_placement_equals_oracle_placement(floor):
    body, sup = bodies.ahmed_mesh(), bodies.stilts_mesh()
    pb, ps = bodies.place_mesh(body, sup, 24, (64, 32, 32), floor)
    ob, os_ = O.place_mesh(body, sup, 24, (64, 32, 32), floor)
    assert np.array_equal(pb.view(np.uint32), ob.view(np.uint32))
    assert np.array_equal(ps.view(np.uint32), os_.view(np.uint32))


@pytest.mark.parametrize("tier", ["S", "M", "L"])
def test_sphere_preset_frontal_cells_match_reference_results(tier):
    row = next(r for r in GOLDEN["laminar"] if r["id"] == f"sphere-re-100-{tier}-trt")
    layout = {"S": (128, 64, 64), "M": (192, 96, 96), "L": (256, 128, 128)}[tier]
    body, _ = O.place_mesh(O.preset_mesh("sphere"), None, row["bodyLengthCells"], layout, False)
    g = O.voxelize(body, None, layout, False, link_fractions=False)
    assert g.frontal_cells == round(row["voxelAreaM2"] / row["cellSizeM"] ** 2)
    assert g.blockage == row["blockage"]
