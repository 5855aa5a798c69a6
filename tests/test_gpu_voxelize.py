"""The GPU voxelizer is bit-exact against the oracle: identical flags (kinds and pull masks),
q offsets, f32 link fractions (compared as bits), boundary list and counts, on the
reference's test meshes, its presets at tier sizes, the benchmark bluff body and a dense
noisy mesh."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from meshes import box, sphere
from pocketwind_amd import bodies
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu


def assert_same(gpu, ref):
    assert gpu.geometry.size == ref.geometry.size
    diff = np.flatnonzero(gpu.geometry != ref.geometry)
    assert diff.size == 0, f"{diff.size} geometry words differ, first at {diff[:5]}"
    assert np.array_equal(gpu.boundary_cells, ref.boundary_cells)
    assert (gpu.solid_cells, gpu.frontal_cells, gpu.blockage, gpu.watertight, gpu.missed_links) == \
        (ref.solid_cells, ref.frontal_cells, ref.blockage, ref.watertight, ref.missed_links)


CASES = {
    "box": lambda: (box([20.2] * 3, [40.2] * 3), None, (64, 64, 64), False),
    "open-box": lambda: (box([20.2] * 3, [40.2] * 3, True), None, (64, 64, 64), False),
    "box-support-floor": lambda: (box([20.25] * 3, [24.25] * 3), box([30.25] * 3, [34.25] * 3), (64, 64, 64), True),
    "icosphere": lambda: (sphere((31.37, 30.23, 32.41), 20, 5), None, (64, 64, 64), False),
    "periodic-image": lambda: (box([20.2, 60.2, 20.2], [24.2, 63.2, 24.2]), None, (64, 64, 64), False),
    "empty": lambda: (np.zeros(0, np.float32), None, (32, 16, 16), True),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("fractions", [True, False])
def test_reference_meshes_bit_exact(case, fractions):
    body, sup, layout, floor = CASES[case]()
    assert_same(voxelize(body, sup, layout, floor, fractions), O.voxelize(body, sup, layout, floor, fractions))


@pytest.mark.parametrize("layout,length", [((128, 64, 64), 16), ((192, 96, 96), 24), ((256, 128, 128), 32)])
def test_sphere_preset_at_tiers_bit_exact(layout, length):
    body, _ = bodies.place_mesh(bodies.sphere_mesh(), None, length, layout, False)
    assert_same(voxelize(body, None, layout, False), O.voxelize(body, None, layout, False))


def test_ahmed_with_stilts_and_floor_bit_exact():
    layout = (64, 32, 32)
    body, sup = bodies.place_mesh(bodies.ahmed_mesh(), bodies.stilts_mesh(), 24, layout, True)
    assert_same(voxelize(body, sup, layout, True), O.voxelize(body, sup, layout, True))


def test_benchmark_bluff_body_bit_exact():
    layout = (256, 128, 128)
    body, _ = bodies.place_mesh(bodies.bluff_body(), None, 64, layout, False)
    g = voxelize(body, None, layout, False)
    assert_same(g, O.voxelize(body, None, layout, False))
    assert g.watertight and g.missed_links == 0 and g.solid_cells > 0


def test_dense_noisy_mesh_bit_exact():
    layout = (128, 64, 64)
    mesh = bodies.noisy_bluff_body(levels=5, segments=12)
    assert mesh.size // 9 == 118784
    body, _ = bodies.place_mesh(mesh, None, 32, layout, False)
    assert_same(voxelize(body, None, layout, False), O.voxelize(body, None, layout, False))


def test_rejects_invalid_input():
    with pytest.raises(ValueError):
        voxelize(box([20] * 3, [24] * 3), None, (1, 64, 64), False)
    bad = box([20] * 3, [24] * 3)
    bad[4] = np.inf
    with pytest.raises(ValueError):
        voxelize(bad, None, (64, 64, 64), False)


def test_imported_3mf_ellipsoid_placed_and_voxelised_bit_exact():
    """A mesh from the 3MF importer (327680 triangles), placed as a body, free and on the floor,
    and voxelised with link fractions on the device, against the oracle."""
    from archives import ellipsoid_archive
    from pocketwind_amd.mesh_io import parse_model
    model = parse_model(ellipsoid_archive(), "ellipsoid.3mf")
    layout = (128, 64, 64)
    for floor in (False, True):
        body, _ = bodies.place_mesh(model.mesh.positions, None, 40, layout, floor)
        assert_same(voxelize(body, None, layout, floor), O.voxelize(body, None, layout, floor))
