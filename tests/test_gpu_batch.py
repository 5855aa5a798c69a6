"""Batched tunnels: every tunnel advanced inside one launch per step (and its replayed
hipGraph) is bit-identical to the same tunnel advanced alone by pw_solver with the same step
variant, in populations, macros and FP64 force samples."""
from __future__ import annotations

import math

import numpy as np
import pytest

from pocketwind_amd import bodies
from pocketwind_amd.batch import TunnelBatch
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu


def tunnel(layout, kind, seed):
    rng = np.random.default_rng(seed)
    mesh = {"sphere": bodies.sphere_mesh, "box": bodies.cube_mesh, "bluff": bodies.bluff_body}[kind]()
    body, _ = bodies.place_mesh(mesh, None, layout[0] * rng.uniform(0.12, 0.2), layout, False)
    grid = voxelize(body, None, layout, False, link_fractions=False)
    re = rng.uniform(20, 200)
    u = rng.uniform(0.03, 0.06)
    d = layout[0] * 0.15
    return grid, 3 * (u * d / re) + 0.5, u


CASES = [((32, 32, 32), "sphere", 1), ((64, 32, 32), "box", 2), ((48, 24, 24), "bluff", 3),
         ((96, 48, 48), "sphere", 4)]


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_batch_equals_single_tunnels(precision):
    tunnels = [tunnel(*c) for c in CASES]
    shared = SolverParams(0.6, 0.05, False, precision=precision)
    batch = TunnelBatch.create([t[0] for t in tunnels], [t[1] for t in tunnels], [t[2] for t in tunnels], shared)
    try:
        samples = batch.step_sampled(203, 8)
        batch.step(1, True)
        for k, (grid, tau, u) in enumerate(tunnels):
            s = GpuSolver.create(grid.layout, grid, SolverParams(tau, u, False, precision=precision))
            s.set_kernel("cell-nt")  # the batch's per-cell step with non-temporal streams
            single = s.step_sampled(203, 8)
            s.step(1, True)
            assert np.array_equal(single, samples[k]), f"tunnel {k} force samples differ"
            assert np.array_equal(s.read_populations().view(np.uint16), batch.read_populations(k).view(np.uint16))
            assert np.array_equal(s.read_macro(), batch.read_macro(k))
            s.destroy()
        assert batch.step_count == 204
        assert batch.cells == sum(g.cells for g, _, _ in tunnels)
    finally:
        batch.destroy()


def test_batch_rejects_invalid_tunnel():
    grid, tau, u = tunnel((32, 32, 32), "sphere", 1)
    with pytest.raises(ValueError, match="Tunnel 1"):
        TunnelBatch.create([grid, grid], [tau, 0.5], [u, u], SolverParams(0.6, 0.05, False))


def test_batch_with_learned_closure_and_particles_equals_single_tunnels():
    """f16 REG + LES tunnels with the learned wall closure (every 8 steps) and 4096 particles
    on each: the batch (row-wavefront and per-cell tunnels in one launch) matches single
    solvers bit for bit, and particles bound to a batched tunnel match particles bound to the
    single solver."""
    import os
    from pocketwind_amd import closure
    from pocketwind_amd.particles import Particles
    net, norm, _ = closure.load(os.path.join(os.path.dirname(closure.__file__), "closure_c2.npz"))
    cases = [((128, 64, 64), "bluff", 5), ((96, 48, 48), "sphere", 6)]
    grids = []
    for layout, kind, seed in cases:
        mesh = {"sphere": bodies.sphere_mesh, "bluff": bodies.bluff_body}[kind]()
        body, _ = bodies.place_mesh(mesh, None, layout[0] / 4, layout, False)
        grids.append((voxelize(body, None, layout, False, link_fractions=False), body))
    shared = SolverParams(0.53, 0.05, True, operator="reg", precision="f16", wall_model="learned")
    batch = TunnelBatch.create([g for g, _ in grids], [0.53] * 2, [0.05] * 2, shared)
    bparts = []
    try:
        batch.set_closure(net, norm)
        batch.set_closure_interval(8)
        for k, (g, body) in enumerate(grids):
            lo, hi = body.reshape(-1, 3).min(0), body.reshape(-1, 3).max(0)
            bparts.append(Particles.create_in_batch(batch, k, (lo, hi), 4096, 7, max_count=4096))
        samples = batch.step_sampled(64, 8)
        batch.step(1, True)
        for p in bparts:
            p.update(1.0)
        for k, (g, body) in enumerate(grids):
            s = GpuSolver.create(g.layout, g, shared)
            s.set_kernel("cell-nt")
            s.set_closure(net, norm)
            s.set_closure_interval(8)
            lo, hi = body.reshape(-1, 3).min(0), body.reshape(-1, 3).max(0)
            sp = Particles.create(s, (lo, hi), 4096, 7, max_count=4096)
            single = s.step_sampled(64, 8)
            s.step(1, True)
            sp.update(1.0)
            used = int(np.sum(s.closure_output()[:, 3] > 0))
            print(f"tunnel {k}: {used} of {s.wall_cells().size} wall cells on the network")
            assert used > 0
            assert np.array_equal(single, samples[k]), f"tunnel {k} force samples differ"
            assert np.array_equal(s.read_populations().view(np.uint16), batch.read_populations(k).view(np.uint16))
            assert np.array_equal(s.read_macro(), batch.read_macro(k))
            assert np.array_equal(sp.read(), bparts[k].read())
            sp.destroy()
            s.destroy()
    finally:
        for p in bparts:
            p.destroy()
        batch.destroy()


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_persistent_step_form_equals_one_block_per_tile(monkeypatch, precision):
    """The batched step's persistent form (resident blocks walking the tiles) gives the
    one-block-per-tile form's populations and force samples bit for bit."""
    # enough tiles (5 216) that the resident blocks (7 per CU for REG + LES) each walk several
    tunnels = [tunnel(*c) for c in [((128, 64, 64), "sphere", 1), ((96, 48, 48), "box", 2),
                                     ((128, 64, 64), "bluff", 3), ((64, 32, 32), "sphere", 4)]]
    shared = SolverParams(0.6, 0.05, True, operator="reg", precision=precision)
    out = {}
    for form in ("cell-nt-stores/16/persistent", "cell-nt-stores/16"):
        monkeypatch.setenv("PW_BATCH_KERNEL", form)
        batch = TunnelBatch.create([t[0] for t in tunnels], [t[1] for t in tunnels], [t[2] for t in tunnels], shared)
        try:
            variant = batch.variant
            samples = batch.step_sampled(101, 8)
            batch.step(1, True)
            out[form] = (variant, samples, [batch.read_populations(k).copy() for k in range(len(tunnels))])
        finally:
            batch.destroy()
    (vp, sp, pp), (vn, sn, pn) = out["cell-nt-stores/16/persistent"], out["cell-nt-stores/16"]
    assert "/persistent-" in vp and "persistent" not in vn
    assert np.array_equal(sp, sn)
    for a, b in zip(pp, pn):
        assert np.array_equal(a.view(np.uint16), b.view(np.uint16))
