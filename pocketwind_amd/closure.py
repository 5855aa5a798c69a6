"""The learned wall closure: its network (bf16 MFMA kernels in csrc/mlp.hip) and helpers.

The network maps 12 wall-cell features to the 3 components of the wall shear stress. It has no
counterpart in the reference, whose wall treatment is the wall law of
src/lbm/shaders/stream_collide.wgsl:1-80 (the closure's fallback).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

INPUTS, HIDDEN, OUTPUTS = 12, 64, 3
N_PARAMS = HIDDEN * INPUTS + HIDDEN + HIDDEN * HIDDEN + HIDDEN + OUTPUTS * HIDDEN + OUTPUTS


def _f32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights (limit sqrt(6 / fan_in)) and zero biases, as float32."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in in ((HIDDEN, INPUTS), (HIDDEN, HIDDEN), (OUTPUTS, HIDDEN)):
        lim = np.sqrt(6.0 / fan_in)
        parts += [rng.uniform(-lim, lim, fan_out * fan_in), np.zeros(fan_out)]
    return np.concatenate(parts).astype(np.float32)


class Mlp:
    """Handle of a network on the device (pw_mlp_*)."""

    def __init__(self, params=None):
        h = C.c_void_p()
        N.check(N.lib().pw_mlp_create(C.byref(h)))
        self._h = h
        if params is not None:
            self.params = params

    def destroy(self):
        if self._h:
            N.lib().pw_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, np.float32)
        N.check(N.lib().pw_mlp_get_params(self._h, out.ctypes.data, out.size))
        return out

    @params.setter
    def params(self, value):
        v = _f32(value, (N_PARAMS,))
        N.check(N.lib().pw_mlp_set_params(self._h, v.ctypes.data, v.size))

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        if x.ndim != 2 or x.shape[1] != INPUTS:
            raise ValueError("inputs must be rows x 12")
        y = np.empty((x.shape[0], OUTPUTS), np.float32)
        N.check(N.lib().pw_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def gradients(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.empty(N_PARAMS, np.float32)
        loss = C.c_float()
        N.check(N.lib().pw_mlp_gradients(self._h, x.ctypes.data, t.ctypes.data, x.shape[0], g.ctypes.data,
                                         C.byref(loss)))
        return float(loss.value), g

    def set_data(self, x, t):
        x, t = _f32(x), _f32(t)
        if x.shape[0] != t.shape[0] or x.shape[1] != INPUTS or t.shape[1] != OUTPUTS:
            raise ValueError("data must be rows x 12 inputs and rows x 3 targets")
        self._data = (x, t)
        N.check(N.lib().pw_mlp_set_data(self._h, x.ctypes.data, t.ctypes.data, x.shape[0]))

    def train(self, steps: int, batch: int, lr: float, seed: int = 0) -> np.ndarray:
        losses = np.empty(steps, np.float32)
        N.check(N.lib().pw_mlp_train(self._h, steps, batch, C.c_float(lr), seed, losses.ctypes.data))
        return losses


# --- Training data from this solver's own coarse/fine runs ---------------------------------

N_NORM = 27  # 12 means, 12 scales, output scale, limit, velocity unit


def random_body(seed: int) -> np.ndarray:
    """Seeded rounded-box / slanted body (bodies.bluff_body with drawn proportions)."""
    from . import bodies
    rng = np.random.default_rng(seed)
    return bodies.bluff_body(width=rng.uniform(0.28, 0.45), height=rng.uniform(0.22, 0.36),
                             radius=rng.uniform(0.03, 0.14), slant_length=rng.uniform(0.1, 0.3),
                             slant_deg=rng.uniform(0.0, 30.0))


def _tunnel(mesh, layout, length, nu, u_in, wall_model="law", precision="f32"):
    from . import bodies
    from .gpu_solver import GpuSolver, SolverParams
    from .voxelize import voxelize
    body, _ = bodies.place_mesh(mesh, None, length, layout, False)
    grid = voxelize(body, None, layout, False, link_fractions=True)
    # Halfway bounce-back: the wall function (stream_collide.wgsl:18-80) assumes it; with
    # Bouzidi walls it drives spurious near-wall velocities (the reference restatement too).
    # The voxelizer's link fractions still reach the closure as its q feature.
    params = SolverParams(3 * nu + 0.5, u_in, True, operator="reg", walls="bb", precision=precision,
                          wall_model=wall_model)
    solver = GpuSolver.create(layout, grid, params)
    if not isinstance(solver, GpuSolver):
        raise MemoryError(solver.message)
    return solver, grid


def _averaged(solver, warmup, window, snapshots, read):
    solver.step(warmup, False)
    every = max(1, window // snapshots)
    acc = None
    for _ in range(snapshots):
        solver.step(every, False)
        v = read()
        acc = v.astype(np.float64) if acc is None else acc + v
    return acc / snapshots


def _coords(cells, layout):
    nx, ny, _ = layout
    cells = np.asarray(cells, np.int64)
    return np.stack([cells % nx, (cells // nx) % ny, cells // (nx * ny)], 1)


def restrict_stress(coarse_cells, coarse_layout, normals, fine_cells, fine_layout, fine_forces, fine_normals,
                    fine_exposed, radius: int = 2, align: float = 0.9):
    """Wall shear stress the fine run puts on the patch of wall of each coarse wall cell: over
    the fine wall cells within `radius` fine cells of the coarse cell's image (2X) whose wall
    normal is within acos(align) of the coarse one, the sum of their forces with each one's
    normal component removed (then the coarse normal's), over the sum of their exposed weights
    (6 sum w = faces). Returns (stress, support), support the fine face count."""
    fc = _coords(fine_cells, fine_layout)
    index = {tuple(c): k for k, c in enumerate(fc)}
    cc = 2 * _coords(coarse_cells, coarse_layout)
    ff = np.asarray(fine_forces, np.float64)
    fn = np.asarray(fine_normals, np.float64)
    tangential = ff - np.sum(ff * fn, 1, keepdims=True) * fn
    stress = np.zeros((len(coarse_cells), 3))
    support = np.zeros(len(coarse_cells))
    span = range(-radius, radius + 1)
    nfx, nfy, nfz = fine_layout
    for k, (x, y, z) in enumerate(cc):
        n = normals[k]
        f = np.zeros(3)
        area = 0.0
        for dz in span:
            for dy in span:
                for dx in span:
                    j = index.get(((x + dx) % nfx, (y + dy) % nfy, (z + dz) % nfz))
                    if j is not None and np.dot(fn[j], n) >= align:
                        f += tangential[j]
                        area += fine_exposed[j]
        if area > 0:
            f = f - np.dot(f, n) * n
            stress[k] = f / area
            support[k] = area
    return stress, support


def coarse_fine_pair(seed: int, coarse_layout=(48, 24, 24), length: int = 16, re_range=(100.0, 400.0),
                     u_in: float = 0.05, flow_throughs: float = 4.0, snapshots: int = 30):
    """One training body: the closure features of its coarse run and the wall shear stress
    its 2x finer run puts on the same wall, both averaged over the last flow-through.
    Returns dict(features (n x 12), law (n x 3), target (n x 3), coarse_stress (n x 3))."""
    rng = np.random.default_rng(seed + 7919)
    re = rng.uniform(*re_range)
    nu = u_in * length / re
    mesh = random_body(seed)
    fine_layout = tuple(2 * v for v in coarse_layout)
    flow = int(np.ceil(coarse_layout[0] / u_in))
    warm = int((flow_throughs - 1) * flow)
    solver, _ = _tunnel(mesh, coarse_layout, length, nu, u_in)
    try:
        rec = _averaged(solver, warm, flow, snapshots, lambda: solver.closure_features(u_in))
        cells = solver.wall_cells()
        coarse_forces = _averaged(solver, 0, flow // 4, max(1, snapshots // 4), solver.wall_forces)
    finally:
        solver.destroy()
    solver, _ = _tunnel(mesh, fine_layout, 2 * length, 2 * nu, u_in)
    try:
        fine_forces = _averaged(solver, 2 * warm, 2 * flow, snapshots, solver.wall_forces)
        fine_cells = solver.wall_cells()
        fine_rec = solver.closure_features(u_in)
    finally:
        solver.destroy()
    normals = rec[:, 8:11]
    normals = normals / np.maximum(np.linalg.norm(normals, axis=1, keepdims=True), 1e-12)
    target, support = restrict_stress(cells, coarse_layout, normals, fine_cells, fine_layout, fine_forces,
                                      fine_rec[:, 8:11], fine_rec[:, 11].astype(np.float64))
    area = rec[:, 11].astype(np.float64)
    cf = coarse_forces - np.sum(coarse_forces * normals, 1, keepdims=True) * normals
    keep = support > 0
    return dict(features=rec[keep, :12], law=rec[keep, 12:15], target=target[keep],
                coarse_stress=(cf / area[:, None])[keep], re=re)


def fit_normalization(features, targets, limit: float = 6.0, u_ref: float = 0.05) -> np.ndarray:
    mean = features.mean(0)
    std = features.std(0)
    scale = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0)
    out = float(np.sqrt(np.mean(targets ** 2))) or 1.0
    return np.concatenate([mean, scale, [out, limit, u_ref]]).astype(np.float32)


def normalise(features, norm):
    return (features - norm[:12]) * norm[12:24]


def train_closure(features, targets, steps: int = 4000, batch: int = 512, lr: float = 2e-3, seed: int = 0,
                  norm=None):
    """Fits the network on the device. Returns (net, norm, loss history)."""
    norm = fit_normalization(features, targets) if norm is None else norm
    net = Mlp(init_params(seed))
    net.set_data(normalise(features, norm), targets / norm[24])
    history = net.train(steps, batch, lr, seed=seed)
    return net, norm, history


def predict(net, norm, features) -> np.ndarray:
    return net.forward(normalise(features, norm)) * norm[24]


# --- Trained closure files ------------------------------------------------------------------

def save(path, net, norm, meta=None):
    import json
    np.savez(path, params=net.params, norm=np.asarray(norm, np.float32),
             meta=np.array(json.dumps(meta or {})))


def load(path):
    """(Mlp, norm, meta) from a file written by save()."""
    import json
    with np.load(path, allow_pickle=False) as d:
        net = Mlp(d["params"])
        return net, d["norm"].astype(np.float32), json.loads(str(d["meta"]))


def evaluate_cd(reynolds: float = 2000.0, layout=(256, 128, 128), length: int = 64, closure_file=None,
                u_in: float = 0.05, flow_throughs: int = 6, average: int = 2, log=print) -> dict:
    """Drag of the slanted bluff body (``bodies.bluff_body``, nose at 0.25 nx, centred) at
    Reynolds number `reynolds` on its length, REG + Smagorinsky LES, halfway bounce-back:
    f16 storage with the learned closure (every 8 steps) and f32 with the wall law on `layout`,
    and f32 with the wall law on the twice finer grid (the reference result the closure is
    judged against). Cd = mean fx over the last `average` of `flow_throughs` flow-throughs
    (runner.ts:234-244) on the voxel frontal area."""
    import os
    from . import bodies, drag
    from .gpu_solver import GpuSolver, SolverParams
    from .voxelize import voxelize
    if closure_file is None:
        closure_file = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closure_c2.npz")
    net, norm, _ = load(closure_file)

    def cd(grid_layout, grid_length, precision, learned):
        body, _ = bodies.place_mesh(bodies.bluff_body(), None, grid_length, grid_layout, False)
        grid = voxelize(body, None, grid_layout, False, link_fractions=True)
        nu = u_in * grid_length / reynolds
        s = GpuSolver.create(grid_layout, grid, SolverParams(3 * nu + 0.5, u_in, True, operator="reg",
                                                             precision=precision,
                                                             wall_model="learned" if learned else "law"))
        try:
            if learned:
                s.set_closure(net, norm)
                s.set_closure_interval(8)
            flow = drag.flow_through_steps(grid.nx, u_in)
            samples = s.step_sampled(flow_throughs * flow, 8)
            fx, _, _ = drag.window_mean_force(samples, flow_throughs * flow, average * flow)
            used = int(np.sum(s.closure_output()[:, 3] > 0)) if learned else 0
            walls = int(s.wall_cells().size)
        finally:
            s.destroy()
        value = fx * drag.cd_per_lattice_force(grid.frontal_cells, u_in)
        log(f"  {grid_layout} {precision} {'learned' if learned else 'wall law'}: Cd {value:.6f}"
            + (f" ({used} of {walls} wall cells on the network)" if learned else ""))
        return value, used, walls

    fine = tuple(2 * n for n in layout)
    learned, used, walls = cd(layout, length, "f16", True)
    law, _, _ = cd(layout, length, "f32", False)
    ref, _, _ = cd(fine, 2 * length, "f32", False)
    net.destroy()
    return {"reynolds": reynolds, "layout": list(layout), "length_cells": length,
            "cd_f16_learned": learned, "cd_f32_law": law, "cd_f32_law_2x_grid": ref,
            "learned_vs_2x_grid_pct": 100 * (learned / ref - 1), "law_vs_2x_grid_pct": 100 * (law / ref - 1),
            "wall_cells": walls, "wall_cells_on_network": used}


def drag_split(forces, normals, frontal_cells: float, u_in: float) -> dict:
    """Cd of per-wall-cell forces (n x 3, lattice units) split into the part along each cell's
    unit wall normal (pressure) and the rest (friction)."""
    f = np.asarray(forces, np.float64)
    n = np.asarray(normals, np.float64)
    n = n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-12)
    normal = np.sum(f * n, 1, keepdims=True) * n
    scale = 1.0 / (0.5 * u_in * u_in * frontal_cells)
    return {"cd": float(f[:, 0].sum() * scale), "pressure": float(normal[:, 0].sum() * scale),
            "friction": float((f - normal)[:, 0].sum() * scale)}


def _floor_run(grid_layout, grid_length, reynolds, u_in, flow_throughs, average, snapshots, prescribed=None,
               source=False, log=print):
    """The slanted bluff body on `grid_layout` with the wall law, or with `prescribed` = (vectors,
    valid) as the closure's output (shear stresses, or momentum sources with `source`): the
    window-mean Cd of the body force (for sources, the sampled momentum-exchange force minus the
    sum of the sources), the time-averaged per-wall-cell momentum-exchange forces over the
    window, and the wall cells' normals and exposed weights."""
    from . import bodies, drag
    from .gpu_solver import GpuSolver, SolverParams
    from .voxelize import voxelize
    body, _ = bodies.place_mesh(bodies.bluff_body(), None, grid_length, grid_layout, False)
    grid = voxelize(body, None, grid_layout, False, link_fractions=True)
    nu = u_in * grid_length / reynolds
    s = GpuSolver.create(grid_layout, grid, SolverParams(3 * nu + 0.5, u_in, True, operator="reg",
                                                         wall_model="law" if prescribed is None else "learned"))
    try:
        sources = np.zeros(3)
        if prescribed is not None:
            s.prescribe_wall_stress(*prescribed, source=source)
            if source:
                vec = np.asarray(prescribed[0], np.float32).astype(np.float64)
                sources = vec[np.asarray(prescribed[1], bool)].sum(0)
        flow = drag.flow_through_steps(grid.nx, u_in)
        total, window = flow_throughs * flow, average * flow
        samples = [s.step_sampled(total - window, 8)]
        chunk = max(8, (window // snapshots) // 8 * 8)
        forces, done = None, total - window
        while done < total:
            k = min(chunk, total - done)
            samples.append(s.step_sampled(k, 8))
            done += k
            f = s.wall_forces()
            forces = f if forces is None else forces + f
        forces /= -(-window // chunk)
        fx, _, _ = drag.window_mean_force(np.concatenate(samples), total, window)
        scale = drag.cd_per_lattice_force(grid.frontal_cells, u_in)
        rec = s.closure_features(u_in)
        out = dict(cells=s.wall_cells(), normals=rec[:, 8:11].astype(np.float64),
                   exposed=rec[:, 11].astype(np.float64), forces=forces, frontal=grid.frontal_cells,
                   cd=(fx - sources[0]) * scale, cd_exchange=fx * scale, cd_sources=-sources[0] * scale,
                   split=drag_split(forces, rec[:, 8:11], grid.frontal_cells, u_in))
    finally:
        s.destroy()
    what = "wall law" if prescribed is None else ("prescribed momentum sources" if source else "prescribed shear")
    log(f"  {grid_layout} {what}: Cd {out['cd']:.6f} (exchange {out['cd_exchange']:.6f}, sources "
        f"{out['cd_sources']:+.6f}; time-averaged wall forces: {out['split']})")
    return out


def shear_floor(reynolds: float = 320.0, layout=(256, 128, 128), length: int = 64, u_in: float = 0.05,
                flow_throughs: int = 6, average: int = 2, snapshots: int = 40, log=print) -> dict:
    """What a wall-shear closure can do for Cd at best: the slanted bluff body of evaluate_cd
    on `layout` with the wall law; on the twice finer grid with the wall law; and on `layout`
    with the fine grid's time-averaged wall shear restricted to each coarse wall cell
    (restrict_stress, the closure's training target) prescribed as the closure output, i.e. a
    perfect network. Each run's window-mean Cd and its time-averaged wall forces split into
    pressure (wall-normal) and friction (tangential) parts."""
    run = lambda grid_layout, grid_length, prescribed=None: _floor_run(
        grid_layout, grid_length, reynolds, u_in, flow_throughs, average, snapshots, prescribed, False, log)
    fine_layout = tuple(2 * v for v in layout)
    coarse = run(layout, length)
    fine = run(fine_layout, 2 * length)
    normals = coarse["normals"] / np.maximum(np.linalg.norm(coarse["normals"], axis=1, keepdims=True), 1e-12)
    target, support = restrict_stress(coarse["cells"], layout, normals, fine["cells"], fine_layout, fine["forces"],
                                      fine["normals"], fine["exposed"])
    perfect = run(layout, length, (target, support > 0))
    pct = lambda a: 100 * (a / fine["cd"] - 1)
    return {"reynolds": reynolds, "layout": list(layout), "length_cells": length,
            "cd_law": coarse["cd"], "cd_prescribed_fine_shear": perfect["cd"], "cd_law_2x_grid": fine["cd"],
            "law_vs_2x_grid_pct": pct(coarse["cd"]), "prescribed_vs_2x_grid_pct": pct(perfect["cd"]),
            "split_law": coarse["split"], "split_prescribed": perfect["split"], "split_2x_grid": fine["split"],
            "wall_cells": int(len(coarse["cells"])), "wall_cells_with_target": int(np.sum(support > 0))}


def restrict_force(coarse_cells, coarse_layout, coarse_normals, fine_cells, fine_layout, fine_forces, fine_normals,
                   scale: float, neighbours: int = 8, align: float = 0.0) -> np.ndarray:
    """Total wall force the fine run puts on the patch of wall of each coarse wall cell: every
    fine wall cell goes to the nearest coarse wall cell (by the fine cell centre's image at half
    scale) among its `neighbours` nearest whose normal is within acos(align) of its own (the
    nearest of all when none is), so the restricted forces sum to the fine total; times `scale`
    (the coarse over the fine frontal area, the factor between the two grids' forces at equal
    Cd). Pressure and friction are both kept."""
    from scipy.spatial import cKDTree
    cc = _coords(coarse_cells, coarse_layout).astype(np.float64)
    fc = (_coords(fine_cells, fine_layout).astype(np.float64) + 0.5) / 2.0 - 0.5
    tree = cKDTree(cc)
    k = min(neighbours, len(cc))
    _, idx = tree.query(fc, k=k)
    idx = idx.reshape(len(fc), k)
    cn = np.asarray(coarse_normals, np.float64)
    fn = np.asarray(fine_normals, np.float64)
    fn = fn / np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-12)
    dots = np.einsum("fkd,fd->fk", cn[idx], fn)
    ok = dots >= align
    pick = np.where(ok.any(1), np.argmax(ok, 1), 0)
    owner = idx[np.arange(len(fc)), pick]
    out = np.zeros((len(cc), 3))
    np.add.at(out, owner, np.asarray(fine_forces, np.float64))
    return out * scale


def source_floor(reynolds: float = 320.0, layout=(256, 128, 128), length: int = 64, u_in: float = 0.05,
                 flow_throughs: int = 6, average: int = 2, snapshots: int = 40, iterations: int = 3,
                 log=print) -> dict:
    """What a closure whose output is a momentum source at each wall cell (pressure and friction
    together) can do for Cd: the coarse run with the wall law; the 2x finer run; then the coarse
    run with per-wall-cell sources C = (its momentum-exchange force) - (the fine run's total wall
    force restricted to that cell, restrict_force), i.e. a network that predicts the fine grid's
    per-cell force correction exactly for the wall-law flow; the flow's response to the sources
    changes the exchange, so the sources are then updated from each run's own per-cell forces
    (C <- its exchange - the restricted fine force) for `iterations` runs in all."""
    run = lambda grid_layout, grid_length, prescribed=None, source=False: _floor_run(
        grid_layout, grid_length, reynolds, u_in, flow_throughs, average, snapshots, prescribed, source, log)
    fine_layout = tuple(2 * v for v in layout)
    coarse = run(layout, length)
    fine = run(fine_layout, 2 * length)
    normals = coarse["normals"] / np.maximum(np.linalg.norm(coarse["normals"], axis=1, keepdims=True), 1e-12)
    target = restrict_force(coarse["cells"], layout, normals, fine["cells"], fine_layout, fine["forces"],
                            fine["normals"], coarse["frontal"] / fine["frontal"])
    scale = 1.0 / (0.5 * u_in * u_in * coarse["frontal"])
    log(f"  restricted fine force: Cd {target[:, 0].sum() * scale:.6f} (fine run {fine['cd']:.6f}; "
        f"its time-averaged wall forces {fine['split']['cd']:.6f})")
    pct = lambda a: 100 * (a / fine["cd"] - 1)
    history = []
    exchange = coarse["forces"]
    valid = np.ones(len(target), bool)
    for it in range(iterations):
        sources = exchange - target
        res = run(layout, length, (sources, valid), source=True)
        history.append({"cd": res["cd"], "vs_2x_grid_pct": pct(res["cd"]), "cd_exchange": res["cd_exchange"],
                        "cd_sources": res["cd_sources"], "split_exchange": res["split"]})
        exchange = res["forces"]
    return {"reynolds": reynolds, "layout": list(layout), "length_cells": length,
            "cd_law": coarse["cd"], "cd_law_2x_grid": fine["cd"], "law_vs_2x_grid_pct": pct(coarse["cd"]),
            "cd_restricted_fine": float(target[:, 0].sum() * scale), "iterations": history,
            "split_law": coarse["split"], "split_2x_grid": fine["split"], "wall_cells": int(len(coarse["cells"]))}


def wall_rule_cd(reynolds: float = 320.0, layout=(256, 128, 128), length: int = 64, u_in: float = 0.05,
                 flow_throughs: int = 6, average: int = 2, log=print) -> list:
    """The slanted bluff body on `layout` and on the twice finer grid with each wall rule (halfway
    bounce-back, Bouzidi), with and without the LES wall model, REG collision, no closure: the
    window-mean Cd of both grids and the largest velocity component of the last state."""
    from . import bodies, drag
    from .gpu_solver import GpuSolver, SolverParams
    from .voxelize import voxelize

    def cd(grid_layout, grid_length, walls, les):
        body, _ = bodies.place_mesh(bodies.bluff_body(), None, grid_length, grid_layout, False)
        grid = voxelize(body, None, grid_layout, False, link_fractions=True)
        nu = u_in * grid_length / reynolds
        s = GpuSolver.create(grid_layout, grid, SolverParams(3 * nu + 0.5, u_in, les, operator="reg", walls=walls))
        try:
            flow = drag.flow_through_steps(grid.nx, u_in)
            samples = s.step_sampled(flow_throughs * flow, 8)
            fx, _, _ = drag.window_mean_force(samples, flow_throughs * flow, average * flow)
            s.step(1, True)
            umax = float(np.abs(s.read_macro()[:, 1:]).max())
        finally:
            s.destroy()
        return fx * drag.cd_per_lattice_force(grid.frontal_cells, u_in), umax

    rows = []
    fine_layout = tuple(2 * v for v in layout)
    for les in (True, False):
        for walls in ("bb", "bouzidi"):
            c, uc = cd(layout, length, walls, les)
            f, uf = cd(fine_layout, 2 * length, walls, les)
            rows.append({"les": les, "walls": walls, "cd": c, "cd_2x_grid": f, "vs_2x_grid_pct": 100 * (c / f - 1),
                         "max_velocity": uc, "max_velocity_2x_grid": uf})
            log(f"  LES {les}, walls {walls}: Cd {c:.5f}, 2x grid {f:.5f} ({100 * (c / f - 1):+.2f}%), "
                f"largest |u| {uc:.3f} / {uf:.3f}")
    return rows


def main(argv=None):
    """Train the closure on coarse/fine pairs of random bodies and save it.

    python -m pocketwind_amd.closure --out closure.npz --coarse 256 128 128 --length 64 \\
        --train 6 --held 2 --re 200 500
    """
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", required=True)
    ap.add_argument("--coarse", type=int, nargs=3, default=[256, 128, 128])
    ap.add_argument("--length", type=int, default=64)
    ap.add_argument("--train", type=int, default=6)
    ap.add_argument("--held", type=int, default=2)
    ap.add_argument("--re", type=float, nargs=2, default=[200.0, 500.0])
    ap.add_argument("--flow-throughs", type=float, default=4.0)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--report", default=None)
    ap.add_argument("--wall-rules", type=float, default=None, metavar="RE",
                    help="Cd of the bluff body on both grids with each wall rule, with and without LES")
    ap.add_argument("--source-floor", type=float, default=None, metavar="RE",
                    help="measure the Cd a perfect momentum-source closure gives at this Reynolds number")
    ap.add_argument("--shear-floor", type=float, default=None, metavar="RE",
                    help="instead of training, measure Cd with the fine grid's wall shear prescribed "
                         "on the coarse grid (a perfect closure output) beside the wall law and the "
                         "twice finer grid at this Reynolds number")
    ap.add_argument("--evaluate", type=float, default=None, metavar="RE",
                    help="instead of training, compare Cd of the learned closure with the wall law "
                         "and the twice finer grid at this Reynolds number (the closure file --out)")
    args = ap.parse_args(argv)
    if args.wall_rules is not None:
        rec = wall_rule_cd(args.wall_rules, tuple(args.coarse), args.length, log=lambda m: print(m, flush=True))
        print(json.dumps(rec))
        if args.report:
            with open(args.report, "w") as f:
                json.dump(rec, f, indent=1)
        return 0
    if args.source_floor is not None:
        rec = source_floor(args.source_floor, tuple(args.coarse), args.length, log=lambda m: print(m, flush=True))
        print(json.dumps(rec, indent=1, default=float))
        if args.report:
            with open(args.report, "w") as f:
                json.dump(rec, f, indent=1, default=float)
        return 0
    if args.shear_floor is not None:
        rec = shear_floor(args.shear_floor, tuple(args.coarse), args.length, log=lambda m: print(m, flush=True))
        print(json.dumps(rec))
        if args.report:
            json.dump(rec, open(args.report, "w"), indent=1)
        return 0
    if args.evaluate is not None:
        rec = evaluate_cd(args.evaluate, tuple(args.coarse), args.length, args.out,
                          log=lambda m: print(m, flush=True))
        print(json.dumps(rec))
        if args.report:
            json.dump(rec, open(args.report, "w"), indent=1)
        return 0
    t0 = time.time()
    pairs = []
    for k, seed in enumerate(list(range(args.train)) + [1000 + j for j in range(args.held)]):
        p = coarse_fine_pair(seed, tuple(args.coarse), args.length, tuple(args.re), flow_throughs=args.flow_throughs)
        pairs.append(p)
        print(f"# pair {k} (seed {seed}, Re {p['re']:.0f}): {len(p['target'])} wall cells, "
              f"{time.time() - t0:.0f} s", flush=True)
    train, held = pairs[:args.train], pairs[args.train:]
    x = np.concatenate([p["features"] for p in train])
    t = np.concatenate([p["target"] for p in train])
    net, norm, history = train_closure(x, t, steps=args.steps, batch=512, lr=2e-3, seed=1)
    rms = lambda a: float(np.sqrt(np.mean(np.sum(a ** 2, axis=1))))
    report = {"train_samples": int(len(t)), "final_loss": float(history[-100:].mean()),
              "coarse": args.coarse, "length": args.length, "re": args.re}
    if held:
        xh = np.concatenate([p["features"] for p in held])
        th = np.concatenate([p["target"] for p in held])
        law = np.concatenate([p["law"] for p in held])
        report.update(held_samples=int(len(th)), held_rms_target=rms(th),
                      held_rms_error_network=rms(predict(net, norm, xh) - th), held_rms_error_wall_law=rms(law - th))
    save(args.out, net, norm, report)
    print(json.dumps(report))
    if args.report:
        json.dump(report, open(args.report, "w"), indent=1)
    net.destroy()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
