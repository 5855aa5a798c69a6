"""f32, f16 and bf16 storage against the FP64 oracle on the 80 x 40 x 40 bluff-body sub-case
(tests/golden/make_bluff_subcase.py): the slanted body at length 20 cells, REG + LES with the
wall law, Re 320, 30 000 steps. Bars: Cd within 1 % of the oracle's for every storage (the
benchmark's own drag criterion); the open cells' total density drifting by no more than 1e-4
more than the oracle's; the window-averaged centreline x-velocity within 0.1 % of u_in at every
point for f32, for f16 within 1 % of u_in in RMS over the line and 5 % at any point, and for
bf16, whose 8 significant bits are 8 times coarser than f16's, within 2 % and 10 % (the same
factor of about ten the seeded-blob bars allow bf16 over f16). 16-bit storage rounds every population's deviation stochastically each step; the
wall-law cells just upstream of the nose (a stagnation region where the velocity falls from
0.03 to 0 over three cells) respond to that noise through the law's nonlinear friction
velocity, so single points there move more than the line as a whole.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from pocketwind_amd import bodies
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from pocketwind_amd.voxelize import voxelize

pytestmark = pytest.mark.gpu

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bluff_subcase_oracle.json")))


@pytest.mark.parametrize("precision", ["f32", "f16", "bf16"])
def test_storage_against_oracle_on_the_bluff_subcase(precision):
    layout = tuple(G["layout"])
    nx, ny, nz = layout
    body, _ = bodies.place_mesh(bodies.bluff_body(), None, G["length_cells"], layout, False)
    grid = voxelize(body, None, layout, False, link_fractions=False)
    assert grid.solid_cells == G["solid_cells"] and grid.frontal_cells == G["frontal_cells"]
    kinds = grid.flags & 7
    open_cells = (kinds == 0) | (kinds == 4)
    line = [x + nx * (ny // 2 + ny * (nz // 2)) for x in range(nx)]
    u = G["u_in"]
    s = GpuSolver.create(layout, grid, SolverParams(G["tau0"], u, True, operator="reg", precision=precision))
    try:
        mass0 = float(s.read_macro()[open_cells, 0].astype(np.float64).sum())
        s.step_sampled(G["steps"] - G["average_steps"], G["sample_every"])
        forces, profiles = [], []
        while s.step_count < G["steps"]:
            every = G["profile_every"]
            chunk = every - s.step_count % every - 1
            forces.append(s.step_sampled(chunk, G["sample_every"]))
            s.step(1, True)
            if s.step_count % G["sample_every"] == 0:
                s.encode_force()
                f = s.read_force()
                forces.append(np.array([[f.step, f.fx, f.fy, f.fz]]))
            profiles.append(s.read_macro()[line, 1].astype(np.float64))
        macro = s.read_macro()
    finally:
        s.destroy()
    samples = np.concatenate(forces)
    assert samples.shape[0] == G["average_steps"] // G["sample_every"]
    cd = samples[:, 1].mean() / (0.5 * u ** 2 * G["frontal_cells"])
    profile = np.mean(profiles, axis=0)
    ref_profile = np.asarray(G["centreline_ux"])
    drift = float(macro[open_cells, 0].astype(np.float64).sum()) / mass0 - 1
    print(f"{precision}: Cd {cd:.6f} oracle {G['cd']:.6f} ({100 * (cd / G['cd'] - 1):+.3f} %), "
          f"centreline max |du|/u {np.abs(profile - ref_profile).max() / u:.2e}, "
          f"mass drift {drift:+.3e} oracle {G['mass_drift']:+.3e}")
    assert len(profiles) == G["profile_samples"]
    assert abs(cd / G["cd"] - 1) < 0.01
    err = np.abs(profile - ref_profile)
    print(f"  centreline RMS |du|/u {np.sqrt(np.mean(err ** 2)) / u:.2e}")
    rms_bar, max_bar = {"f32": (1e-3, 1e-3), "f16": (0.01, 0.05), "bf16": (0.02, 0.10)}[precision]
    assert np.sqrt(np.mean(err ** 2)) < rms_bar * u and err.max() < max_bar * u
    assert abs(drift) <= abs(G["mass_drift"]) + 1e-4
