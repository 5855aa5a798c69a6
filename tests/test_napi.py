"""The Node-API boundary (pocketwind_amd/napi): a TypeScript/Node host's GpuSolver
(src/lbm/gpu-solver.ts:135-465) over the C-ABI.

CPU: tests/napi/host.test.js checks the ABI version, the device count, packSolverParams against
tests/unit/gpu-solver.test.ts:9-62 (with the layouts only JS can express: -1, 1.5, NaN,
Infinity) and that create rejects invalid input with a RangeError before any device call.

GPU: tests/napi/gpu_run.js steps a REG + LES sphere tunnel (64 x 32 x 32, f32 and f16) in blocks
of FORCE_EVERY_STEPS with a force sample after each; its populations, macros and force samples
equal those of the same sequence through the Python host bit for bit (both call one library),
and its f32 populations match the FP64 oracle within the spec's 1e-5 (gpu-vs-cpu.spec.ts:14-19).
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from grids import grid_from_kinds
from pocketwind_amd import build
from pocketwind_amd.gpu_solver import GpuSolver, SolverParams
from test_gpu_row_parity import sphere_tunnel_kinds
from test_gpu_solver import normalized_peak

HERE = os.path.dirname(os.path.abspath(__file__))
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.isfile(build.NAPI_ADDON),
                                reason="node or the built pw_napi.node addon is absent")


def run_node(*args, env=None, timeout=120):
    return subprocess.run([NODE, *args], capture_output=True, text=True, timeout=timeout,
                          env={**os.environ, **(env or {})})


@needs_node
def test_host_side_boundary_without_a_device():
    env = {"PW_EXPECT_DEVICES": "0"} if not os.path.exists("/dev/kfd") else {}
    r = run_node(os.path.join(HERE, "napi", "host.test.js"), env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "13/13 passed" in r.stdout


LAYOUT = (64, 32, 32)
BLOCKS = 13  # 104 steps


def python_run(grid, params):
    solver = GpuSolver.create(LAYOUT, grid, params)
    try:
        samples = []
        for k in range(BLOCKS):
            solver.step(8, k == BLOCKS - 1)
            solver.encode_force()
            samples.append(solver.read_force())
        return samples, solver.read_populations(), solver.read_macro()
    finally:
        solver.destroy()


@pytest.mark.gpu
@needs_node
@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_node_host_equals_the_python_host_and_the_oracle(tmp_path, precision):
    grid = grid_from_kinds(LAYOUT, sphere_tunnel_kinds(LAYOUT))
    (tmp_path / "geometry.u32").write_bytes(np.ascontiguousarray(grid.geometry, np.uint32).tobytes())
    (tmp_path / "boundary.u32").write_bytes(np.ascontiguousarray(grid.boundary_cells, np.uint32).tobytes())
    spec = {"layout": dict(zip("nx ny nz".split(), LAYOUT)), "blocks": BLOCKS,
            "params": {"tau0": 0.8, "uIn": 0.06, "les": True, "operator": "reg", "precision": precision}}
    (tmp_path / "case.json").write_text(json.dumps(spec))
    print("node run", precision, flush=True)
    r = run_node(os.path.join(HERE, "napi", "gpu_run.js"), str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    result = json.loads((tmp_path / "result.json").read_text())
    assert result["busyBeforeEncode"] == "busy"
    assert result["stepCount"] == 8 * BLOCKS and result["stepCountAfterReset"] == 0
    assert "destroyed" in result["stepAfterDestroy"]

    samples, pops, macro = python_run(grid, SolverParams(0.8, 0.06, True, operator="reg", precision=precision))
    node_pops = np.frombuffer((tmp_path / "populations.bin").read_bytes(), pops.dtype).reshape(pops.shape)
    node_macro = np.frombuffer((tmp_path / "macro.f32").read_bytes(), np.float32).reshape(macro.shape)
    word = np.uint16 if pops.itemsize == 2 else np.uint32
    assert np.array_equal(node_pops.view(word), pops.view(word))
    assert np.array_equal(node_macro.view(np.uint32), macro.view(np.uint32))
    assert [(s["step"], s["fx"], s["fy"], s["fz"]) for s in result["samples"]] == \
        [(s.step, s.fx, s.fy, s.fz) for s in samples]

    if precision == "f32":
        cpu = oracle.Solver(LAYOUT, 0.8, oracle.REG, True, oracle.BB, flags=grid.geometry[:grid.cells],
                            equilibrium_state=(1.0, (0.06, 0.0, 0.0)), threads=8)
        cpu.step(8 * BLOCKS)
        err = normalized_peak(node_pops, cpu.populations.copy().reshape(19, grid.cells))
        print("node f32 populations vs oracle:", err)
        assert err < 1e-5
