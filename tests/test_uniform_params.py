"""The solver uniform against tests/unit/gpu-solver.test.ts:9-62: pw_pack_params (the host-only
restatement of packSolverParams, src/lbm/gpu-solver.ts:80-116) writes the reference's 48-byte
struct Params with the same integer and f32 values, and rejects the same inputs. The values
asserted are the reference test's own: nu = fround(0.1) at tau0 0.8, cs_term =
fround(18 sqrt(2) 0.01), lambda = 3/16, and the first f32 relaxation time above one half."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from pocketwind_amd import _native as N

LAYOUT = (32, 16, 8)


def params(tau0=0.8, u_in=0.04, les=False, walls="bb"):
    return N.Params(tau0=tau0, u_in=u_in, les=int(les), op=0, walls=1 if walls == "bouzidi" else 0, storage=0,
                    ramp_steps=0, wall_model=0, sides=0, outlet=0)


def pack(layout=LAYOUT, write_macro=False, step=0, **kw):
    out = np.zeros(12, np.uint32)
    status = N.lib().pw_pack_params(C.byref(N.Layout(*layout)), C.byref(params(**kw)), int(write_macro),
                                    step, out.ctypes.data, out.nbytes)
    return status, out


def fround(v: float) -> float:
    return float(np.float32(v))


def test_packs_the_twelve_scalars_with_their_integer_and_float_representations():
    status, data = pack(write_macro=True)
    assert status == N.PW_OK
    floats = data.view(np.float32)
    assert list(data[:4]) == [32, 16, 8, 4096]
    assert floats[4] == fround(0.04)
    assert floats[5] == fround(0.1)
    assert floats[6] == fround(18 * math.sqrt(2) * 0.1 ** 2)
    assert floats[7] == 3 / 16
    assert list(data[8:]) == [0, 0, 1, 0]


def test_macro_output_and_les_are_independent():
    assert list(pack(les=True)[1][8:]) == [1, 0, 0, 0]


@pytest.mark.parametrize("walls,mode", [("bb", 0), ("bouzidi", 1)])
def test_walls_pack_as_their_mode(walls, mode):
    assert list(pack(walls=walls)[1][8:]) == [0, mode, 0, 0]


def test_step_packs_into_the_last_word():
    assert pack(step=12345)[1][11] == 12345


@pytest.mark.parametrize("nx", [0])
def test_rejects_an_empty_grid_dimension(nx):
    # gpu-solver.test.ts:35-37 (-1, 1.5, NaN and Infinity cannot reach a u32 layout)
    assert pack(layout=(nx, 16, 8))[0] == N.PW_EINVAL


@pytest.mark.parametrize("tau0", [0.5, 0, math.nan, math.inf, 1e100, 0.5 + np.finfo(float).eps, 0.5 + 2 ** -26])
def test_rejects_invalid_relaxation_times(tau0):
    assert pack(tau0=tau0)[0] == N.PW_EINVAL
    assert "Relaxation time" in N.last_error()


@pytest.mark.parametrize("u_in", [1e100, -1e100, math.inf, -math.inf, math.nan])
def test_rejects_speeds_that_overflow_f32(u_in):
    assert pack(u_in=u_in)[0] == N.PW_EINVAL


def test_accepts_the_first_f32_relaxation_time_above_one_half():
    status, data = pack(tau0=0.5 + 2 ** -24)
    assert status == N.PW_OK
    nu = float(data.view(np.float32)[5])
    assert fround(fround(3 * nu) + 0.5) == 0.5 + 2 ** -24


def test_rejects_population_indices_outside_u32():
    assert pack(layout=(65536, 65536, 1))[0] == N.PW_EINVAL


def test_planted_viscosity_mistake_is_caught():
    """nu computed in double and rounded once (no f32 tau0 first) differs from the reference's
    packing at a tau0 that is not an f32 value; the packed word must be the reference's."""
    tau0 = 0.800000016
    ref = fround((fround(tau0) - 0.5) / 3)
    wrong = fround((tau0 - 0.5) / 3)
    assert ref != wrong
    assert pack(tau0=tau0)[1].view(np.float32)[5] == ref
