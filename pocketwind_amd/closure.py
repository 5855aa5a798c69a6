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
