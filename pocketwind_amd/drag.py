"""Drag coefficient accumulation in FP64 on the host.

* :func:`create_drag_stats` restates createDragStats (src/forces/drag-stats.ts:161-234): a
  one-flow-through window, two flow-throughs of settling, steady if the window's relative
  standard deviation is below 3 % and its half-to-half drift below 1 %, and a batch-means
  standard error (``sem_pct``, drag-stats.ts:120-146) that withholds Cd when fewer than five
  independent batches fit.
* :func:`cd_per_lattice_force` restates cdPerLatticeForce (src/lbm/units.ts:42-45) and
  :func:`window_mean_force` the validation average over the last flow-throughs
  (src/validation/runner.ts:174-179, :234-244).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

FORCE_EVERY_STEPS = 8
U_LU = 0.05
WINDOW_FLOW_THROUGHS = 1
SETTLING_FLOW_THROUGHS = 2
STEADY_REL_STD_PCT = 3.0
STEADY_DRIFT_PCT = 1.0
MIN_BATCHES = 5
MAX_BATCHES = 16
TOO_FEW_INDEPENDENT_SAMPLES = "too-few-independent-samples"


def _mean(values):
    if not values:
        return None
    s = 0.0
    for v in values:
        s += v
    return s / len(values)


def _sample_std(values):
    if len(values) < 2:
        return 0.0
    m = _mean(values)
    sq = 0.0
    for v in values:
        d = v - m
        sq += d * d
    return math.sqrt(sq / (len(values) - 1))


def _rel_std_pct(values):
    m = _mean(values)
    if m is None:
        return 0.0
    std = _sample_std(values)
    if m == 0:
        return 0.0 if std == 0 else math.inf
    return 100 * std / abs(m)


def _half_drift_pct(values):
    n = len(values)
    if n < 2:
        return 0.0
    half = n // 2
    first, second, m = _mean(values[:half]), _mean(values[half:]), _mean(values)
    if first is None or second is None or m is None:
        return 0.0
    change = abs(first - second)
    if m == 0:
        return 0.0 if change == 0 else math.inf
    return 100 * change / abs(m)


def tau_int(values, sample_spacing_steps: float = 1) -> float:
    """Integral time in steps from the autocorrelation summed to its first zero crossing
    (drag-stats.ts:80-108)."""
    if not math.isfinite(sample_spacing_steps) or sample_spacing_steps <= 0:
        raise ValueError("Sample spacing must be finite and positive.")
    n = len(values)
    if n < 2:
        return float(sample_spacing_steps)
    m = sum(values) / n
    centred = [v - m for v in values]
    variance = sum(c * c for c in centred)
    if variance == 0:
        return float(sample_spacing_steps)
    total = 1.0
    for lag in range(1, n // 2 + 1):
        cov = 0.0
        for i in range(n - lag):
            cov += centred[i] * centred[i + lag]
        if cov <= 0:
            break
        total += cov / variance
    return total * sample_spacing_steps


def batch_length(values, sample_spacing_steps: float = 1) -> float:
    return max(2 * tau_int(values, sample_spacing_steps), 8 * FORCE_EVERY_STEPS)


def sem_pct(values, sample_spacing_steps: float = 1, window_steps: Optional[float] = None):
    """Batch-means standard error of the mean in percent, or None (drag-stats.ts:120-146)."""
    n = len(values)
    if window_steps is None:
        window_steps = n * sample_spacing_steps
    m = _mean(values)
    if m is None or not math.isfinite(window_steps) or window_steps <= 0:
        return None
    length = batch_length(values, sample_spacing_steps)
    per_batch = max(1, round_half_up(length / sample_spacing_steps))
    available = n // per_batch
    batches = min(MAX_BATCHES, math.floor(window_steps / length), available)
    if batches < MIN_BATCHES:
        return None
    start = n - batches * per_batch
    means = []
    for b in range(batches):
        lo = start + b * per_batch
        s = 0.0
        for i in range(lo, lo + per_batch):
            s += values[i]
        means.append(s / per_batch)
    std = _sample_std(means)
    if m == 0:
        return 0.0 if std == 0 else math.inf
    return 100 * (std / math.sqrt(batches)) / abs(m)


def round_half_up(v: float) -> int:
    """JavaScript Math.round."""
    return math.floor(v + 0.5)


@dataclass
class DragSummary:
    cd: Optional[float]
    status: dict
    samples: int
    sem_pct: Optional[float] = None
    reason: Optional[str] = None


@dataclass
class DragStats:
    flow_through_steps: float
    cd_per_force: float = 1.0
    _samples: list = field(default_factory=list)

    def __post_init__(self):
        if not math.isfinite(self.flow_through_steps) or self.flow_through_steps <= 0:
            raise ValueError("Flow-through steps must be finite and positive.")
        if not math.isfinite(self.cd_per_force):
            raise ValueError("Cd per unit force must be finite.")

    @property
    def window_steps(self):
        return WINDOW_FLOW_THROUGHS * self.flow_through_steps

    @property
    def settle_steps(self):
        return SETTLING_FLOW_THROUGHS * self.flow_through_steps

    def push(self, step: float, fx: float) -> None:
        if not math.isfinite(step) or not math.isfinite(fx):
            raise ValueError("Drag samples require a finite step and force.")
        if self._samples and step <= self._samples[-1][0]:
            raise ValueError("Drag sample steps must strictly increase.")
        self._samples.append((step, fx))
        while self._samples and self._samples[0][0] <= step - self.settle_steps:
            self._samples.pop(0)

    def _spacing(self):
        s = self._samples
        if len(s) < 2:
            return FORCE_EVERY_STEPS
        spacing = (s[-1][0] - s[0][0]) / (len(s) - 1)
        return spacing if math.isfinite(spacing) and spacing > 0 else FORCE_EVERY_STEPS

    def summary(self, steps_per_s: Optional[float]) -> DragSummary:
        elapsed = self._samples[-1][0] if self._samples else 0
        values = [fx for step, fx in self._samples if step > elapsed - self.window_steps]
        all_values = [fx for _, fx in self._samples]
        m = _mean(values)
        sem = sem_pct(all_values, self._spacing(), self.settle_steps)
        cd = m * self.cd_per_force if (m is not None and sem is not None) else None
        reason = TOO_FEW_INDEPENDENT_SAMPLES if sem is None else None
        if elapsed < self.settle_steps:
            eta = None
            if steps_per_s is not None and math.isfinite(steps_per_s) and steps_per_s > 0:
                eta = max(0.0, (self.settle_steps - elapsed) / steps_per_s)
            return DragSummary(cd, {"kind": "settling", "etaS": eta}, len(values), sem, reason)
        fluct, drift = _rel_std_pct(values), _half_drift_pct(values)
        kind = "steady" if (fluct < STEADY_REL_STD_PCT and drift < STEADY_DRIFT_PCT) else "unsteady"
        return DragSummary(cd, {"kind": kind, "relStdPct": fluct, "driftPct": drift}, len(values), sem,
                           reason)


def create_drag_stats(flow_through_steps: float, cd_per_force: float = 1.0) -> DragStats:
    return DragStats(flow_through_steps, cd_per_force)


def cd_per_lattice_force(frontal_cells: float, u_lu: float = U_LU) -> float:
    """Cd = fx / (1/2 u^2 A) in lattice units (units.ts:42-45)."""
    if not math.isfinite(frontal_cells) or frontal_cells <= 0:
        raise ValueError("Frontal area in cells must be finite and positive.")
    return 1 / (0.5 * u_lu * u_lu * frontal_cells)


def flow_through_steps(nx: int, u_lu: float = U_LU) -> int:
    """validation/cases.ts:129-133."""
    if nx < 1:
        raise ValueError("Grid length must be positive.")
    if not (u_lu > 0):
        raise ValueError("Inlet speed must be positive.")
    return math.ceil(nx / u_lu)


def window_mean_force(samples, total_steps: int, average_steps: int):
    """Mean (fx, fy, fz) over samples with total - average < step <= total
    (runner.ts:174-179, :234-244)."""
    window = [s for s in samples if total_steps - average_steps < s[0] <= total_steps]
    if not window:
        raise ValueError("No force samples in the averaging window.")
    n = len(window)
    return tuple(sum(s[k] for s in window) / n for k in (1, 2, 3))
