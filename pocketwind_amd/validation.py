"""Validation harness: case specifications, result rows, grid convergence and scoring, as the
reference defines them, so that this solver's drag rows are judged by the same rules and can
be written in the same file shape.

grid_convergence   gridConvergence (src/forces/convergence.ts:29-48): observed order by
                   fixed-point iteration, Richardson extrapolation, fine-grid GCI (Fs 1.25)
grid_band          gridBand (src/forces/grid-band.ts:14-22)
CaseSpec and the case builders
                   src/validation/cases.ts:45-627 (Ahmed, laminar sphere, shape classes,
                   confinement, sensitivity and smoke cases; flow-through and window lengths)
row_from_run       rowFromRun (src/validation/runner.ts:258-321): window mean of the force,
                   Cd on the stated and voxel areas, Maskell bound, reference error, pass rules
case_scales, case_meshes, case_grid, run_case
                   one case run on this solver as runCase does it (src/app/case-run.ts:97-250,
                   src/validation/links.ts:14-35): Reynolds number and lattice viscosity, the
                   placed and voxelised preset with its floor and stilts, forces in newtons
                   every 8 steps
run_body_drag      runBodyDrag (src/lbm/reference/cases.ts:264-321) as scripts/drag-cpu.ts runs
                   it, on this solver: Cd per flow-through of the Ahmed body at 64 x 32 x 32
run_grid_check     the app's grid check (src/app/grid-check.ts, case-run.ts:257-320) on this
                   solver: S, M, L at the live run's Re and body share, band from the means
score_results      scripts/score-results.mjs: match rows to reference values by case
main               `python -m pocketwind_amd.validation score <file>`, and
                   `python -m pocketwind_amd.validation run <out.json> [ids] [--tiers ...]`
"""
from __future__ import annotations

import json
import math
import sys
import time
from dataclasses import asdict, dataclass, field, replace
from decimal import ROUND_HALF_UP, Decimal
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import drag
from .field_export import js_number
from .lattice import TIERS

# src/lbm/units.ts, src/forces/physical.ts, src/validation/cases.ts:80-113
U_LU = 0.05
NU_AIR = 1.5e-5
RHO_AIR = 1.204
FORCE_SAMPLE_STEPS = 8
MASKELL_THETA = 2.5
FLOW_THROUGHS = 6
AVERAGE_THROUGHS = 2
PASS_TOLERANCE_PCT = 10
AHMED_25_REFERENCE_CD = 0.285
AHMED_35_REFERENCE_CD = None
SPHERE_RE_100_REFERENCE_CD = 1.09
AHMED_FRONTAL_AREA_M2 = 0.112
AHMED_BODY_LENGTH_MM = 1044
AHMED_VALIDATION_SPEED_MPS = 40
AHMED_BODY_LENGTH_FRACTION = 0.5
CUBE_SIDE_FRACTION = 0.2
SPHERE_DIAMETER_FRACTION = 0.226
LAMINAR_SPHERE_DIAMETER_MM = 100
LAMINAR_SPHERE_RE = 100
SHAPE_SPHERE_RE = 20000
SHAPE_SPEED_MPS = 40
DEFAULT_TIERS = ("S", "M", "L", "XL")
SHAPE_TIERS = ("S", "M", "L")
OPERATORS = ("trt", "reg")
DEFAULT_WALLS = "bb"
DEFAULT_STORAGE = "f32"

# Reference drag coefficients by case (scripts/score-results.mjs); None: known, no value.
REFERENCES: Dict[str, Optional[float]] = {
    "ahmed-25": 0.285, "ahmed-35": None, "sphere-re-100": 1.09, "cube": None, "sphere": None}
SCORED_SECTIONS = ("laminar", "validation")


def default_operator(les: bool) -> str:
    """src/lbm/params.ts:15: REG with the sub-grid model, TRT without."""
    return "reg" if les else "trt"


# --------------------------------------------------------------------------------------------
# Grid convergence

MAX_ROUNDS = 100
SETTLE_TOLERANCE = 1e-10
SAFETY_FACTOR = 1.25


def _sign(x: float) -> float:
    return float(x > 0) - float(x < 0)


def _observed_order(ratio: float, r21: float, r32: float, s: float) -> Optional[float]:
    """Fixed-point iteration for p in |ln|e32/e21| + ln((r21^p - s)/(r32^p - s))| / ln r21;
    the closed form holds only for equal refinement ratios."""
    order = 1.0
    for _ in range(MAX_ROUNDS):
        try:
            q = math.log((r21 ** order - s) / (r32 ** order - s))
        except (ValueError, ZeroDivisionError):
            return None
        following = abs(math.log(abs(ratio)) + q) / math.log(r21)
        if abs(following - order) < SETTLE_TOLERANCE:
            return following
        order = following
    return None


def grid_convergence(rows: Sequence[Dict[str, float]]) -> Dict[str, Union[str, float]]:
    """gridConvergence: rows of {h, value}; only the three finest levels are used."""
    ordered = sorted(({"h": float(r["h"]), "value": float(r["value"])} for r in rows), key=lambda r: r["h"])
    if len(ordered) < 3:
        return {"kind": "insufficient"}
    fine, mid, coarse = ordered[:3]
    r21 = mid["h"] / fine["h"]
    r32 = coarse["h"] / mid["h"]
    e21 = mid["value"] - fine["value"]
    e32 = coarse["value"] - mid["value"]
    if e21 == 0 or e32 == 0:
        return {"kind": "insufficient"}
    ratio = e32 / e21
    if ratio < 0:
        return {"kind": "oscillatory"}
    if abs(e21) >= abs(e32):  # differences growing toward the fine grid
        return {"kind": "divergent"}
    order = _observed_order(ratio, r21, r32, _sign(ratio))
    if order is None or not math.isfinite(order):
        return {"kind": "divergent"}
    factor = r21 ** order - 1
    extrapolated = (r21 ** order * fine["value"] - mid["value"]) / factor
    gci = 100 * SAFETY_FACTOR * abs((fine["value"] - mid["value"]) / fine["value"]) / factor
    return {"kind": "monotone", "order": order, "extrapolated": extrapolated, "gciFinePct": gci}


def grid_band(rows: Sequence[Dict[str, float]]) -> Dict[str, Union[str, float]]:
    """gridBand: rows of {h, cd}; a band only for a monotone sequence of usable rows."""
    usable = all(math.isfinite(r["cd"]) and r["h"] > 0 and math.isfinite(r["h"]) for r in rows)
    if not usable:
        return {"kind": "no-band", "reason": "insufficient"}
    result = grid_convergence([{"h": r["h"], "value": r["cd"]} for r in rows])
    if result["kind"] != "monotone":
        return {"kind": "no-band", "reason": result["kind"]}
    return {"kind": "band", "gciPct": result["gciFinePct"], "extrapolatedCd": result["extrapolated"]}


# --------------------------------------------------------------------------------------------
# Cases

def cd_per_newton(area_m2: float, speed_mps: float, density: float = RHO_AIR) -> float:
    """cdPerNewton (physical.ts:48-51): 1 / (q A)."""
    for v, name in ((area_m2, "Reference area in square metres"), (speed_mps, "Speed in metres per second"),
                    (density, "Air density in kilograms per cubic metre")):
        if not math.isfinite(v) or v <= 0:
            raise ValueError(f"{name} must be finite and positive.")
    return 1 / (0.5 * density * speed_mps * speed_mps * area_m2)


@dataclass(frozen=True)
class CaseSpec:
    """CaseSpec (cases.ts:45-80)."""

    case: str
    id: str
    group: str
    label: str
    tier: str
    grid: Dict[str, int]
    cells: int
    operator: str
    walls: str
    storage: str
    variant: Optional[str]
    sensitivity: Optional[str]
    shape: str
    shapeClass: str
    speedMps: float
    re: Optional[float]
    les: bool
    floor: bool
    stilts: bool
    bodyLengthMm: float
    bodyLengthCells: float
    cellSizeM: float
    referenceAreaM2: float
    voxelAreaM2: float
    blockage: float
    baselineBlockage: Optional[float]
    uLu: float
    steps: int
    flowThroughSteps: int
    averageSteps: int
    referenceCd: Optional[float]
    isDefault: bool


def _grid(tier: str) -> Dict[str, int]:
    t = TIERS[tier]
    return {"nx": t.nx, "ny": t.ny, "nz": t.nz}


def _default_ahmed_length(tier: str) -> float:
    return TIERS[tier].nx * AHMED_BODY_LENGTH_FRACTION


def is_default_row(spec: CaseSpec) -> bool:
    """isDefaultRow (cases.ts:172-206) with the app's defaults (bb walls, f32 storage, the
    1044 mm Ahmed body at half the streamwise domain)."""
    factors = (spec.operator == default_operator(spec.les) and spec.walls == DEFAULT_WALLS
               and spec.storage == DEFAULT_STORAGE)
    if spec.group in ("sensitivity", "smoke"):
        return False
    if spec.group != "validation":
        return factors
    if not spec.shape.startswith("ahmed-"):
        return False
    return (factors and spec.bodyLengthMm == AHMED_BODY_LENGTH_MM
            and spec.bodyLengthCells == _default_ahmed_length(spec.tier) and spec.grid == _grid(spec.tier))


def _case_spec(area_m2: float, steps: Optional[int] = None, flow_steps: Optional[int] = None, **fields) -> CaseSpec:
    """caseSpec (cases.ts:225-243): cells, areas, blockage and step counts from the inputs."""
    grid = fields["grid"]
    flow = flow_steps if flow_steps is not None else drag.flow_through_steps(grid["nx"], fields["uLu"])
    cell = fields["cellSizeM"]
    voxel_area = max(1, drag.round_half_up(area_m2 / (cell * cell))) * cell * cell
    spec = CaseSpec(cells=grid["nx"] * grid["ny"] * grid["nz"], referenceAreaM2=area_m2, voxelAreaM2=voxel_area,
                    blockage=area_m2 / (grid["ny"] * grid["nz"] * cell * cell),
                    steps=steps if steps is not None else FLOW_THROUGHS * flow, flowThroughSteps=flow,
                    averageSteps=AVERAGE_THROUGHS * flow, isDefault=False, **fields)
    return replace(spec, isDefault=is_default_row(spec))


def ahmed_case(tier: str, operator: str, walls: str = DEFAULT_WALLS, storage: str = DEFAULT_STORAGE, *,
               body_length_cells: Optional[float] = None, u_lu: float = U_LU, stilts: bool = True,
               variant: Optional[str] = None, sensitivity: Optional[str] = None, group: str = "validation",
               baseline_blockage: Optional[float] = None, reference_cd="default", case_id: Optional[str] = None,
               preset: str = "ahmed-25", steps: Optional[int] = None) -> CaseSpec:
    """ahmedCase (cases.ts:245-296)."""
    length = body_length_cells if body_length_cells is not None else _default_ahmed_length(tier)
    return _case_spec(
        AHMED_FRONTAL_AREA_M2, steps=steps,
        case=preset, id=case_id or f"{preset}-{tier}-{operator}-{variant or 'default'}", group=group,
        label=f"Ahmed body at {25 if preset == 'ahmed-25' else 35} degrees", tier=tier, grid=_grid(tier),
        operator=operator, walls=walls, storage=storage, variant=variant, sensitivity=sensitivity, shape=preset,
        shapeClass="road-vehicle", speedMps=AHMED_VALIDATION_SPEED_MPS,
        re=AHMED_VALIDATION_SPEED_MPS * AHMED_BODY_LENGTH_MM / 1000 / NU_AIR, les=True, floor=True, stilts=stilts,
        bodyLengthMm=AHMED_BODY_LENGTH_MM, bodyLengthCells=length, cellSizeM=AHMED_BODY_LENGTH_MM / 1000 / length,
        baselineBlockage=baseline_blockage, uLu=u_lu,
        referenceCd=AHMED_25_REFERENCE_CD if reference_cd == "default" else reference_cd)


def laminar_sphere_case(tier: str, operator: str, walls: str = DEFAULT_WALLS,
                        storage: str = DEFAULT_STORAGE) -> CaseSpec:
    """laminarSphereCase (cases.ts:298-338): D = nx / 8 cells stands for 100 mm at Re 100."""
    grid = _grid(tier)
    d_cells = grid["nx"] / 8
    cell = LAMINAR_SPHERE_DIAMETER_MM / 1000 / d_cells
    d_m = d_cells * cell
    return _case_spec(
        math.pi * (d_m / 2) ** 2,
        case="sphere-re-100", id=f"sphere-re-100-{tier}-{operator}", group="laminar", label="Sphere at Re 100",
        tier=tier, grid=grid, operator=operator, walls=walls, storage=storage, variant=None, sensitivity=None,
        shape="sphere", shapeClass="laminar", speedMps=LAMINAR_SPHERE_RE * NU_AIR / d_m, re=LAMINAR_SPHERE_RE,
        les=False, floor=False, stilts=False, bodyLengthMm=LAMINAR_SPHERE_DIAMETER_MM, bodyLengthCells=d_cells,
        cellSizeM=cell, baselineBlockage=None, uLu=U_LU, referenceCd=SPHERE_RE_100_REFERENCE_CD)


def shape_case(kind: str, tier: str) -> CaseSpec:
    """shapeCase (cases.ts:340-374): one metre across the tunnel, 4 % blockage."""
    grid = _grid(tier)
    cell = 1 / grid["ny"]
    size = (CUBE_SIDE_FRACTION if kind == "cube" else SPHERE_DIAMETER_FRACTION) * grid["ny"]
    d_m = size * cell
    area = d_m ** 2 if kind == "cube" else math.pi * (d_m / 2) ** 2
    return _case_spec(
        area, case=kind, id=f"{kind}-{tier}", group="shape-class",
        label="Cube face-on" if kind == "cube" else "Sphere at Re 20000", tier=tier, grid=grid,
        operator=default_operator(True), walls=DEFAULT_WALLS, storage=DEFAULT_STORAGE, variant=None,
        sensitivity=None, shape=kind, shapeClass="sharp-edged" if kind == "cube" else "smooth-subcritical",
        speedMps=SHAPE_SPHERE_RE * NU_AIR / d_m if kind == "sphere" else SHAPE_SPEED_MPS,
        re=SHAPE_SPHERE_RE if kind == "sphere" else None, les=True, floor=False, stilts=False,
        bodyLengthMm=d_m * 1000, bodyLengthCells=size, cellSizeM=cell, baselineBlockage=None, uLu=U_LU,
        referenceCd=None)


def confinement_case(kind: str, operator: str) -> CaseSpec:
    """confinementSphereCase / confinementCubeCase (cases.ts:376-438): the tier-L body in the
    XL domain."""
    grid = _grid("XL")
    if kind == "sphere":
        d_cells = TIERS["L"].nx / 8
        cell = LAMINAR_SPHERE_DIAMETER_MM / 1000 / d_cells
        d_m = d_cells * cell
        return _case_spec(
            math.pi * (d_m / 2) ** 2, case="sphere-confinement", id=f"sphere-confinement-{operator}",
            group="confinement", label="Sphere at Re 100 in the larger domain", tier="XL", grid=grid,
            operator=operator, walls=DEFAULT_WALLS, storage=DEFAULT_STORAGE, variant=None, sensitivity=None,
            shape="sphere", shapeClass="laminar", speedMps=LAMINAR_SPHERE_RE * NU_AIR / d_m, re=LAMINAR_SPHERE_RE,
            les=False, floor=False, stilts=False, bodyLengthMm=LAMINAR_SPHERE_DIAMETER_MM, bodyLengthCells=d_cells,
            cellSizeM=cell, baselineBlockage=None, uLu=U_LU, referenceCd=SPHERE_RE_100_REFERENCE_CD)
    cell = 1 / TIERS["L"].ny
    size = CUBE_SIDE_FRACTION * TIERS["L"].ny
    return _case_spec(
        (size * cell) ** 2, case="cube-confinement", id=f"cube-confinement-{operator}", group="confinement",
        label="Cube face-on in the larger domain", tier="XL", grid=grid, operator=operator, walls=DEFAULT_WALLS,
        storage=DEFAULT_STORAGE, variant=None, sensitivity=None, shape="cube", shapeClass="sharp-edged",
        speedMps=SHAPE_SPEED_MPS, re=None, les=True, floor=False, stilts=False, bodyLengthMm=size * cell * 1000,
        bodyLengthCells=size, cellSizeM=cell, baselineBlockage=None, uLu=U_LU, referenceCd=None)


def sensitivity_cases() -> List[CaseSpec]:
    """sensitivityRows (cases.ts:497-545)."""
    op = default_operator(True)
    length = _default_ahmed_length("L")
    baseline = ahmed_case("L", op)
    common = dict(group="sensitivity", variant=None, reference_cd=None)
    return [
        ahmed_case("XL", op, sensitivity="domain", body_length_cells=length, baseline_blockage=baseline.blockage,
                   case_id="sensitivity-domain", **common),
        ahmed_case("L", op, sensitivity="body-length", body_length_cells=0.3 * TIERS["L"].nx,
                   case_id="sensitivity-body-length-0.3", **common),
        ahmed_case("M", op, sensitivity="inlet-speed", u_lu=0.1, case_id="sensitivity-u-lu-0.1", **common),
        ahmed_case("L", op, sensitivity="stilts", stilts=True, case_id="sensitivity-stilts-included", **common),
        ahmed_case("L", op, sensitivity="stilts", stilts=False, case_id="sensitivity-stilts-excluded", **common),
    ]


def smoke_case() -> CaseSpec:
    """smokeCase (cases.ts:587-595): tier S at the defaults for two flow-throughs."""
    return ahmed_case("S", default_operator(True), group="smoke", case_id="smoke-ahmed-25-S", reference_cd=None,
                      steps=AVERAGE_THROUGHS * drag.flow_through_steps(TIERS["S"].nx))


def build_cases(tiers: Sequence[str] = DEFAULT_TIERS, operators: Sequence[str] = OPERATORS,
                walls: Sequence[str] = (DEFAULT_WALLS,), storages: Sequence[str] = ("f32", "f16")) -> List[CaseSpec]:
    """buildCases (cases.ts:552-579). Sorted by group then id (case-insensitively, which is how
    the reference's locale comparison orders these ids)."""
    op, wl, st = default_operator(True), DEFAULT_WALLS, DEFAULT_STORAGE
    rows: List[CaseSpec] = []
    for tier in tiers:
        if tier != "L":
            rows.append(ahmed_case(tier, op, wl, st))
            continue
        rows.append(ahmed_case(tier, op, wl, st, variant="default"))
        for factor, levels, current in (("operator", operators, op), ("walls", walls, wl), ("storage", storages, st)):
            other = [v for v in levels if v != current] if len(levels) == 2 else []
            if other:
                kw = {"operator": op, "walls": wl, "storage": st, factor: other[0]}
                rows.append(ahmed_case("L", kw["operator"], kw["walls"], kw["storage"], variant=f"flip-{factor}",
                                       case_id=f"ahmed-25-L-{factor}-{other[0]}"))
    if "L" in tiers:
        rows += sensitivity_cases()
    for operator in operators:
        rows += [laminar_sphere_case(t, operator) for t in SHAPE_TIERS if t in tiers]
        rows += [confinement_case("cube", operator), confinement_case("sphere", operator)]
    for tier in SHAPE_TIERS:
        if tier in tiers:
            rows += [shape_case("cube", tier), shape_case("sphere", tier)]
    return sorted(rows, key=lambda r: (r.group, r.id.lower()))


# --------------------------------------------------------------------------------------------
# Runs and rows

@dataclass(frozen=True)
class ForceSample:
    step: int
    fx: float
    fy: float
    fz: float


@dataclass
class SolverRun:
    """SolverRun (runner.ts:25-34): force samples in newtons."""

    forces: List[ForceSample]
    steps: int
    msPerStep: float
    voxelAreaM2: Optional[float] = None
    blockage: Optional[float] = None
    adapter: Optional[dict] = None
    userAgent: Optional[str] = None


@dataclass
class RunContext:
    commit: str
    date: str
    adapter: dict = field(default_factory=dict)
    userAgent: str = ""


def _check_run(spec: CaseSpec, run: SolverRun) -> None:
    if not isinstance(run.steps, int) or run.steps < spec.steps:
        raise ValueError(f"Case {spec.id} ran {run.steps} steps, expected at least {spec.steps}.")
    if not math.isfinite(run.msPerStep):
        raise ValueError("Milliseconds per step must be finite.")
    for s in run.forces:
        if not all(math.isfinite(v) for v in (s.step, s.fx, s.fy, s.fz)):
            raise ValueError(f"Case {spec.id} returned a non-finite force sample.")


def _mean(values: List[float]) -> float:
    total = 0.0
    for v in values:
        total += v
    return total / len(values)


def _fluctuations(spec: CaseSpec, run: SolverRun) -> Tuple[float, float]:
    stats = drag.create_drag_stats(spec.flowThroughSteps)
    for s in run.forces:
        stats.push(s.step, s.fx)
    status = stats.summary(None).status
    if status["kind"] == "settling":
        return 0.0, 0.0
    return status["relStdPct"], status["driftPct"]


def _evaluate(cd, reference_cd, error_pct, rel_std, drift) -> Tuple[bool, List[str]]:
    reasons = []
    if not math.isfinite(cd):
        reasons.append("Cd is not finite")
    if not math.isfinite(rel_std):
        reasons.append("relative standard deviation is not finite")
    if not math.isfinite(drift):
        reasons.append("drift is not finite")
    if reference_cd is not None:
        if error_pct is None or not math.isfinite(error_pct):
            reasons.append("reference error is missing")
        elif abs(error_pct) > PASS_TOLERANCE_PCT:
            reasons.append(f"reference error {to_fixed(error_pct, 1)}% exceeds {PASS_TOLERANCE_PCT}%")
    return not reasons, reasons


def maskell(cd_raw: float, blockage: float) -> float:
    """Maskell's blockage bound with theta 2.5 (runner.ts:185-187)."""
    return cd_raw / (1 + MASKELL_THETA * cd_raw * blockage)


def row_from_run(spec: CaseSpec, run: SolverRun, context: RunContext) -> dict:
    """rowFromRun: the reported row of one run, its mean force over the last two flow-throughs
    and its history at the 8-step readback cadence."""
    _check_run(spec, run)
    window = [s for s in run.forces if spec.steps - spec.averageSteps < s.step <= spec.steps]
    if not window:
        raise ValueError(f"Case {spec.id} returned no force samples in the averaging window.")
    fx, fy, fz = (_mean([getattr(s, k) for s in window]) for k in ("fx", "fy", "fz"))
    cd = fx * cd_per_newton(spec.referenceAreaM2, spec.speedMps)
    voxel_area = run.voxelAreaM2 if run.voxelAreaM2 is not None else spec.voxelAreaM2
    cd_voxel = fx * cd_per_newton(voxel_area, spec.speedMps)
    rel_std, drift = _fluctuations(spec, run)
    error_pct = None if spec.referenceCd is None else 100 * (cd - spec.referenceCd) / spec.referenceCd
    passed, reasons = _evaluate(cd, spec.referenceCd, error_pct, rel_std, drift)
    measured_shape = spec.shape in ("cube", "sphere")
    blockage = run.blockage if run.blockage is not None else spec.blockage
    row = {k: v for k, v in asdict(spec).items() if k not in ("shape", "averageSteps")}
    row.update(
        voxelAreaM2=voxel_area, blockage=blockage, steps=run.steps, msPerStep=run.msPerStep, cd=cd,
        cdVoxel=cd_voxel, cdRaw=cd_voxel if measured_shape else None,
        cdMaskell=maskell(cd_voxel, blockage) if measured_shape else None, errorPct=error_pct, fx=fx, fy=fy,
        fz=fz, relStdPct=rel_std, driftPct=drift,
        forceHistory=[[s.step, s.fx, s.fy, s.fz] for s in run.forces
                      if s.step % FORCE_SAMPLE_STEPS == 0 and s.step <= run.steps],
        adapter=run.adapter if run.adapter is not None else context.adapter,
        userAgent=run.userAgent if run.userAgent is not None else context.userAgent,
        commit=context.commit, date=context.date)
    row.update({"pass": passed, "reasons": reasons})
    order = ["case", "id", "group", "label", "tier", "grid", "cells", "operator", "walls", "storage", "variant",
             "sensitivity", "shapeClass", "speedMps", "re", "les", "floor", "stilts", "bodyLengthMm",
             "bodyLengthCells", "cellSizeM", "referenceAreaM2", "voxelAreaM2", "blockage", "baselineBlockage", "uLu",
             "steps", "flowThroughSteps", "msPerStep", "cd", "cdVoxel", "cdRaw", "cdMaskell", "referenceCd",
             "errorPct", "fx", "fy", "fz", "relStdPct", "driftPct", "forceHistory", "adapter", "userAgent", "commit",
             "date", "isDefault", "pass", "reasons"]
    return {k: row[k] for k in order}


def case_scales(spec: CaseSpec) -> Tuple[float, float, float]:
    """caseScales (case-run.ts:97-112) on the parameters caseHash (links.ts:14-35) passes:
    (Re, speed in m/s, lattice viscosity). Only the sphere cases pass their Reynolds number;
    the others take it from the speed and the body length in millimetres (six decimals)."""
    length_m = float(f"{spec.bodyLengthMm:.6f}") / 1000
    passed_re = spec.re if spec.shape == "sphere" else None
    re = passed_re if passed_re is not None else spec.speedMps * length_m / NU_AIR
    speed = spec.speedMps if passed_re is None else re * NU_AIR / length_m
    return re, speed, spec.uLu * spec.bodyLengthCells / re


def case_meshes(spec: CaseSpec):
    """voxelBody's placement (case-run.ts:121-147): (body, supports) of the case's preset
    placed at bodyLengthCells with the case's floor. The Ahmed body's stilts are support
    cells, left out of the drag, unless the stilts sensitivity row counts them (case-run.ts:93,
    links.ts:32): then they are joined to the body."""
    from . import bodies
    mesh = bodies.preset_mesh(spec.shape)
    if mesh is None:
        raise ValueError(f"The preset {spec.shape} has no built-in geometry.")
    layout = (spec.grid["nx"], spec.grid["ny"], spec.grid["nz"])
    body, supports = bodies.place_mesh(mesh, bodies.preset_supports(spec.shape), spec.bodyLengthCells, layout,
                                       spec.floor)
    if spec.sensitivity == "stilts" and spec.stilts and supports is not None:
        return np.concatenate([body, supports]), None
    return body, supports


def case_grid(spec: CaseSpec):
    """The case's grid from the device voxelizer (link fractions for interpolated walls)."""
    from .voxelize import voxelize
    body, supports = case_meshes(spec)
    layout = (spec.grid["nx"], spec.grid["ny"], spec.grid["nz"])
    grid = voxelize(body, supports, layout, spec.floor, link_fractions=spec.walls == "bouzidi")
    if grid.solid_cells == 0:
        raise ValueError("The body covers no grid cells at this size.")
    return grid


def run_case(spec: CaseSpec, steps: Optional[int] = None) -> SolverRun:
    """One case on this solver, as runCase does it (case-run.ts:200-250): the case's grid,
    tau0 = 3 nu + 1/2 from its own Reynolds number (no interactive floor), its operator, walls,
    storage and LES switch, a force sample every 8 steps converted to newtons with
    rho_air (speed / u_lu)^2 dx^2, and the measured frontal area and blockage. A non-finite
    force ends the run with an error, as the reference's does."""
    from .gpu_solver import GpuSolver, SolverParams
    layout = (spec.grid["nx"], spec.grid["ny"], spec.grid["nz"])
    grid = case_grid(spec)
    _, speed, viscosity = case_scales(spec)
    params = SolverParams(3 * viscosity + 0.5, spec.uLu, spec.les, operator=spec.operator, walls=spec.walls,
                          precision=spec.storage)
    solver = GpuSolver.create(layout, grid, params)
    if not isinstance(solver, GpuSolver):
        raise MemoryError(solver.message)
    total = spec.steps if steps is None else steps
    newtons = RHO_AIR * (speed / spec.uLu) ** 2 * spec.cellSizeM ** 2
    try:
        started = time.perf_counter()
        samples = solver.step_sampled(total, FORCE_SAMPLE_STEPS)
        elapsed = time.perf_counter() - started
    finally:
        solver.destroy()
    forces = []
    for step, fx, fy, fz in samples:
        if not all(math.isfinite(v) for v in (fx, fy, fz)):
            raise ValueError(f"The force was not finite at step {int(step)} of {total}.")
        forces.append(ForceSample(int(step), fx * newtons, fy * newtons, fz * newtons))
    return SolverRun(forces, total, 1000 * elapsed / total, voxelAreaM2=grid.frontal_cells * spec.cellSizeM ** 2,
                     blockage=grid.frontal_cells / (layout[1] * layout[2]))


def run_laminar_sphere(spec: CaseSpec, steps: Optional[int] = None) -> SolverRun:
    """The laminar sphere cases (the 24 x 16 lat-long preset, D = bodyLengthCells, Re 100)."""
    if spec.case not in ("sphere-re-100", "sphere-confinement"):
        raise ValueError("run_laminar_sphere runs the laminar sphere cases.")
    return run_case(spec, steps)


# --------------------------------------------------------------------------------------------
# Body drag per flow-through (scripts/drag-cpu.ts on this solver)

DRAG_LAYOUT = (64, 32, 32)  # drag-cpu.ts:11
DRAG_BODY_LENGTH_CELLS = 24  # drag-cpu.ts:12
DRAG_FLOW_THROUGHS = 4  # drag-cpu.ts:13


def drag_flow_throughs(samples, steps_per_flow_through: int, flow_throughs: int, cd_per_force: float) -> List[dict]:
    """The per-flow-through binning of runBodyDrag (cases.ts:293-321): a sample at step k goes
    to flow-through floor(k / steps_per_flow_through) (the last step's sample falls past the
    end and is dropped); each bin reports its last step, sample count, mean Cd and relative
    standard deviation (n - 1, in percent; None below two samples)."""
    bins: List[List[float]] = [[] for _ in range(flow_throughs)]
    last = [0] * flow_throughs
    for step, fx in samples:
        k = int(step) // steps_per_flow_through
        if k < flow_throughs:
            bins[k].append(fx * cd_per_force)
            last[k] = int(step)
    out = []
    for k, values in enumerate(bins):
        mean = _mean(values) if values else None
        rel = None
        if mean is not None and len(values) >= 2:
            rel = 100 * math.sqrt(sum((v - mean) ** 2 for v in values) / (len(values) - 1)) / abs(mean)
        out.append({"flowThrough": k + 1, "step": last[k], "samples": len(values), "meanCd": mean, "relStdPct": rel})
    return out


def run_body_drag(operator: str = "trt", layout=DRAG_LAYOUT, body_length_cells: float = DRAG_BODY_LENGTH_CELLS,
                  flow_throughs: int = DRAG_FLOW_THROUGHS, les: bool = True, floor: bool = True,
                  precision: str = "f32") -> dict:
    """runBodyDrag (cases.ts:264-321) as scripts/drag-cpu.ts drives it (drag-cpu.ts:71-110), on
    this solver: the Ahmed preset with its stilts as supports, placed at `body_length_cells` on
    the floor, tau0 from the viscosity at 40 m/s (the interactive floor 1e-2 at this size), the
    uniform inflow state, a force sample every 8 steps binned per flow-through of
    ceil(nx / 0.05) steps, Cd on the voxel frontal cells."""
    from . import bodies
    from .gpu_solver import GpuSolver, SolverParams
    from .voxelize import voxelize
    nu = max(NU_MIN_LU, U_LU * body_length_cells / (DEFAULT_SPEED_MPS * (AHMED_BODY_LENGTH_MM / 1000) / NU_AIR))
    tau0 = nu * 3 + 0.5
    body, supports = bodies.place_mesh(bodies.ahmed_mesh(), bodies.stilts_mesh(), body_length_cells, layout, floor)
    grid = voxelize(body, supports, layout, floor, link_fractions=False)
    if grid.solid_cells == 0:
        raise ValueError("The body does not cover any grid cells.")
    per_flow = math.ceil(layout[0] / U_LU)
    cd_per_force = 1 / (0.5 * U_LU * U_LU * grid.frontal_cells)
    solver = GpuSolver.create(layout, grid, SolverParams(tau0, U_LU, les, operator=operator, precision=precision))
    if not isinstance(solver, GpuSolver):
        raise MemoryError(solver.message)
    try:
        samples = solver.step_sampled(flow_throughs * per_flow, FORCE_SAMPLE_STEPS)
    finally:
        solver.destroy()
    finite = bool(np.isfinite(samples[:, 1:]).all())
    return {"operator": operator, "tau0": tau0, "frontalCells": grid.frontal_cells,
            "blockagePct": 100 * grid.frontal_cells / (layout[1] * layout[2]), "stepsPerFlowThrough": per_flow,
            "finite": finite, "flowThroughs": drag_flow_throughs(samples[:, :2], per_flow, flow_throughs, cd_per_force)}


# --------------------------------------------------------------------------------------------
# Grid check (the app's "Check grid" button)

GRID_CHECK_TIERS = ("S", "M", "L")  # grid-check.ts:13
GRID_CHECK_FLOW_THROUGHS = 4  # grid-check.ts:15
NU_MIN_LU = 1e-2  # units.ts:7, the interactive viscosity floor
LENGTH_CAP_NX = 0.4  # transform.ts:17
BLOCKAGE_MAX = 0.08  # transform.ts:18
DEFAULT_SPEED_MPS = 40  # params.ts:4
# Preset body length and registry reference area (params.ts:18-19).
PRESET_SIZES = {"ahmed-25": (1044.0, 112_000.0), "sphere": (100.0, math.pi * 50 * 50)}


def default_length_cells(mesh, layout) -> float:
    """defaultLengthCells (transform.ts:80-91): the longest extent at 0.4 nx, or less so that
    the frontal box stays below 8 % blockage (rounded down to whole cells)."""
    from . import bodies
    nx, ny, nz = layout
    lo, hi = bodies.mesh_bounds(mesh)
    ext = hi - lo
    longest = float(max(ext))
    frontal = float(min(ext[0], ext[1]) * ext[2])
    blockage_length = math.inf if frontal == 0 else math.floor(longest * math.sqrt(BLOCKAGE_MAX * ny * nz / frontal))
    return min(LENGTH_CAP_NX * nx, blockage_length)


def run_length_cells(preset: str, mesh, layout) -> float:
    """runLengthCells (run-plan.ts:34-45): the Ahmed body at half the streamwise domain."""
    if preset == "ahmed-25":
        return max(1, drag.round_half_up(layout[0] * AHMED_BODY_LENGTH_FRACTION))
    return default_length_cells(mesh, layout)


def grid_check_length_cells(length_cells: float, nx: int, tier: str) -> int:
    """gridCheckLengthCells (grid-check.ts:65-71): the same body share of the domain on every grid."""
    if not (length_cells > 0 and nx > 0):
        raise ValueError("The run being checked needs a positive body length and grid length.")
    return max(1, drag.round_half_up(length_cells / nx * TIERS[tier].nx))


def grid_check_steps(tier: str, flow_throughs: float = GRID_CHECK_FLOW_THROUGHS) -> Tuple[int, int]:
    """gridCheckSteps (grid-check.ts:74-86): (total steps, newest steps averaged)."""
    if not (math.isfinite(flow_throughs) and flow_throughs > 0):
        raise ValueError("The run length must be a positive number of flow-throughs.")
    flow = drag.flow_through_steps(TIERS[tier].nx)
    return math.ceil(flow_throughs * flow), math.ceil(flow_throughs / 2 * flow)


def mean_of_last(history: Sequence[Tuple[int, float]], average_steps: int) -> Optional[float]:
    """meanOfLast (grid-check.ts:131-141): mean Cd of the points after the last step minus
    `average_steps`, or None when nothing finite is left."""
    if not history:
        return None
    last = history[-1][0]
    window = [cd for step, cd in history if step > last - average_steps]
    mean = _mean(window)
    return mean if math.isfinite(mean) else None


def run_grid_check(preset: str, current_tier: str = "S", floor: bool = True, storage: str = "f16",
                   flow_throughs: float = GRID_CHECK_FLOW_THROUGHS, log=None) -> dict:
    """runBodyGridCheck (case-run.ts:257-320) with runGridCheck (grid-check.ts:201-228) on this
    solver, at the app's defaults (40 m/s, floor on, REG + LES, halfway walls, 16-bit storage
    where the device has it): the run at `current_tier` fixes Re = u_lu L / nu with the
    interactive floor nu >= 1e-2; tiers S, M and L then run the body at the same share of the
    domain at that Re for `flow_throughs` flow-throughs each, Cd on the preset's registry area
    in cells, averaged over the newest half; the band is read off the three means."""
    from . import bodies
    from .gpu_solver import GpuSolver, SolverParams
    from .voxelize import voxelize
    length_mm, area_mm2 = PRESET_SIZES[preset]
    mesh, supports = bodies.preset_mesh(preset), bodies.preset_supports(preset)
    current = TIERS[current_tier]
    current_cells = run_length_cells(preset, mesh, (current.nx, current.ny, current.nz))
    nu_live = max(NU_MIN_LU, U_LU * current_cells / (DEFAULT_SPEED_MPS * (length_mm / 1000) / NU_AIR))
    re = U_LU * current_cells / nu_live
    rows = []
    for tier in GRID_CHECK_TIERS:
        t = TIERS[tier]
        layout = (t.nx, t.ny, t.nz)
        length = grid_check_length_cells(current_cells, current.nx, tier)
        body, placed = bodies.place_mesh(mesh, supports, length, layout, floor)
        grid = voxelize(body, placed, layout, floor, link_fractions=False)
        cell_mm = length_mm / length
        cd_per_force = 1 / (0.5 * U_LU * U_LU * (area_mm2 / (cell_mm * cell_mm)))
        params = SolverParams(3 * (U_LU * length / re) + 0.5, U_LU, True, operator=default_operator(True),
                              walls=DEFAULT_WALLS, precision=storage)
        solver = GpuSolver.create(layout, grid, params)
        if not isinstance(solver, GpuSolver):
            raise MemoryError(solver.message)
        total, average = grid_check_steps(tier, flow_throughs)
        try:
            samples = solver.step_sampled(total, FORCE_SAMPLE_STEPS)
        finally:
            solver.destroy()
        history = [(int(step), fx * cd_per_force) for step, fx, _, _ in samples]
        cd = mean_of_last(history, average)
        if log is not None:
            log(f"grid check {preset} tier {tier}: {length} cells, Re {re:g}, {total} steps, Cd {cd}")
        rows.append({"tier": tier, "lengthCells": length, "cd": cd})
    band = grid_band([{"h": 1 / r["lengthCells"], "cd": math.nan if r["cd"] is None else r["cd"]} for r in rows])
    return {"case": preset, **band, "re": re, "rows": rows}


def convergence_rows(rows: Sequence[dict], operators: Sequence[str] = OPERATORS) -> List[dict]:
    """convergenceRows (results.ts:208-260): h = 1 / nx over the default Ahmed rows at M, L, XL
    and the laminar sphere at S, M, L per operator; a series missing a tier is skipped."""
    series = [("ahmed-25", default_operator(True), ("M", "L", "XL"),
               lambda r: r["case"] == "ahmed-25" and r["group"] == "validation" and r["isDefault"])]
    for op in operators:
        series.append(("sphere-re-100", op, ("S", "M", "L"),
                       lambda r, op=op: r["case"] == "sphere-re-100" and r["group"] == "laminar"
                       and r["operator"] == op))
    out = []
    for case, op, tiers, match in series:
        samples = []
        for tier in tiers:
            row = next((r for r in rows if r["tier"] == tier and match(r)), None)
            if row is not None:
                samples.append({"h": 1 / row["grid"]["nx"], "value": row["cd"]})
        if len(samples) == len(tiers):
            out.append({**grid_convergence(samples), "case": case, "group": "convergence", "operator": op,
                        "tiers": list(tiers)})
    return out


# --------------------------------------------------------------------------------------------
# Scoring

def to_fixed(x: float, digits: int) -> str:
    """Number.prototype.toFixed: the exact binary value rounded half away from zero, a sign
    only for negative values (not for -0)."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if abs(x) >= 1e21:
        return js_number(x)
    if x == 0:
        x = 0.0
    q = Decimal(x).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP)
    return f"{q:f}"


def _read_entry(row, where: str):
    if not isinstance(row, dict):
        return f"{where} is not an object."
    if not isinstance(row.get("case"), str) or row["case"] == "":
        return f"{where} has no case name."
    cd = row.get("cd")
    if isinstance(cd, bool) or not isinstance(cd, (int, float)) or not math.isfinite(cd):
        return f"{where} ({row['case']}) has no finite cd."
    return {"case": row["case"], "cd": float(cd)}


def score_results(parsed) -> Tuple[List[str], List[str], int]:
    """score-results.mjs: (stdout lines, stderr lines, exit status) for a parsed document, either
    a foreign file {solver, rows} or a results.json (its isDefault rows of the laminar and
    validation sections)."""
    if not isinstance(parsed, dict):
        return [], ["score-results: the file must be a JSON object."], 1
    if isinstance(parsed.get("rows"), list):
        if not isinstance(parsed.get("solver"), str) or parsed["solver"] == "":
            return [], ["score-results: solver must name the solver and its version."], 1
        rows = parsed["rows"]
    elif "schemaVersion" in parsed:
        rows = []
        for name in SCORED_SECTIONS:
            section = parsed.get(name)
            if isinstance(section, list):
                rows += [r for r in section if isinstance(r, dict) and r.get("isDefault") is True]
    else:
        return [], ["score-results: the file is neither a results.json nor a foreign results file."], 1
    entries, errors = [], []
    for i, row in enumerate(rows):
        e = _read_entry(row, f"row {i}")
        (errors if isinstance(e, str) else entries).append(e)
    if errors:
        return [], [f"score-results: {e}" for e in errors], 1
    out, err, unmatched = [], [], 0
    for e in entries:
        if e["case"] not in REFERENCES:
            err.append(f"score-results: unknown case {e['case']}.")
            unmatched += 1
            continue
        ref = REFERENCES[e["case"]]
        error_pct = None if ref is None else (e["cd"] - ref) / ref * 100
        out.append(f"{e['case']}, {to_fixed(e['cd'], 4)}, {'-' if ref is None else js_number(ref)}, "
                   f"{'-' if error_pct is None else to_fixed(error_pct, 2)}")
    out.append(f"unmatched: {unmatched}")
    return out, err, 0


def foreign_results(solver: str, rows: Sequence[dict]) -> dict:
    """The foreign-solver results document (docs/validation.md "Results row format")."""
    return {"solver": solver, "rows": [{"case": r["case"], "tier": r["tier"], "cd": r["cd"],
                                        "relStdPct": r["relStdPct"], "steps": r["steps"]} for r in rows]}


def run_cases(specs: Sequence[CaseSpec], context: RunContext, log=print) -> List[dict]:
    """Run each case on this solver and turn it into a result row (the force history kept)."""
    rows = []
    for spec in specs:
        started = time.perf_counter()
        row = row_from_run(spec, run_case(spec), context)
        log(f"{spec.id}: Cd {row['cd']:.6f} (voxel {row['cdVoxel']:.6f}), relStd {row['relStdPct']:.3f} %, "
            f"drift {row['driftPct']:.3f} %, {row['msPerStep']:.3f} ms/step, "
            f"{time.perf_counter() - started:.1f} s")
        rows.append(row)
    return rows


def _score(path: str) -> int:
    try:
        with open(path, encoding="utf-8") as f:
            parsed = json.load(f)
    except (OSError, ValueError) as e:
        sys.stderr.write(f"score-results: cannot read {path}: {e}\n")
        return 1
    out, err, status = score_results(parsed)
    for line in err:
        sys.stderr.write(line + "\n")
    for line in out:
        sys.stdout.write(line + "\n")
    return status


def _run(argv: List[str]) -> int:
    import argparse
    import datetime
    ap = argparse.ArgumentParser(prog="python -m pocketwind_amd.validation run",
                                 description="Run validation cases on the device and write their rows.")
    ap.add_argument("out", help="output file: a foreign results document with the full rows under 'results'")
    ap.add_argument("ids", nargs="*", help="case ids (default: every default row of the chosen tiers)")
    ap.add_argument("--tiers", nargs="+", default=list(DEFAULT_TIERS), choices=list(TIERS))
    a = ap.parse_args(argv)
    cases = build_cases(a.tiers) + [smoke_case()]
    if a.ids:
        known = {c.id: c for c in cases}
        unknown = [i for i in a.ids if i not in known]
        if unknown:
            sys.stderr.write(f"Unknown case ids: {', '.join(unknown)}\n")
            return 2
        specs = [known[i] for i in a.ids]
    else:
        specs = [c for c in cases if c.isDefault]
    date = datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%S.000Z")
    rows = run_cases(specs, RunContext("pocketwind_amd", date), log=lambda m: print(m, flush=True))
    doc = foreign_results("pocketwind_amd", rows)
    doc["results"] = [{k: v for k, v in r.items() if k != "forceHistory"} for r in rows]
    with open(a.out, "w", encoding="utf-8") as f:
        json.dump(doc, f, indent=1)
    return 0


def main(argv: Optional[List[str]] = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) == 2 and argv[0] == "score":
        return _score(argv[1])
    if argv and argv[0] == "run":
        return _run(argv[1:])
    sys.stderr.write("Usage: python -m pocketwind_amd.validation score <results.json | foreign-results.json>\n"
                     "       python -m pocketwind_amd.validation run <out.json> [case ids] [--tiers S M L XL]\n")
    return 2


if __name__ == "__main__":
    sys.exit(main())
