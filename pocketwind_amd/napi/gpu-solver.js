'use strict';
/*
 * GpuSolver for a Node host (src/lbm/gpu-solver.ts:135-465): the reference's class, method for
 * method, over the Node-API addon pw_napi.node, which calls libpocketwind's C-ABI. A command
 * encoder is not part of this interface: step/reset/encodeForce are stream-ordered on the
 * solver's HIP stream, as the reference's are queue-ordered.
 *
 *   const { GpuSolver } = require('pocketwind_amd/napi/gpu-solver');
 *   const solver = await GpuSolver.create({nx, ny, nz}, grid, {tau0: 0.8, uIn: 0.06, les: false});
 *   solver.step(8, true); solver.encodeForce(); const f = await solver.readForce();
 */
const path = require('path');

const addon = require(path.join(__dirname, 'pw_napi.node'));

/** src/lbm/params.ts:12: steps between force samples. */
const FORCE_EVERY_STEPS = 8;

/** src/lbm/gpu-resources.ts:3-5: the one failure that permits a smaller allocation. */
class AllocationFailure extends Error {
  constructor(message) {
    super(message);
    this.kind = 'allocation-failure';
  }
}

/** src/lbm/gpu-solver.ts:80-116: the 48-byte uniform, packed by the library. */
function packSolverParams(layout, params, writeMacro = false, step = 0) {
  return addon.packParams(layout, params, writeMacro, step);
}

/** src/lbm/gpu-solver.ts:119-123. */
function rampFactor(step, rampSteps) {
  if (!Number.isFinite(rampSteps) || rampSteps <= 0) return 1;
  if (!Number.isFinite(step) || step <= 0) return 0;
  return Math.min(1, step / rampSteps);
}

class GpuSolver {
  constructor(handle, layout) {
    this._handle = handle;
    this._layout = { nx: layout.nx, ny: layout.ny, nz: layout.nz };
    this._destroyed = false;
  }

  /**
   * src/lbm/gpu-solver.ts:158-335. `grid` carries `geometry` (Uint32Array, the packed
   * [flags N][q offsets N][f32 q per link]) and `boundaryCells` (Uint32Array); `params` is
   * SolverParams {tau0, uIn, les, rampSteps?, walls?, operator?, precision?} plus this library's
   * wallModel, sides, outlet and slipFloor. Resolves to the solver, or to an AllocationFailure; invalid
   * input throws a RangeError.
   */
  static async create(layout, grid, params) {
    const boundary = grid.boundaryCells || new Uint32Array(0);
    const handle = addon.create(layout, grid.geometry, boundary, params);
    if (handle && handle.kind === 'allocation-failure') return new AllocationFailure(handle.message);
    return new GpuSolver(handle, layout);
  }

  /** src/lbm/gpu-solver.ts:342-344. */
  get stepCount() {
    return addon.stepCount(this._handle);
  }

  get layout() {
    return this._layout;
  }

  /** src/lbm/gpu-solver.ts:347-355. */
  reset() {
    addon.reset(this._handle);
  }

  /** src/lbm/gpu-solver.ts:378-395. */
  step(count, writeMacroOnLast) {
    addon.step(this._handle, count, Boolean(writeMacroOnLast));
  }

  /** src/lbm/gpu-solver.ts:448-451. */
  encodeForce() {
    addon.encodeForce(this._handle);
  }

  /** src/lbm/gpu-solver.ts:454-456: {step, fx, fy, fz} or 'busy'. */
  async readForce() {
    return addon.readForce(this._handle);
  }

  /** The latest state's 19 planes in storage format (populationBuffers readback). */
  readPopulations() {
    return addon.readPopulations(this._handle);
  }

  /** The test-side writeBuffer into populationBuffers (tests/browser/force.spec.ts:145-153). */
  writePopulations(data) {
    addon.writePopulations(this._handle, data);
  }

  /** macroBuffer readback: Float32Array of (rho, ux, uy, uz) per cell. */
  readMacro() {
    return addon.readMacro(this._handle);
  }

  /** src/lbm/gpu-solver.ts:459-464: idempotent. */
  destroy() {
    if (this._destroyed) return;
    this._destroyed = true;
    addon.destroy(this._handle);
  }
}

module.exports = {
  AllocationFailure,
  FORCE_EVERY_STEPS,
  GpuSolver,
  abiVersion: addon.abiVersion,
  deviceCount: addon.deviceCount,
  packSolverParams,
  rampFactor,
};
