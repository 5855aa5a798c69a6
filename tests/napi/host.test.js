'use strict';
// Host-only checks of the Node-API boundary (no device call): the library's ABI, the device
// count on a machine without a GPU, and packSolverParams against tests/unit/gpu-solver.test.ts:9-62,
// including the layouts JS can express and a u32 layout cannot (-1, 1.5, NaN, Infinity).
const assert = require('assert');
const path = require('path');

const pw = require(path.join(__dirname, '..', '..', 'pocketwind_amd', 'napi', 'gpu-solver.js'));

const fround = Math.fround;
const layout = { nx: 32, ny: 16, nz: 8 };
const base = { tau0: 0.8, uIn: 0.04, les: false };

function words(buffer) {
  return { u: new Uint32Array(buffer), f: new Float32Array(buffer) };
}

const checks = [];
function check(name, fn) {
  checks.push([name, fn]);
}

check('abi version 8', () => assert.strictEqual(pw.abiVersion(), 8));

check('device count is a non-negative integer', () => {
  const n = pw.deviceCount();
  assert.ok(Number.isInteger(n) && n >= 0);
  if (process.env.PW_EXPECT_DEVICES !== undefined) assert.strictEqual(n, Number(process.env.PW_EXPECT_DEVICES));
});

check('packs the twelve scalars', () => {
  const { u, f } = words(pw.packSolverParams(layout, base, true));
  assert.strictEqual(u.byteLength, 48);
  assert.deepStrictEqual(Array.from(u.slice(0, 4)), [32, 16, 8, 4096]);
  assert.strictEqual(f[4], fround(0.04));
  assert.strictEqual(f[5], fround(0.1));
  assert.strictEqual(f[6], fround(18 * Math.SQRT2 * 0.01));
  assert.strictEqual(f[7], 3 / 16);
  assert.deepStrictEqual(Array.from(u.slice(8)), [0, 0, 1, 0]);
});

check('les and macro output are independent', () => {
  const { u } = words(pw.packSolverParams(layout, { ...base, les: true }, false));
  assert.deepStrictEqual(Array.from(u.slice(8)), [1, 0, 0, 0]);
});

check('walls pack as their mode', () => {
  assert.strictEqual(words(pw.packSolverParams(layout, { ...base, walls: 'bb' })).u[9], 0);
  assert.strictEqual(words(pw.packSolverParams(layout, { ...base, walls: 'bouzidi' })).u[9], 1);
});

check('step packs into the last word', () => {
  assert.strictEqual(words(pw.packSolverParams(layout, base, false, 12345)).u[11], 12345);
});

check('rejects invalid grid dimensions', () => {
  for (const nx of [0, -1, 1.5, NaN, Infinity]) {
    assert.throws(() => pw.packSolverParams({ ...layout, nx }, base), RangeError, `nx ${nx}`);
  }
});

check('rejects invalid relaxation times', () => {
  for (const tau0 of [0.5, 0, NaN, Infinity, 1e100, 0.5 + Number.EPSILON, 0.5 + 2 ** -26]) {
    assert.throws(() => pw.packSolverParams(layout, { ...base, tau0 }), /Relaxation time/, `tau0 ${tau0}`);
    assert.throws(() => pw.packSolverParams(layout, { ...base, tau0 }), RangeError);
  }
});

check('rejects speeds that overflow f32', () => {
  for (const uIn of [1e100, -1e100, Infinity, -Infinity, NaN]) {
    assert.throws(() => pw.packSolverParams(layout, { ...base, uIn }), RangeError, `uIn ${uIn}`);
  }
});

check('accepts the first f32 relaxation time above one half', () => {
  const { f } = words(pw.packSolverParams(layout, { ...base, tau0: 0.5 + 2 ** -24 }));
  assert.ok(f[5] > 0);
  assert.strictEqual(f[5], fround(fround(fround(0.5 + 2 ** -24) - 0.5) / 3));
});

check('rejects unknown option strings', () => {
  assert.throws(() => pw.packSolverParams(layout, { ...base, walls: 'mirror' }), RangeError);
  assert.throws(() => pw.packSolverParams(layout, { ...base, precision: 'f8' }), RangeError);
});

check('create validates before touching a device', async () => {
  const cells = 4 * 4 * 4;
  const geometry = new Uint32Array(2 * cells);
  geometry.fill(0xffffffff, cells);
  await assert.rejects(pw.GpuSolver.create({ nx: 4, ny: 4, nz: 4 }, { geometry }, { ...base, tau0: 0.5 }), RangeError);
  await assert.rejects(
    pw.GpuSolver.create({ nx: 4, ny: 4, nz: 4 }, { geometry: geometry.subarray(1) }, base), RangeError);
  await assert.rejects(
    pw.GpuSolver.create({ nx: 4, ny: 4, nz: 4 }, { geometry: new Float32Array(2 * cells) }, base), RangeError);
});

check('rampFactor follows the reference', () => {
  assert.strictEqual(pw.rampFactor(5, 0), 1);
  assert.strictEqual(pw.rampFactor(0, 10), 0);
  assert.strictEqual(pw.rampFactor(5, 10), 0.5);
  assert.strictEqual(pw.rampFactor(50, 10), 1);
});

(async () => {
  let failed = 0;
  for (const [name, fn] of checks) {
    try {
      await fn();
      console.log(`ok - ${name}`);
    } catch (e) {
      failed += 1;
      console.log(`not ok - ${name}: ${e && e.stack}`);
    }
  }
  console.log(`${checks.length - failed}/${checks.length} passed`);
  process.exit(failed ? 1 : 0);
})();
