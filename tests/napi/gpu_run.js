'use strict';
// One solver run through the Node-API boundary, driven by tests/test_napi.py: reads
// <dir>/case.json ({layout, params, blocks}), <dir>/geometry.u32 and <dir>/boundary.u32, steps
// `blocks` times 8 steps with a force sample after each (gpu-solver.ts:378-456 call order), and
// writes <dir>/populations.bin, <dir>/macro.f32 and <dir>/result.json.
const assert = require('assert');
const fs = require('fs');
const path = require('path');

const pw = require(path.join(__dirname, '..', '..', 'pocketwind_amd', 'napi', 'gpu-solver.js'));

function u32(file) {
  const b = fs.readFileSync(file);
  return new Uint32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}

(async () => {
  const dir = process.argv[2];
  const spec = JSON.parse(fs.readFileSync(path.join(dir, 'case.json'), 'utf8'));
  const grid = { geometry: u32(path.join(dir, 'geometry.u32')), boundaryCells: u32(path.join(dir, 'boundary.u32')) };
  console.log('napi case', JSON.stringify(spec));
  const solver = await pw.GpuSolver.create(spec.layout, grid, spec.params);
  assert.ok(solver instanceof pw.GpuSolver, `create returned ${solver && solver.kind}`);
  const result = { busyBeforeEncode: await solver.readForce(), samples: [] };
  for (let k = 0; k < spec.blocks; ++k) {
    solver.step(pw.FORCE_EVERY_STEPS, k === spec.blocks - 1);
    solver.encodeForce();
    result.samples.push(await solver.readForce());
  }
  result.stepCount = solver.stepCount;
  fs.writeFileSync(path.join(dir, 'populations.bin'), Buffer.from(solver.readPopulations()));
  const macro = solver.readMacro();
  fs.writeFileSync(path.join(dir, 'macro.f32'), Buffer.from(macro.buffer, macro.byteOffset, macro.byteLength));
  solver.reset();
  result.stepCountAfterReset = solver.stepCount;
  solver.destroy();
  solver.destroy();
  try {
    solver.step(1, false);
    result.stepAfterDestroy = 'no error';
  } catch (e) {
    result.stepAfterDestroy = e.message;
  }
  fs.writeFileSync(path.join(dir, 'result.json'), JSON.stringify(result));
  console.log('napi done', result.stepCount);
})().catch((e) => {
  console.error(e && e.stack);
  process.exit(1);
});
