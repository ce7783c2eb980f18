/*
 * Host-side checks of the Node binding that need no GPU: the addon loads
 * and exports the whole C ABI surface, the facade rejects what this build
 * does not have, and constructing a world without a GPU fails loudly (the
 * library has no CPU path).
 */
'use strict';
var path = require('path');
var root = path.join(__dirname, '..', '..', 'watersim-v1_amd');
var failures = 0;
function check(name, ok, detail) {
  console.log((ok ? '  PASS ' : '  FAIL ') + name + (detail !== undefined ? '  [' + detail + ']' : ''));
  if (!ok) failures++;
}
function throws(fn, re) {
  try { fn(); } catch (e) { return re.test(e.message) ? e.message : 'wrong message: ' + e.message; }
  return null;
}

var addon = require(path.join(root, '_build', 'watersim_napi.node'));
var want = ['create', 'destroy', 'setF', 'getF', 'upload', 'download', 'substep', 'stage', 'stepFrame',
  'runSubsteps', 'splatDensity', 'mtBuild', 'autotuneIso', 'kineticEnergy', 'sync', 'worldBuild', 'setBalls',
  'getBalls', 'metaballs', 'readRender', 'markFluid', 'sampleVel', 'planetGravity'];
check('addon exports the C ABI surface', want.every(function (k) { return typeof addon[k] === 'function'; }),
  want.filter(function (k) { return typeof addon[k] !== 'function'; }).join(','));

var FluidSolver = require(path.join(root, 'js', 'solver.js'));
// the reference defaults to the box pool (solver.js:78); this build has the
// planet only: constructing is allowed (callers set mode afterwards, as the
// reference's tests do), device work in pool mode is refused
var pool = new FluidSolver({ nx: 8, ny: 8, nz: 8, dx: 0.1 });
var m = throws(function () { pool.resetWater(0.3); }, /sphere/);
var m2 = throws(function () { pool.step(1 / 60); }, /sphere/);
check('box-pool mode is refused with a clear error', pool.mode === 'pool' && m && !/wrong/.test(m) && m2 &&
  !/wrong/.test(m2), m);
var s = new FluidSolver({ nx: 24, ny: 24, nz: 24, dx: 5.1 / 24, targetParticles: 4000, mode: 'sphere',
  coreR: 1.275 });
check('reference defaults', s.pic === 0.15 && s.omega === 1.55 && s.pressureIters === 20 && s.cfl === 3 &&
  s.gravity === 9.81 && s.maxSpeed === 3.2 && s.Tamb === 0.32 && Math.abs(s.meltT - 0.105) < 1e-12, 'pic ' + s.pic);
s.addBall(1, 1, 1, 0.3, 350);
var b = s.balls[0];
check('addBall keeps the reference ball object (solver.js:3910-3912)', s.balls.length === 1 && b.x === 1 &&
  b.vx === 0 && b.r === 0.3 && b.rho === 350 && b.submerged === 0);
s.balls.length = 0;
if (!process.env.WATERSIM_EXPECT_GPU) {
  // before resetWater a device call creates an empty world, which needs a GPU
  m = throws(function () { s.step(1 / 60); }, /ws_create|GPU/);
  check('no GPU: a device call before resetWater throws (no CPU fallback)', m && !/wrong/.test(m), m);
  m = throws(function () { s.resetWater(0.765); }, /ws_create|GPU/);
  check('no GPU: building the world throws (no CPU fallback)', m && !/wrong/.test(m), m);
}
var w = addon.worldBuild({ nx: 24, ny: 24, nz: 24, dx: 5.1 / 24, coreR: 1.275, depth: 0.765, bumpiness: 1,
  seed: 20260114, targetParticles: 4000 });
check('host world builder through the addon', w.nP > 2000 && w.nP <= 4000 && w.solid.length === w.tn * w.tn * w.tn &&
  w.gravity.length === 3 * 25 * 24 * 24, 'nP ' + w.nP + ', tn ' + w.tn + ', land ' + w.landFrac.toFixed(3));
console.log(failures ? failures + ' FAILED' : 'ALL BOUNDARY CHECKS PASSED');
process.exitCode = failures ? 1 : 0;
