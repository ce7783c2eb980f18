'use strict';
/*
 * Times the reference's own CPU substep (FluidSolver.prototype._substep,
 * js/solver.js:2568-2624) in Node, single-threaded, the way
 * test/performance.js times step(): warm-up first, then the median of the
 * timed calls, measured with process.hrtime.bigint() around each call.
 *
 *   node tests/golden/time_reference.js <reference> cfg1
 *   node tests/golden/time_reference.js <reference> cfg2 <velocities.bin>
 *
 * cfg1: 32^3, 10k target particles, seed 20260114, FLIP 0.95, omega 1.7,
 *       40 lexicographic SOR sweeps, from the fill at rest, dt from the CFL
 *       law of _stepSync (solver.js:2689-2690).
 * cfg2: 64^3, 232k target particles, seed 20260115, vorticity 0.05, the
 *       curl-noise start velocities read from the file (float32 x, y, z
 *       components, nP each), and SOR to relative residual 1e-4: the
 *       reference has a fixed sweep count, so _pressureSolve is called with
 *       pressureIters = 8 until the residual (FP64, computed here from the
 *       solver's own rhs, diag, q and cell types) is at most 1e-4; every
 *       call re-assembles the same system and continues the sweeps from q.
 *
 * Prints one JSON line: the config, particle count, median / mean ms per
 * substep, the pressure share, and the sweeps per substep.
 * Run only in the development container, where the reference checkout is.
 */
var path = require('path');
var fs = require('fs');

var REF = process.argv[2] || '/root/reference';
var CFG = process.argv[3] || 'cfg1';
var VEL = process.argv[4] || '';
global.MarchingTetrahedra = require(path.join(REF, 'js/surface.js'));
var FluidSolver = require(path.join(REF, 'js/solver.js'));

var W = 5.1;
var spec = CFG === 'cfg2'
  ? { n: 64, target: 232000, seed: 20260115, vort: 0.05, tol: 1e-4, warm: 30, timed: 90 }
  : { n: 32, target: 10000, seed: 20260114, vort: 0, tol: 0, warm: 30, timed: 90 };
var n = spec.n, dx = W / n;
var s = new FluidSolver({ nx: n, ny: n, nz: n, dx: dx, targetParticles: spec.target, mode: 'sphere',
  coreR: 0.25 * W, terrainSeed: spec.seed });
s.resetWater(0.4 * W - 0.25 * W);
s.pic = 0.05; s.omega = 1.7; s.gravity = 9.8; s.maxSpeed = 3.2; s.vorticity = spec.vort;
s.pressureIters = 40;
if (VEL) {
  var buf = fs.readFileSync(VEL);
  var v = new Float32Array(buf.buffer, buf.byteOffset, buf.length / 4);
  if (v.length !== 3 * s.nP) throw new Error('velocity file holds ' + v.length / 3 + ' particles, the solver ' + s.nP);
  for (var p = 0; p < s.nP; p++) { s.pvx[p] = v[p]; s.pvy[p] = v[s.nP + p]; s.pvz[p] = v[2 * s.nP + p]; }
}

function dtLaw() {
  var dt = Math.min((s.cfl * s.dx) / Math.max(s.umax, 0.08), 1 / 100);
  return dt < 1e-5 ? 1e-5 : dt;
}

// ||rhs - A q|| / ||rhs|| over the fluid cells (solid neighbours Neumann)
function residual() {
  var t = s.cellType, q = s.q, rhs = s.rhs, diag = s.diag, nxy = n * n, rr = 0, bb = 0;
  for (var c = 0; c < n * n * n; c++) {
    if (t[c] !== 1) continue;
    var i = c % n, j = ((c / n) | 0) % n, k = (c / nxy) | 0, sum = 0;
    if (i > 0 && t[c - 1] === 1) sum += q[c - 1];
    if (i < n - 1 && t[c + 1] === 1) sum += q[c + 1];
    if (j > 0 && t[c - n] === 1) sum += q[c - n];
    if (j < n - 1 && t[c + n] === 1) sum += q[c + n];
    if (k > 0 && t[c - nxy] === 1) sum += q[c - nxy];
    if (k < n - 1 && t[c + nxy] === 1) sum += q[c + nxy];
    var r = rhs[c] - (diag[c] * q[c] - sum);
    rr += r * r;
    bb += rhs[c] * rhs[c];
  }
  return bb > 0 ? Math.sqrt(rr) / Math.sqrt(bb) : 0;
}

var sweeps = 0, pressNs = 0n;
var orig = FluidSolver.prototype._pressureSolve;
s._pressureSolve = function () {
  var t0 = process.hrtime.bigint();
  if (spec.tol > 0) {
    this.pressureIters = 0;
    orig.call(this);
    var res = residual(), done = 0;
    while (res > spec.tol && done < 4000) {
      this.pressureIters = 8;
      orig.call(this);
      done += 8;
      res = residual();
    }
    sweeps += done;
  } else {
    orig.call(this);
    sweeps += this.pressureIters;
  }
  pressNs += process.hrtime.bigint() - t0;
};

for (var i = 0; i < spec.warm; i++) s._substep(dtLaw());
sweeps = 0; pressNs = 0n;
var times = [];
for (i = 0; i < spec.timed; i++) {
  var dt = dtLaw();
  var t0 = process.hrtime.bigint();
  s._substep(dt);
  times.push(Number(process.hrtime.bigint() - t0) / 1e6);
}
var mean = times.reduce(function (a, b) { return a + b; }, 0) / times.length;
var sorted = times.slice().sort(function (a, b) { return a - b; });
console.log(JSON.stringify({
  config: CFG, n: n, particles: s.nP, warmup: spec.warm, timed: spec.timed,
  median_ms: sorted[sorted.length >> 1], mean_ms: mean,
  pressure_ms_mean: Number(pressNs) / 1e6 / spec.timed, sweeps_per_substep: sweeps / spec.timed,
  ke_end: s.kineticEnergy(), node: process.version
}));
