'use strict';
/*
 * Golden-vector generator: runs the reference FluidSolver (js/solver.js,
 * js/surface.js of johnnybuggy/watersim-v1) in Node on small seeded scenes and
 * writes per-stage input/output dumps of one substep, plus trajectory
 * summaries and a marching-tetrahedra result, into tests/golden/*.wsg.gz.
 *
 * Run only where the reference checkout exists:
 *     node tests/golden/gen_golden.js /root/reference
 * The fixtures it writes are data (inputs and expected outputs); the
 * reference itself never travels with them.
 *
 * File format (.wsg.gz, gzip of): "WSG1", u32 LE header length, JSON header
 * {meta, arrays:[{name,dtype,shape,offset,nbytes}]}, then the raw arrays
 * (little endian) back to back.
 */
var path = require('path');
var fs = require('fs');
var zlib = require('zlib');

var REF = process.argv[2] || '/root/reference';
var ONLY = process.argv[3] || '';   // optional: write only this scene
var OUT = path.dirname(__filename);

global.MarchingTetrahedra = require(path.join(REF, 'js/surface.js'));
var FluidSolver = require(path.join(REF, 'js/solver.js'));

function Writer() { this.arrays = []; this.meta = {}; }
Writer.prototype.add = function (name, arr) {
  var dt;
  if (arr instanceof Float32Array) dt = 'f4';
  else if (arr instanceof Float64Array) dt = 'f8';
  else if (arr instanceof Uint8Array) dt = 'u1';
  else if (arr instanceof Int32Array) dt = 'i4';
  else if (arr instanceof Uint32Array) dt = 'u4';
  else throw new Error('dtype ' + name);
  this.arrays.push({ name: name, dtype: dt, data: Buffer.from(arr.buffer.slice(arr.byteOffset, arr.byteOffset + arr.byteLength)) });
};
Writer.prototype.save = function (file) {
  var off = 0, hdr = { meta: this.meta, arrays: [] };
  this.arrays.forEach(function (a) {
    hdr.arrays.push({ name: a.name, dtype: a.dtype, offset: off, nbytes: a.data.length });
    off += a.data.length;
  });
  var hj = Buffer.from(JSON.stringify(hdr));
  var head = Buffer.alloc(8);
  head.write('WSG1', 0, 'ascii');
  head.writeUInt32LE(hj.length, 4);
  var all = Buffer.concat([head, hj].concat(this.arrays.map(function (a) { return a.data; })));
  fs.writeFileSync(path.join(OUT, file), zlib.gzipSync(all, { level: 9 }));
  console.log(file, (all.length / 1e6).toFixed(2), 'MB raw');
};

function copy(a, n) { var c = new a.constructor(n === undefined ? a.length : n); c.set(n === undefined ? a : a.subarray(0, n)); return c; }

var PART_F = ['px', 'py', 'pz', 'pvx', 'pvy', 'pvz', 'pT', 'pAir', 'pLight', 'pWx', 'pWy', 'pWz', 'pPh'];
var PART_U8 = ['pflag', 'pcool', 'pDepth'];

function dumpWorld(w, s) {
  var t = s.terrain;
  w.add('terrain_solid', copy(t.solid));
  w.add('terrain_R', copy(t.R));
  w.add('static_types', copy(s._staticTypes));
  w.add('gu', copy(s._gu)); w.add('gv', copy(s._gv)); w.add('gw', copy(s._gw));
  w.meta.world = {
    nx: s.nx, ny: s.ny, nz: s.nz, dx: s.dx, W: s.W, cx: s.cx, cy: s.cy, cz: s.cz,
    coreR: s.coreR, oceanR: s.oceanR, domainR: s.domainR, spacing: s.spacing,
    particleVolume: s.particleVolume, tn: t.n, dv: t.dv, Rsl: t.Rsl, Rlo: t.Rlo, Rhi: t.Rhi,
    landFrac: t.landFrac, atmosphereH: s.atmosphereH
  };
}

function dumpParticles(w, s, tag) {
  var n = s.nP;
  PART_F.forEach(function (k) { w.add(tag + k, copy(s[k], n)); });
  PART_U8.forEach(function (k) { w.add(tag + k, copy(s[k], n)); });
}

function dumpGrid(w, s, tag, names) {
  names.forEach(function (k) { w.add(tag + k, copy(s[k])); });
}

function params(s) {
  return { pic: s.pic, omega: s.omega, pressureIters: s.pressureIters, gravity: s.gravity,
    maxSpeed: s.maxSpeed, viscosity: s.viscosity, vorticity: s.vorticity, cfl: s.cfl,
    gRamp: s._gRamp, nP: s.nP, iso: s.iso };
}

function dtLaw(s) {
  var dt = Math.min((s.cfl * s.dx) / Math.max(s.umax, 0.08), 1 / 100);
  return dt < 1e-5 ? 1e-5 : dt;
}

// One substep, the stage order of FluidSolver.prototype._substep (sphere mode),
// dumping what each stage writes.
function stagedSubstep(w, s, dt) {
  var G = ['u', 'v', 'w'];
  w.meta.stage_dt = dt;
  w.meta.stage_params = params(s);
  dumpParticles(w, s, 'in_');
  w.add('in_q', copy(s.q));
  var nb = s.balls.length;
  if (nb) w.add('in_balls', ballState(s));
  s.u.fill(0); s.v.fill(0); s.w.fill(0);
  s.uW.fill(0); s.vW.fill(0); s.wW.fill(0);
  s._rasterizeStatic();
  if (nb) dumpGrid(w, s, 'rast_', ['cellType', 'svx', 'svy', 'svz']);
  s._markFluidCells(0, s.nP);
  w.add('mark_cellType', copy(s.cellType));
  w.add('mark_pflag', copy(s.pflag, s.nP));
  s._pushOutOfBalls();
  if (nb) dumpParticles(w, s, 'pushb_');
  s._p2g();
  dumpGrid(w, s, 'p2g_', ['u', 'v', 'w', 'uW', 'vW', 'wW']);
  s._applyBC();
  dumpGrid(w, s, 'bc_', G);
  s.uO.set(s.u); s.vO.set(s.v); s.wO.set(s.w);
  s._viscosity(dt);
  dumpGrid(w, s, 'visc_', G);
  w.meta.gRamp_before_gravity = s._gRamp;
  s._applyPlanetGravity(dt);
  s._applyBC();
  dumpGrid(w, s, 'grav_', G);
  s._pressureSolve();
  w.add('pres_q', copy(s.q)); w.add('pres_rhs', copy(s.rhs)); w.add('pres_diag', copy(s.diag));
  w.meta.nFluid = s.nFluid;
  s._project();
  dumpGrid(w, s, 'proj_', G);
  if (s.vorticity > 0) s._vorticityConfinement(dt);
  dumpGrid(w, s, 'vort_', G);
  s.umax = s._clampUmSlice(0, s.uN + s.vN + s.wN);
  w.meta.umax = s.umax;
  dumpGrid(w, s, 'clamp_', G);
  s._extrapolate();
  dumpGrid(w, s, 'ext_', ['u', 'v', 'w', 'validU', 'validV', 'validW']);
  s._g2p(dt);
  w.meta.airborneCount = s.airborneCount;
  dumpParticles(w, s, 'g2p_');
  w.add('g2p_advU', copy(s._advU, s.nP)); w.add('g2p_advV', copy(s._advV, s.nP)); w.add('g2p_advW', copy(s._advW, s.nP));
  s._advect(dt);
  s._gridSampleValid = false;
  s._updateBalls(dt);
  s.dtLast = dt;
  dumpParticles(w, s, 'adv_');
  if (nb) {
    w.add('out_balls', ballState(s));
    w.add('ball_hist', copy(s._hist)); w.add('ball_surf', copy(s._surfY));
  }
  w.meta.rng_after = { evS: s._evS, cfS: s._cfS, bfS: s._bfS };
}

function ke(s) { return s.kineticEnergy(); }

// ball objects as 9 doubles each: x y z vx vy vz r rho submerged
function ballState(s) {
  var a = new Float64Array(9 * s.balls.length);
  s.balls.forEach(function (b, i) {
    a.set([b.x, b.y, b.z, b.vx, b.vy, b.vz, b.r, b.rho, b.submerged], 9 * i);
  });
  return a;
}

function trajectory(s, nSub, rec) {
  for (var i = 0; i < nSub; i++) {
    var dt = dtLaw(s);
    s._substep(dt);
    rec.push({ dt: dt, umax: s.umax, ke: ke(s), airborne: s.airborneCount });
  }
}

function seededVel(s, seed, amp) {
  var st = seed >>> 0;
  function r() { st = (Math.imul(st, 1664525) + 1013904223) >>> 0; return st / 4294967296; }
  for (var p = 0; p < s.nP; p++) {
    s.pvx[p] = amp * (r() - 0.5); s.pvy[p] = amp * (r() - 0.5); s.pvz[p] = amp * (r() - 0.5);
  }
}

// ---------------------------------------------------------------- scene A
// Plain method, small: 32^3, 10k particles, FLIP 0.95, omega 1.7, 40 sweeps.
(function sceneA() {
  if (ONLY && ONLY !== 'cfg1_32') return;
  var n = 32, W = 5.1, dx = W / n;
  var s = new FluidSolver({ nx: n, ny: n, nz: n, dx: dx, targetParticles: 10000, mode: 'sphere',
    coreR: 0.25 * W, terrainSeed: 20260114 });
  s.resetWater(0.4 * W - 0.25 * W);
  s.pic = 0.05; s.omega = 1.7; s.pressureIters = 40; s.gravity = 9.8; s.maxSpeed = 3.2; s.vorticity = 0;
  var w = new Writer();
  w.meta.scene = 'cfg1_32';
  w.meta.fill = params(s);
  s._rasterizeStatic();
  dumpWorld(w, s);
  dumpParticles(w, s, 'fill_');
  w.add('fill_dens', copy(s.dens));
  var rec = [];
  trajectory(s, 30, rec);
  stagedSubstep(w, s, dtLaw(s));
  rec.push({ dt: s.dtLast, umax: s.umax, ke: ke(s), airborne: s.airborneCount });
  // surface on the post-substep state
  s.splatDensity();
  w.add('surf_dens', copy(s.dens));
  var m = global.MarchingTetrahedra.build(s.dens, n, n, n, dx, s.iso);
  w.meta.surf = { iso: s.iso, count: m.count, volume: m.volume };
  w.add('surf_pos', copy(m.pos, m.count * 3)); w.add('surf_nrm', copy(m.nrm, m.count * 3));
  w.meta.surf.autotune_iso = s.autotuneIso();
  w.meta.surf.Vt = s.nP * s.particleVolume;
  trajectory(s, 100 - 31, rec);
  w.meta.trajectory = rec;
  dumpParticles(w, s, 'end_');
  w.save('cfg1_32.wsg.gz');
})();

// ---------------------------------------------------------------- scene B
// Micro scene: 16^3, 2k particles, random velocities, full gravity,
// vorticity confinement on (exercises viscosity, confinement, spray paths).
(function sceneB() {
  if (ONLY && ONLY !== 'micro_16') return;
  var n = 16, W = 5.1, dx = W / n;
  var s = new FluidSolver({ nx: n, ny: n, nz: n, dx: dx, targetParticles: 2000, mode: 'sphere',
    coreR: 0.25 * W, terrainSeed: 20260115, vorticity: 0.05 });
  s.resetWater(0.4 * W - 0.25 * W);
  s.pic = 0.05; s.omega = 1.7; s.pressureIters = 40; s.gravity = 9.8; s.maxSpeed = 3.2;
  s._gRamp = 1.5;
  seededVel(s, 99, 2.0);
  var w = new Writer();
  w.meta.scene = 'micro_16';
  w.meta.fill = params(s);
  s._rasterizeStatic();
  dumpWorld(w, s);
  dumpParticles(w, s, 'fill_');
  var rec = [];
  trajectory(s, 5, rec);
  stagedSubstep(w, s, dtLaw(s));
  w.meta.trajectory = rec;
  w.save('micro_16.wsg.gz');
})();

// The frame-end passes (_vaporCollisions, _iceContacts, _pushSurfaceParticles;
// js/solver.js:3290-3587) on a state perturbed so that every branch runs:
// particles pushed into the rock, above the atmosphere cap and outside the
// domain; a tight line of vapor along the sea surface (chained contacts and
// vapor-liquid contacts); a cluster of snow, part of it stuck.
function frameEnd(w, s) {
  var dx = s.dx, cx = s.cx, cy = s.cy, cz = s.cz, k = 0, p;
  for (p = 7; p < s.nP && k < 90; p += 41, k++) {
    var ex = s.px[p] - cx, ey = s.py[p] - cy, ez = s.pz[p] - cz;
    var er = Math.sqrt(ex * ex + ey * ey + ez * ez) || 1e-9, r;
    if (k % 3 === 0) { r = er - (0.5 + (k % 5) * 0.5) * dx; s.pflag[p] = 0; }
    else if (k % 3 === 1) { r = s.oceanR + s.atmosphereH + 0.05 + 0.01 * (k % 7); s.pflag[p] = 2; }
    else { r = s.domainR + 0.02 * (1 + (k % 4)); s.pflag[p] = 0; }
    s.px[p] = cx + ex / er * r; s.py[p] = cy + ey / er * r; s.pz[p] = cz + ez / er * r;
    s.pvx[p] = -ex / er * 0.3 + 0.01 * (k % 5); s.pvy[p] = -ey / er * 0.3; s.pvz[p] = -ez / er * 0.3;
  }
  // a line of vapor just under the sea surface, spaced 0.45 spacing apart
  var p0 = 1500, dir = [0.6, 0.64, 0.48], r0 = s.oceanR - 0.3 * dx;
  for (k = 0; k < 80 && p0 + k < s.nP; k++) {
    p = p0 + k;
    s.px[p] = cx + dir[0] * r0 + k * 0.45 * s.spacing; s.py[p] = cy + dir[1] * r0 + 0.02 * s.spacing * (k % 3);
    s.pz[p] = cz + dir[2] * r0;
    s.pflag[p] = k % 5 === 4 ? 3 : 2;
    s.pvx[p] = 0.2 * ((k % 4) - 1.5); s.pvy[p] = 0.1 * ((k % 3) - 1); s.pvz[p] = 0.05 * (k % 2);
    s.pWx[p] = 0.3 * (k % 2); s.pWy[p] = 0; s.pWz[p] = -0.2 * (k % 3 === 0 ? 1 : 0);
  }
  // a cluster of snow, every fourth grain stuck
  var q0 = 3000;
  for (k = 0; k < 60 && q0 + k < s.nP; k++) {
    p = q0 + k;
    s.px[p] = s.px[q0] + (k % 6) * 0.5 * s.spacing; s.py[p] = s.py[q0] + ((k / 6) | 0) * 0.5 * s.spacing;
    s.pz[p] = s.pz[q0];
    s.pflag[p] = 5; s.pAir[p] = k % 4 === 0 ? 1 : 0.2;
    s.pvx[p] = 0.1 * ((k % 3) - 1); s.pvy[p] = -0.2; s.pvz[p] = 0.05 * (k % 2);
  }
  s.spinOn = true;
  w.meta.frame_end = { spinOn: true, sunPos: s.sunPos, waterTopY: s.waterTopY };
  dumpParticles(w, s, 'fe0_');
  s._vaporCollisions();
  w.meta.frame_end.vaporCount = s.vaporCount; w.meta.frame_end.nightVaporCount = s.nightVaporCount;
  dumpParticles(w, s, 'fe1_');
  s._iceContacts();
  dumpParticles(w, s, 'fe2_');
  s._pushSurfaceParticles();
  dumpParticles(w, s, 'fe3_');
}

// ---------------------------------------------------------------- scene C
// Full physics at 32^3: sun, evaporation, currents, vorticity, thermal tick
// and the phase pass; temperatures from a 270-300 K latitude profile on the
// reference's normalised scale (T = 0.105 + (T_K - 273.15) / 60).
(function sceneC() {
  if (ONLY && ONLY !== 'thermal_32') return;
  var n = 32, W = 5.1, dx = W / n;
  var s = new FluidSolver({ nx: n, ny: n, nz: n, dx: dx, targetParticles: 10000, mode: 'sphere',
    coreR: 0.25 * W, terrainSeed: 20260116, sunActivity: 0.6, evapIntensity: 1.5, vorticity: 0.05,
    currents: 0.3 });
  s.resetWater(0.4 * W - 0.25 * W);
  s.pic = 0.05; s.omega = 1.7; s.pressureIters = 40; s.gravity = 9.8; s.maxSpeed = 3.2;
  s.sunPos = [9, 4, 4];
  for (var p = 0; p < s.nP; p++) {
    var rx = s.px[p] - s.cx, ry = s.py[p] - s.cy, rz = s.pz[p] - s.cz;
    var r = Math.sqrt(rx * rx + ry * ry + rz * rz) || 1e-9, sl = ry / r;
    var TK = 270 + 30 * (1 - sl * sl);
    s.pT[p] = 0.105 + (TK - 273.15) / 60;
  }
  var w = new Writer();
  w.meta.scene = 'thermal_32';
  w.meta.fill = params(s);
  s._rasterizeStatic();
  dumpWorld(w, s);
  w.add('terrain_solid_full', copy(s.terrain.solid));
  for (var f = 0; f < 40; f++) { s.step(1 / 60); s.splatDensity(); }
  var tick = { dtTick: 1 / 25, simTime: s._simTime, sunActivity: s.sunActivity, evapIntensity: s.evapIntensity,
    heatK: s.heatK, Tamb: s.Tamb, shadeCool: s.shadeCool, snowT: s.snowT, meltT: s.meltT, rainT: s.rainT,
    cloudT: s.cloudT, cloudP: s.cloudP, rainLiftAge: s.rainLiftAge, currents: s.currents, sunPos: s.sunPos,
    Rlo: s.terrain.Rlo, Rhi: s.terrain.Rhi, evS_before: s._evS };
  dumpParticles(w, s, 't0_');
  w.add('t0_dens', copy(s.dens));
  w.add('t0_cellType', copy(s.cellType));
  dumpGrid(w, s, 't0_', ['u', 'v', 'w']);
  s._updateHeat(1 / 25);
  dumpParticles(w, s, 't1_');
  s._updateEvaporation(1 / 25);
  tick.evapCount = s.evapCount; tick.evS_after = s._evS;
  dumpParticles(w, s, 't2_');
  s._updateCurrents(1 / 25);
  dumpParticles(w, s, 't3_');
  s._updatePhaseChanges(1 / 60);
  dumpParticles(w, s, 't4_');
  w.meta.tick = tick;
  var nv = 0;
  for (p = 0; p < s.nP; p++) if (s.pflag[p] === 2 || s.pflag[p] === 3) nv++;
  w.meta.vapor = nv;
  w.meta.rng_before = { evS: s._evS, cfS: s._cfS, bfS: s._bfS };
  stagedSubstep(w, s, dtLaw(s));
  frameEnd(w, s);
  w.save('thermal_32.wsg.gz');
})();

// ---------------------------------------------------------------- scene D
// Balls on the 32^3 planet (solver.js:918-938, 1005-1026, 2294-2564): one
// floating at sea level, one dropped from above the sea, an overlapping pair
// in the water (ball-ball contact) and one sunk into the rock of a continent
// (terrain contact).
(function sceneD() {
  if (ONLY && ONLY !== 'balls_32') return;
  var n = 32, W = 5.1, dx = W / n;
  var s = new FluidSolver({ nx: n, ny: n, nz: n, dx: dx, targetParticles: 10000, mode: 'sphere',
    coreR: 0.25 * W, terrainSeed: 20260117 });
  s.resetWater(0.4 * W - 0.25 * W);
  s.pic = 0.05; s.omega = 1.7; s.pressureIters = 40; s.gravity = 9.8; s.maxSpeed = 3.2; s.vorticity = 0;
  // candidate directions on a Fibonacci sphere, split into deep sea and land
  var sea = [], land = [], ga = Math.PI * (3 - Math.sqrt(5));
  for (var i = 0; i < 300; i++) {
    var y = 1 - 2 * (i + 0.5) / 300, rr = Math.sqrt(1 - y * y), th = ga * i;
    var d = [Math.cos(th) * rr, y * 0.9, Math.sin(th) * rr];
    var dl = Math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d = [d[0] / dl, d[1] / dl, d[2] / dl];
    var Rt = s.terrainRadiusAt(s.cx + d[0] * s.oceanR, s.cy + d[1] * s.oceanR, s.cz + d[2] * s.oceanR);
    if (Rt < s.oceanR - 0.35) sea.push(d); else if (Rt > s.oceanR + 0.1) land.push([d, Rt]);
  }
  function at(d, r) { return [s.cx + d[0] * r, s.cy + d[1] * r, s.cz + d[2] * r]; }
  var p0 = at(sea[3], s.oceanR);
  s.addBall(p0[0], p0[1], p0[2], 0.22, 500);
  var p1 = at(sea[Math.floor(sea.length / 2)], s.oceanR + 0.45);
  s.addBall(p1[0], p1[1], p1[2], 0.2, 900);
  var d1 = sea[Math.floor(sea.length / 2)];
  s.balls[1].vx = -d1[0]; s.balls[1].vy = -d1[1]; s.balls[1].vz = -d1[2];
  var d2 = sea[sea.length - 4], t2 = [d2[1], -d2[0], 0], tl = Math.sqrt(t2[0] * t2[0] + t2[1] * t2[1]) || 1;
  var c2 = at(d2, s.oceanR - 0.12);
  s.addBall(c2[0] + 0.15 * t2[0] / tl, c2[1] + 0.15 * t2[1] / tl, c2[2], 0.18, 1200);
  s.addBall(c2[0] - 0.15 * t2[0] / tl, c2[1] - 0.15 * t2[1] / tl, c2[2], 0.18, 700);
  var L = land[Math.floor(land.length / 3)], p4 = at(L[0], L[1] + 0.05);
  s.addBall(p4[0], p4[1], p4[2], 0.2, 800);
  var w = new Writer();
  w.meta.scene = 'balls_32';
  w.meta.fill = params(s);
  s._rasterizeStatic();
  dumpWorld(w, s);
  dumpParticles(w, s, 'fill_');
  w.add('fill_balls', ballState(s));
  w.add('fill_q', copy(s.q));
  var rec = [], tb = [];
  for (var k = 0; k < 20; k++) {
    var dt = dtLaw(s);
    s._substep(dt);
    rec.push({ dt: dt, umax: s.umax, ke: ke(s), airborne: s.airborneCount });
    tb.push(ballState(s));
  }
  var all = new Float64Array(20 * 9 * s.balls.length);
  tb.forEach(function (a, k) { all.set(a, k * a.length); });
  w.add('traj_balls', all);
  w.meta.trajectory = rec;
  dumpParticles(w, s, 'traj_end_');
  stagedSubstep(w, s, dtLaw(s));
  w.save('balls_32.wsg.gz');
})();

// ---------------------------------------------------------------- scene E
// Metaballs.build (js/surface.js:178-284) on a settled 32^3 planet for a
// sweep of tension and detail settings, and on the 7^3 particle block of
// the reference's own metaball test (test/metaball.test.js:31-46).
(function sceneE() {
  if (ONLY && ONLY !== 'metaballs') return;
  var MB = global.MarchingTetrahedra.Metaballs;
  var w = new Writer();
  w.meta.scene = 'metaballs';
  var n = 32, W = 5.1, dx = W / n;
  var s = new FluidSolver({ nx: n, ny: n, nz: n, dx: dx, targetParticles: 10000, mode: 'sphere',
    coreR: 0.25 * W, terrainSeed: 20260118 });
  s.resetWater(0.4 * W - 0.25 * W);
  s.pic = 0.05; s.omega = 1.7; s.pressureIters = 40; s.gravity = 9.8; s.maxSpeed = 3.2; s.vorticity = 0;
  for (var k = 0; k < 40; k++) s._substep(dtLaw(s));
  w.meta.planet = { nx: n, dx: dx, W: s.W, spacing: s.spacing, nP: s.nP };
  w.add('planet_px', copy(s.px, s.nP)); w.add('planet_py', copy(s.py, s.nP)); w.add('planet_pz', copy(s.pz, s.nP));
  w.add('planet_pflag', copy(s.pflag, s.nP));
  var cases = [[0, null], [0.5, null], [1, null], [0.5, 0.6], [0.5, 1.4], [0.25, 0.3], [0.75, 2.5]];
  w.meta.planet_cases = [];
  cases.forEach(function (c, ci) {
    var m = c[1] === null ? MB.build(s, c[0]) : MB.build(s, c[0], c[1]);
    w.meta.planet_cases.push({ tension: c[0], detail: c[1], count: m.count, volume: m.volume, iso: MB.iso(c[0]) });
    if (m.count < 70000) { w.add('planet_pos_' + ci, copy(m.pos, m.count * 3)); w.add('planet_nrm_' + ci, copy(m.nrm, m.count * 3)); }
  });
  // the reference test's stub: 16^3 lattice, spacing dx/4, a 7^3 block
  var sn = 16, sdx = 1 / sn, sp = sdx * 0.25, part = [];
  for (var kk = 0; kk < 7; kk++) for (var jj = 0; jj < 7; jj++) for (var ii = 0; ii < 7; ii++)
    part.push((0.5 - 3 * sp) + ii * sp, (0.5 - 3 * sp) + jj * sp, (0.5 - 3 * sp) + kk * sp);
  var np = part.length / 3, spx = new Float32Array(np), spy = new Float32Array(np), spz = new Float32Array(np);
  for (var q = 0; q < np; q++) { spx[q] = part[3 * q]; spy[q] = part[3 * q + 1]; spz[q] = part[3 * q + 2]; }
  var stub = { nx: sn, ny: sn, nz: sn, dx: sdx, W: sn * sdx, H: sn * sdx, D: sn * sdx, spacing: sp, nP: np,
    px: spx, py: spy, pz: spz, pflag: new Int8Array(np) };
  w.meta.stub = { nx: sn, dx: sdx, spacing: sp, nP: np, cases: [] };
  w.add('stub_px', spx); w.add('stub_py', spy); w.add('stub_pz', spz);
  [0, 0.25, 0.5, 0.75, 1].forEach(function (t) {
    var m = MB.build(stub, t);
    w.meta.stub.cases.push({ tension: t, detail: null, count: m.count, volume: m.volume });
  });
  [0.6, 1.4, 0.1, 9].forEach(function (d) {
    var m = MB.build(stub, 0.5, d);
    w.meta.stub.cases.push({ tension: 0.5, detail: d, count: m.count, volume: m.volume });
  });
  w.save('metaballs.wsg.gz');
})();
