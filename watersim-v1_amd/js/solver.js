/*
 * FluidSolver on MI355X: the construct / step / readback surface of the
 * reference solver (js/solver.js:45-3924, ocean-planet mode) over the HIP
 * library's C ABI, reached through the N-API addon in ../addon.
 *
 * The simulation state lives in device memory. The typed arrays callers
 * know (px .. pPh, cellType, u/v/w, q, dens ...) are host mirrors:
 *  - reading a field after a device call downloads it once;
 *  - a field that was read may have been written, so it is uploaded before
 *    the next device call;
 *  - fields nobody touches never cross PCIe.
 * Scalar tunables are plain properties pushed to the device before every
 * call. There is no CPU path: without the addon or a GPU, construction of
 * the world throws.
 */
'use strict';

var path = require('path');

var addon = null;
function lib() {
  if (!addon) {
    var where = path.join(__dirname, '..', '_build', 'watersim_napi.node');
    try {
      addon = require(where);
    } catch (e) {
      throw new Error('watersim: the HIP addon is not built (' + where + '): ' + e.message);
    }
  }
  return addon;
}

var AIR = 0, FLUID = 1, SOLID = 2;

// ws_stage identifiers (include/watersim.h)
var STAGE = {
  clear: 0, rasterize: 1, p2g: 2, bc: 3, snapshot: 4, viscosity: 5, gravity: 6, pressure: 7,
  project: 8, vorticity: 9, clamp: 10, extrapolate: 11, g2p: 12, advect: 13, heat: 14,
  evaporation: 15, currents: 16, phase: 17, pushSurface: 18, vaporCollisions: 19, iceContacts: 20,
  pushBalls: 21, updateBalls: 22
};

// scalars mirrored to the device before each call (js/solver.js:56-75, 137-193)
var TUNABLES = ['cfl', 'pic', 'gravity', 'viscosity', 'pressureIters', 'omega', 'maxSpeed', 'vorticity',
  'currents', 'shadeCool', 'substeps', 'heatK', 'Tamb', 'cloudT', 'cloudP', 'rainT', 'rainLiftAge', 'snowT',
  'meltT', 'evapIntensity', 'sunActivity', 'atmosphereH', 'iso', 'pressureTol', 'maxPressureIters'];
var CEIL = { linear: 0, quadratic: 1, exponential: 2 };

// device fields: [name on the device, constructor, extent]
var PARTICLE_FIELDS = {
  px: Float32Array, py: Float32Array, pz: Float32Array, pvx: Float32Array, pvy: Float32Array,
  pvz: Float32Array, pT: Float32Array, pAir: Float32Array, pLight: Float32Array, pWx: Float32Array,
  pWy: Float32Array, pWz: Float32Array, pPh: Float32Array, pflag: Uint8Array, pcool: Uint8Array,
  pDepth: Uint8Array
};
var GRID_FIELDS = {
  cellType: [Uint8Array, 'C'], q: [Float32Array, 'C'], rhs: [Float32Array, 'C'],
  u: [Float32Array, 'U'], v: [Float32Array, 'V'], w: [Float32Array, 'W'],
  uW: [Float32Array, 'U'], vW: [Float32Array, 'V'], wW: [Float32Array, 'W'],
  uO: [Float32Array, 'U'], vO: [Float32Array, 'V'], wO: [Float32Array, 'W'],
  validU: [Uint8Array, 'U'], validV: [Uint8Array, 'V'], validW: [Uint8Array, 'W'],
  dens: [Float32Array, 'D']
};

function FluidSolver(opts) {
  opts = opts || {};
  // solver.js:78; the box pool is not part of this build: device calls in
  // 'pool' mode throw (callers set mode = 'sphere', as the reference's
  // tests do after construction)
  this.mode = opts.mode === 'sphere' ? 'sphere' : 'pool';
  this.nx = opts.nx | 0; this.ny = opts.ny | 0; this.nz = opts.nz | 0;
  this.dx = opts.dx;
  // tunables with the reference defaults
  this.substeps = 1; this.cfl = 3.0; this.pic = 0.15; this.gravity = 9.81; this.viscosity = 5e-4;
  this.pressureIters = 20; this.omega = 1.55; this.maxSpeed = 3.2;
  this.vorticity = opts.vorticity || 0;
  this.currents = opts.currents || 0;
  this.shadeCool = 0.035;
  this.targetParticles = opts.targetParticles || 110000;
  this.coreR = opts.coreR || 0;
  this.oceanR = 0;
  this.terrainSeed = (opts.terrainSeed || 20260114) >>> 0;
  this.bumpiness = opts.bumpiness === undefined ? 1 : Math.max(0, Math.min(opts.bumpiness, 2.5));
  this.sunActivity = opts.sunActivity !== undefined ? opts.sunActivity
    : (opts.evaporation !== undefined ? opts.evaporation : 0);
  this.atmosphereH = opts.atmosphereH || 0.35;
  this.ceilReflect = opts.ceilReflect || 'linear';
  this.sunPos = null;
  this.heatK = 0.6; this.Tamb = 0.32;
  this.cloudT = opts.cloudT !== undefined ? opts.cloudT : 0.34;
  this.cloudP = opts.cloudP !== undefined ? opts.cloudP : 0.35;
  this.rainT = opts.rainT !== undefined ? opts.rainT : 0.22;
  this.rainLiftAge = opts.rainLiftAge !== undefined ? opts.rainLiftAge : 0.4;
  this.snowT = opts.snowT !== undefined ? opts.snowT : 0.10;
  this.meltT = opts.meltT !== undefined ? opts.meltT : this.snowT * 1.05;
  this.evapIntensity = opts.evapIntensity !== undefined ? opts.evapIntensity : 0.05;
  this.spinOn = false;
  // solver-to-residual mode (0 = fixed pressureIters sweeps)
  this.pressureTol = 0; this.maxPressureIters = 4000;

  var nx = this.nx, ny = this.ny, nz = this.nz;
  this.W = nx * this.dx; this.H = ny * this.dx; this.D = nz * this.dx;
  this.cx = this.W / 2; this.cy = this.H / 2; this.cz = this.D / 2;
  this.domainR = 0.5 * Math.min(this.W, this.H, this.D) - this.dx;
  this.nCells = nx * ny * nz;
  this.uN = (nx + 1) * ny * nz; this.vN = nx * (ny + 1) * nz; this.wN = nx * ny * (nz + 1);
  this.spacing = this.dx * 0.5;
  this.particleVolume = this.spacing * this.spacing * this.spacing;
  this.iso = 1.8;
  this.balls = [];
  this.terrain = null;
  this.capacity = 0;

  this._h = null;                 // device context
  this._nP = 0;
  this._host = {};                // host mirrors
  this._stale = {};               // device newer than the mirror
  this._dirty = {};               // mirror handed out, upload before the next call
  this._staticTypes = null;
}

// ---------------------------------------------------------------- mirrors
FluidSolver.prototype._extent = function (kind) {
  switch (kind) {
    case 'C': return this.nCells;
    case 'U': return this.uN;
    case 'V': return this.vN;
    case 'W': return this.wN;
    default: return (this.nx + 1) * (this.ny + 1) * (this.nz + 1);
  }
};

FluidSolver.prototype._mirror = function (name) {
  var a = this._host[name];
  if (!a) {
    if (PARTICLE_FIELDS[name]) a = new PARTICLE_FIELDS[name](Math.max(1, this.capacity));
    else a = new GRID_FIELDS[name][0](this._extent(GRID_FIELDS[name][1]));
    this._host[name] = a;
  }
  return a;
};

FluidSolver.prototype._view = function (name) {
  var a = this._mirror(name);
  return PARTICLE_FIELDS[name] ? a.subarray(0, this._nP) : a;
};

// reading hands the caller the mirror: fetch it if the device is newer and
// assume it may be written
FluidSolver.prototype._touch = function (name) {
  var a = this._mirror(name);
  if (this._h && this._stale[name]) {
    lib().download(this._h, name, this._view(name));
    this._stale[name] = false;
  }
  if (this._h) this._dirty[name] = true;
  return a;
};

FluidSolver.prototype._push = function () {
  var L = lib(), h = this._h;
  for (var i = 0; i < TUNABLES.length; i++) L.setF(h, TUNABLES[i], +this[TUNABLES[i]]);
  L.setF(h, 'ceilReflect', CEIL[this.ceilReflect] || 0);
  L.setF(h, 'spinOn', this.spinOn ? 1 : 0);
  var s = this.sunPos;
  L.setF(h, 'hasSun', s ? 1 : 0);
  if (s) { L.setF(h, 'sunX', s[0]); L.setF(h, 'sunY', s[1]); L.setF(h, 'sunZ', s[2]); }
  for (var name in this._dirty) {
    if (!this._dirty[name]) continue;
    L.upload(h, name, this._view(name));
    this._dirty[name] = false;
  }
};

FluidSolver.prototype._invalidate = function () {
  for (var name in PARTICLE_FIELDS) this._stale[name] = true;
  for (var g in GRID_FIELDS) this._stale[g] = true;
  this._dirty = {};
};

// a device context for calls made before resetWater (a solver whose
// particles were set by hand after _allocParticles): an empty static world
// (every cell AIR, no gravity geometry, no terrain) and every host mirror
// written so far uploaded with the next call
FluidSolver.prototype._ensure = function () {
  if (this.mode !== 'sphere') {
    throw new Error('watersim: only the ocean-planet world (mode "sphere") is part of this build');
  }
  if (this._h) return;
  var L = lib();
  this._h = L.create(this.nx, this.ny, this.nz, this.dx, Math.max(1, this.capacity));
  this._ballsOnDevice = false;
  var h = this._h;
  L.setF(h, 'coreR', this.coreR); L.setF(h, 'oceanR', this.oceanR); L.setF(h, 'spacing', this.spacing);
  L.setF(h, 'particleVolume', this.particleVolume);
  L.upload(h, 'staticTypes', new Uint8Array(this.nCells));
  L.upload(h, 'gu', new Float32Array(this.uN));
  L.upload(h, 'gv', new Float32Array(this.vN));
  L.upload(h, 'gw', new Float32Array(this.wN));
  L.setF(h, 'nP', this._nP);
  this._stale = {};
  this._dirty = {};
  for (var name in this._host) this._dirty[name] = true;
};

FluidSolver.prototype._call = function (fn) {
  this._ensure();
  this._push();
  var L = lib();
  var nb = this.balls.length;
  // the ball objects are the state callers read and write (solver.js:115);
  // the device works on a copy taken before and returned after each call
  if (nb || this._ballsOnDevice) {
    var st = new Float64Array(9 * nb);
    for (var i = 0; i < nb; i++) {
      var b = this.balls[i];
      st.set([b.x, b.y, b.z, b.vx, b.vy, b.vz, b.r, b.rho, b.submerged], 9 * i);
    }
    L.setBalls(this._h, st);
    this._ballsOnDevice = nb > 0;
  }
  var r = fn(L, this._h);
  if (nb) {
    var out = new Float64Array(9 * nb);
    L.getBalls(this._h, out);
    for (var j = 0; j < nb; j++) {
      var o = this.balls[j], q = 9 * j;
      o.x = out[q]; o.y = out[q + 1]; o.z = out[q + 2]; o.vx = out[q + 3]; o.vy = out[q + 4]; o.vz = out[q + 5];
      o.submerged = out[q + 8];
    }
  }
  this._invalidate();
  return r;
};

Object.keys(PARTICLE_FIELDS).concat(Object.keys(GRID_FIELDS)).forEach(function (name) {
  Object.defineProperty(FluidSolver.prototype, name, {
    get: function () { return this._touch(name); },
    set: function (arr) {
      var m = this._mirror(name);
      m.set(arr.subarray ? arr.subarray(0, m.length) : arr);
      if (this._h) { this._dirty[name] = true; this._stale[name] = false; }
    }
  });
});

// the reference caches grid samples as Float64Array (solver.js:199); the
// device keeps them as float, so reads return float64 copies
['U', 'V', 'W'].forEach(function (c) {
  Object.defineProperty(FluidSolver.prototype, '_adv' + c, {
    get: function () {
      var f = new Float32Array(Math.max(1, this._nP));
      if (this._h && this._nP) lib().download(this._h, 'adv' + c, f.subarray(0, this._nP));
      return Float64Array.from(f);
    }
  });
});

Object.defineProperty(FluidSolver.prototype, 'nP', {
  get: function () { return this._nP; },
  set: function (n) {
    n = n | 0;
    if (n < 0 || n > this.capacity) throw new RangeError('nP exceeds the particle capacity ' + this.capacity);
    if (this._h) {
      // keep the fields of the particles that stay valid on the host side
      for (var name in PARTICLE_FIELDS) if (this._stale[name]) this._touch(name);
      lib().setF(this._h, 'nP', n);
      for (var k in PARTICLE_FIELDS) this._dirty[k] = true;
    }
    this._nP = n;
  }
});

// _allocParticles(cap), solver.js:218-231: particle capacity. The device
// context holds its capacity for life, so growing it moves the state to a
// new context (static world, particle and grid fields and scalars).
FluidSolver.prototype._allocParticles = function (cap) {
  cap = cap | 0;
  if (this.capacity >= cap) return;
  if (!this._h) { this.capacity = cap; return; }
  var L = lib(), old = this._h, names = Object.keys(PARTICLE_FIELDS).concat(Object.keys(GRID_FIELDS));
  for (var i = 0; i < names.length; i++) this._touch(names[i]);
  var keep = {};
  ['gRamp', 'simTime', 'heatAcc', 'tn', 'dv', 'Rlo', 'Rhi'].forEach(function (k) { keep[k] = L.getF(old, k); });
  var st = new Uint8Array(this.nCells), gu = new Float32Array(this.uN), gv = new Float32Array(this.vN),
    gw = new Float32Array(this.wN);
  L.download(old, 'staticTypes', st);
  L.download(old, 'gu', gu); L.download(old, 'gv', gv); L.download(old, 'gw', gw);
  var hosts = this._host;
  this._host = {};
  for (var n in hosts) {
    var a = hosts[n];
    if (PARTICLE_FIELDS[n]) { var b = new PARTICLE_FIELDS[n](cap); b.set(a.subarray(0, Math.min(a.length, cap))); a = b; }
    this._host[n] = a;
  }
  L.destroy(old);
  this._h = null;
  this.capacity = cap;
  this._ensure();
  var h = this._h;
  L.upload(h, 'staticTypes', st);
  L.upload(h, 'gu', gu); L.upload(h, 'gv', gv); L.upload(h, 'gw', gw);
  L.setF(h, 'tn', keep.tn); L.setF(h, 'dv', keep.dv); L.setF(h, 'Rlo', keep.Rlo); L.setF(h, 'Rhi', keep.Rhi);
  if (this.terrain) { L.upload(h, 'terrainR', this.terrain.R); L.upload(h, 'terrainSolid', this.terrain.solid); }
  L.setF(h, 'gRamp', keep.gRamp); L.setF(h, 'simTime', keep.simTime); L.setF(h, 'heatAcc', keep.heatAcc);
};

// the gravity ramp timer (solver.js:2657-2658), read and written by callers
Object.defineProperty(FluidSolver.prototype, '_gRamp', {
  get: function () { return this._h ? lib().getF(this._h, 'gRamp') : 0; },
  set: function (v) { this._ensure(); lib().setF(this._h, 'gRamp', +v); }
});

function devScalar(name, dev) {
  Object.defineProperty(FluidSolver.prototype, name, {
    get: function () { return this._h ? lib().getF(this._h, dev) : 0; }
  });
}
devScalar('umax', 'umax');
devScalar('dtLast', 'dtLast');
devScalar('airborneCount', 'airborneCount');
devScalar('_simTime', 'simTime');
devScalar('evapCount', 'evapCount');

// ---------------------------------------------------------------- world
// resetWater(depth), solver.js:611 / _resetWaterSphere :651-731: the voxel
// planet, the ocean fill at rest, radial gravity and the iso calibration
FluidSolver.prototype.resetWater = function (depth) {
  if (this.mode !== 'sphere') {
    throw new Error('watersim: only the ocean-planet world (mode "sphere") is part of this build');
  }
  var L = lib();
  this.balls.length = 0;
  var w = L.worldBuild({
    nx: this.nx, ny: this.ny, nz: this.nz, dx: this.dx, coreR: this.coreR, depth: depth,
    bumpiness: this.bumpiness, seed: this.terrainSeed, targetParticles: this.targetParticles
  });
  this.oceanR = w.oceanR;
  this.spacing = w.spacing;
  this.particleVolume = w.particleVolume;
  this.landFrac = w.landFrac;
  this.terrain = {
    n: w.tn, dv: w.dv, solid: w.solid, R: w.R, Rlo: w.Rlo, Rhi: w.Rhi, landFrac: w.landFrac,
    seed: this.terrainSeed
  };
  this._staticTypes = w.staticTypes;
  if (this._h) { L.destroy(this._h); this._h = null; }
  this.capacity = w.capacity;
  this._host = {};
  this._h = L.create(this.nx, this.ny, this.nz, this.dx, w.capacity);
  this._ballsOnDevice = false;
  var h = this._h;
  L.setF(h, 'coreR', this.coreR); L.setF(h, 'oceanR', w.oceanR); L.setF(h, 'spacing', w.spacing);
  L.setF(h, 'particleVolume', w.particleVolume); L.setF(h, 'dv', w.dv); L.setF(h, 'tn', w.tn);
  L.setF(h, 'Rlo', w.Rlo); L.setF(h, 'Rhi', w.Rhi);
  L.upload(h, 'staticTypes', w.staticTypes);
  L.upload(h, 'gu', w.gravity.subarray(0, this.uN));
  L.upload(h, 'gv', w.gravity.subarray(this.uN, this.uN + this.vN));
  L.upload(h, 'gw', w.gravity.subarray(this.uN + this.vN));
  L.upload(h, 'terrainR', w.R);
  L.upload(h, 'terrainSolid', w.solid);
  this._nP = w.nP;
  L.setF(h, 'nP', w.nP);
  this._stale = {}; this._dirty = {};
  this._mirror('px').set(w.px); this._mirror('py').set(w.py); this._mirror('pz').set(w.pz);
  this._mirror('pT').fill(this.Tamb);
  for (var name in PARTICLE_FIELDS) L.upload(h, name, this._view(name));
  L.setF(h, 'simTime', 0); L.setF(h, 'heatAcc', 0);
  L.setF(h, 'gRamp', 0);                    // gravity eases in (solver.js:727)
  this.splatDensity();
  this.iso = this.autotuneIso();
};

// ---------------------------------------------------------------- dynamics
FluidSolver.prototype.step = function (dtFrame) {
  this.substepsLast = this._call(function (L, h) { return L.stepFrame(h, dtFrame); });
};
FluidSolver.prototype._substep = function (dt) { this._call(function (L, h) { L.substep(h, dt); }); };
FluidSolver.prototype.runSubsteps = function (n) { this._call(function (L, h) { L.runSubsteps(h, n); }); };

function stageMethod(id) {
  return function (dt) {
    var d = dt === undefined ? 0 : dt;
    this._call(function (L, h) { L.stage(h, id, d); });
  };
}
FluidSolver.prototype._rasterize = stageMethod(STAGE.rasterize);
FluidSolver.prototype._p2g = stageMethod(STAGE.p2g);
FluidSolver.prototype._applyBC = stageMethod(STAGE.bc);
FluidSolver.prototype._viscosity = stageMethod(STAGE.viscosity);
FluidSolver.prototype._applyPlanetGravity = stageMethod(STAGE.gravity);
FluidSolver.prototype._pressureSolve = stageMethod(STAGE.pressure);
FluidSolver.prototype._project = stageMethod(STAGE.project);
FluidSolver.prototype._vorticityConfinement = stageMethod(STAGE.vorticity);
FluidSolver.prototype._extrapolate = stageMethod(STAGE.extrapolate);
FluidSolver.prototype._g2p = stageMethod(STAGE.g2p);
FluidSolver.prototype._advect = stageMethod(STAGE.advect);
FluidSolver.prototype._updateHeat = stageMethod(STAGE.heat);
FluidSolver.prototype._updateEvaporation = stageMethod(STAGE.evaporation);
FluidSolver.prototype._updateCurrents = stageMethod(STAGE.currents);
FluidSolver.prototype._updatePhaseChanges = stageMethod(STAGE.phase);
FluidSolver.prototype._pushSurfaceParticles = stageMethod(STAGE.pushSurface);
FluidSolver.prototype._vaporCollisions = stageMethod(STAGE.vaporCollisions);
FluidSolver.prototype._iceContacts = stageMethod(STAGE.iceContacts);
FluidSolver.prototype._pushOutOfBalls = stageMethod(STAGE.pushBalls);
FluidSolver.prototype._updateBalls = stageMethod(STAGE.updateBalls);

// ---------------------------------------------------------------- surface
FluidSolver.prototype.splatDensity = function () {
  this._call(function (L, h) { L.splatDensity(h); });
};
FluidSolver.prototype.autotuneIso = function () {
  return this._call(function (L, h) { return L.autotuneIso(h); });
};
FluidSolver.prototype.kineticEnergy = function () {
  return this._call(function (L, h) { return L.kineticEnergy(h); });
};

// ---------------------------------------------------------------- host helpers
// terrainRadiusAt, solver.js:511-535: trilinear on the radius lattice
FluidSolver.prototype.terrainRadiusAt = function (x, y, z) {
  var t = this.terrain;
  if (!t) return this.coreR;
  var n = t.n, R = t.R;
  var gx = x / t.dv - 0.5, gy = y / t.dv - 0.5, gz = z / t.dv - 0.5;
  var i = Math.min(Math.max(gx | 0, 0), n - 2);
  var j = Math.min(Math.max(gy | 0, 0), n - 2);
  var k = Math.min(Math.max(gz | 0, 0), n - 2);
  var fx = Math.min(Math.max(gx - i, 0), 1), fy = Math.min(Math.max(gy - j, 0), 1);
  var fz = Math.min(Math.max(gz - k, 0), 1);
  var b = (k * n + j) * n + i, sj = n, sk = n * n;
  var x00 = R[b] + (R[b + 1] - R[b]) * fx;
  var x10 = R[b + sj] + (R[b + sj + 1] - R[b + sj]) * fx;
  var x01 = R[b + sk] + (R[b + sk + 1] - R[b + sk]) * fx;
  var x11 = R[b + sk + sj] + (R[b + sk + sj + 1] - R[b + sk + sj]) * fx;
  var y0 = x00 + (x10 - x00) * fy, y1 = x01 + (x11 - x01) * fy;
  return y0 + (y1 - y0) * fz;
};

// _rockCellAt, solver.js:542-559: the static raster at the (clamped) cell
FluidSolver.prototype._rockCellAt = function (x, y, z) {
  if (!this._staticTypes) return false;
  var dx = this.dx, nx = this.nx, ny = this.ny, nz = this.nz;
  var i = Math.min(Math.max((x / dx) | 0, 1), nx - 2);
  var j = Math.min(Math.max((y / dx) | 0, 1), ny - 2);
  var k = Math.min(Math.max((z / dx) | 0, 1), nz - 2);
  return this._staticTypes[(k * ny + j) * nx + i] === SOLID;
};

// _voxelRockAt, solver.js:565-574: the exact voxel
FluidSolver.prototype._voxelRockAt = function (x, y, z) {
  var t = this.terrain;
  if (!t) return false;
  var n = t.n;
  var i = Math.min(Math.max((x / t.dv) | 0, 0), n - 1);
  var j = Math.min(Math.max((y / t.dv) | 0, 0), n - 1);
  var k = Math.min(Math.max((z / t.dv) | 0, 0), n - 1);
  return t.solid[(k * n + j) * n + i] === 1;
};

// _onRockBelow, solver.js:581-587
FluidSolver.prototype._onRockBelow = function (x, y, z) {
  if (!this.terrain) return false;
  var ex = x - this.cx, ey = y - this.cy, ez = z - this.cz;
  var s = this.dx * 0.9 / (Math.sqrt(ex * ex + ey * ey + ez * ez) || 1e-9);
  return this._rockCellAt(x - ex * s, y - ey * s, z - ez * s);
};

// applyImpulseSphere, solver.js:3852-3864: host-side edit of the particle
// velocities (uploaded with the next device call)
FluidSolver.prototype.applyImpulseSphere = function (cx, cy, cz, R, fx, fy, fz) {
  var px = this.px, py = this.py, pz = this.pz, vx = this.pvx, vy = this.pvy, vz = this.pvz;
  var fl = this.pflag, cool = this.pcool, R2 = R * R;
  for (var p = 0; p < this._nP; p++) {
    var ex = px[p] - cx, ey = py[p] - cy, ez = pz[p] - cz, d2 = ex * ex + ey * ey + ez * ez;
    if (d2 > R2) continue;
    var f = 1 - Math.sqrt(d2) / R;
    vx[p] += fx * f; vy[p] += fy * f; vz[p] += fz * f;
    if (fl[p] !== 0) fl[p] = 0;
    cool[p] = 0;
  }
};

// stirAt, solver.js:3895-3908
FluidSolver.prototype.stirAt = function (cx, cy, cz, ux, uy, uz, R, dt) {
  var px = this.px, py = this.py, pz = this.pz, vx = this.pvx, vy = this.pvy, vz = this.pvz;
  var fl = this.pflag, R2 = R * R, k = Math.min(1, 14 * dt);
  for (var p = 0; p < this._nP; p++) {
    var ex = px[p] - cx, ey = py[p] - cy, ez = pz[p] - cz, d2 = ex * ex + ey * ey + ez * ez;
    if (d2 > R2) continue;
    var f = 1 - Math.sqrt(d2) / R;
    vx[p] += (ux * 1.25 - vx[p]) * k * f;
    vy[p] += (uy * 1.25 - vy[p]) * k * f;
    vz[p] += (uz * 1.25 - vz[p]) * k * f;
    fl[p] = 0;
  }
};

// ---------------------------------------------------------------- sampling
// sampleVel / _sampleVel3, solver.js:811-854: the device sampler G2P uses, so
// sampleVel at a particle equals its _advU/_advV/_advW entry after _g2p
var _pt = new Float32Array(3), _pv = new Float32Array(3);
FluidSolver.prototype._sampleVel3 = function (x, y, z, out) {
  _pt[0] = x; _pt[1] = y; _pt[2] = z;
  this._call(function (L, h) { L.sampleVel(h, _pt, _pv); });
  out[0] = _pv[0]; out[1] = _pv[1]; out[2] = _pv[2];
  return out;
};
FluidSolver.prototype.sampleVel = function (x, y, z, out) {
  return this._sampleVel3(x, y, z, out || [0, 0, 0]);
};

// _faceCoords, solver.js:737-748: base index and fractions of component
// comp's face grid at (x, y, z), with the reference's clamps (i0 in
// [0, nx-1] with tx = 1 at the top, j0 and k0 in [0, n-2])
FluidSolver.prototype._faceCoords = function (comp, x, y, z) {
  var dx = this.dx;
  var a = comp === 0 ? x / dx : x / dx - 0.5;
  var b = comp === 1 ? y / dx : y / dx - 0.5;
  var c = comp === 2 ? z / dx : z / dx - 0.5;
  var f = { i0: Math.floor(a), j0: Math.floor(b), k0: Math.floor(c), tx: 0, ty: 0, tz: 0 };
  if (f.i0 < 0) { f.i0 = 0; f.tx = 0; } else if (f.i0 > this.nx - 1) { f.i0 = this.nx - 1; f.tx = 1; } else f.tx = a - f.i0;
  if (f.j0 < 0) { f.j0 = 0; f.ty = 0; } else if (f.j0 > this.ny - 2) { f.j0 = this.ny - 2; f.ty = 1; } else f.ty = b - f.j0;
  if (f.k0 < 0) { f.k0 = 0; f.tz = 0; } else if (f.k0 > this.nz - 2) { f.k0 = this.nz - 2; f.tz = 1; } else f.tz = c - f.k0;
  return f;
};

// row and slab strides and base index of the face grid of comp
FluidSolver.prototype._faceIndex = function (comp, f) {
  var nx = this.nx, ny = this.ny;
  if (comp === 0) return { sj: nx + 1, sk: ny * (nx + 1), idx: (f.k0 * ny + f.j0) * (nx + 1) + f.i0 };
  if (comp === 1) return { sj: nx, sk: (ny + 1) * nx, idx: (f.k0 * (ny + 1) + f.j0) * nx + f.i0 };
  return { sj: nx, sk: ny * nx, idx: (f.k0 * ny + f.j0) * nx + f.i0 };
};

// the trilinear sum of solver.js:777-782 / :800-805 (x, then y, then z) in
// double over any array of the face grid's layout
function trilinear(arr, o, f) {
  var s0 = 1 - f.tx, s1 = f.tx, t0 = 1 - f.ty, t1 = f.ty, r0 = 1 - f.tz, r1 = f.tz;
  var i = o.idx, a = i + o.sj, b = i + o.sk, d = b + o.sj;
  return ((arr[i] * s0 + arr[i + 1] * s1) * t0 + (arr[a] * s0 + arr[a + 1] * s1) * t1) * r0 +
    ((arr[b] * s0 + arr[b + 1] * s1) * t0 + (arr[d] * s0 + arr[d + 1] * s1) * t1) * r1;
}

// _sampleFace, solver.js:750-783 (host arithmetic, any caller array)
FluidSolver.prototype._sampleFace = function (arr, comp, x, y, z) {
  var f = this._faceCoords(comp, x, y, z);
  return trilinear(arr, this._faceIndex(comp, f), f);
};

// _sampleFacePair, solver.js:788-809: two grids at one position
var _pair = [0, 0];
FluidSolver.prototype._sampleFacePair = function (arrA, arrB, comp, x, y, z) {
  var f = this._faceCoords(comp, x, y, z), o = this._faceIndex(comp, f);
  _pair[0] = trilinear(arrA, o, f);
  _pair[1] = trilinear(arrB, o, f);
  return _pair;
};

// _markFluidCells(p0, p1), solver.js:945-1000, on the device
FluidSolver.prototype._markFluidCells = function (p0, p1) {
  var a = p0 === undefined ? 0 : p0, b = p1 === undefined ? this._nP : p1;
  this._call(function (L, h) { L.markFluid(h, a, b); });
};

// _buildPlanetGravity, solver.js:1177-1210: the unit-g radial geometry for
// the current ocean radius (the gravity slider scales it when applied)
FluidSolver.prototype._buildPlanetGravity = function () {
  var g = lib().planetGravity(this.nx, this.ny, this.nz, this.dx, this.oceanR);
  this._gu = g.subarray(0, this.uN);
  this._gv = g.subarray(this.uN, this.uN + this.vN);
  this._gw = g.subarray(this.uN + this.vN);
  this._ensure();
  var L = lib(), h = this._h;
  L.upload(h, 'gu', this._gu); L.upload(h, 'gv', this._gv); L.upload(h, 'gw', this._gw);
};

// dryLandFraction, solver.js:3593-3622: share of 1200 spiral directions
// whose rock stands above sea level with no fluid cell on it (host, over
// the terrain and the cell types)
FluidSolver.prototype.dryLandFraction = function () {
  if (this.mode !== 'sphere' || !this.terrain) return 0;
  var cx = this.cx, cy = this.cy, cz = this.cz, dx = this.dx;
  var type = this.cellType, nx = this.nx, ny = this.ny, nz = this.nz;
  var t = this.terrain, n = t.n, dv = t.dv, solid = t.solid;
  var N = 1200, dry = 0;
  for (var i = 0; i < N; i++) {
    var ang = i * 2.399963, u = 1 - 2 * (i + 0.5) / N, sq = Math.sqrt(1 - u * u);
    var lx = sq * Math.cos(ang), ly = u, lz = sq * Math.sin(ang);
    var top = -1;
    for (var r = t.Rhi + dv; r > t.Rlo - dv; r -= dv * 0.5) {
      var sx = (cx + lx * r) / dv | 0, sy = (cy + ly * r) / dv | 0, sz = (cz + lz * r) / dv | 0;
      if (sx < 0 || sy < 0 || sz < 0 || sx >= n || sy >= n || sz >= n) continue;
      if (solid[(sz * n + sy) * n + sx]) { top = r; break; }
    }
    if (top < 0 || top <= this.oceanR) continue;
    var wet = false;
    for (var k = 1; k <= 4; k++) {
      var rp = top + k * dx * 0.55;
      var ii = (cx + lx * rp) / dx | 0, jj = (cy + ly * rp) / dx | 0, kk = (cz + lz * rp) / dx | 0;
      if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) break;
      var ct = type[(kk * ny + jj) * nx + ii];
      if (ct === SOLID) continue;
      if (ct === FLUID) wet = true;
      break;
    }
    if (!wet) dry++;
  }
  return dry / N;
};

// waveImpulse(dvx), solver.js:3866-3888 (planet branch): a tidal bulge,
// water pushed radially outward around a random direction (host, over the
// particle mirrors; uploaded with the next device call)
FluidSolver.prototype.waveImpulse = function (dvx) {
  var px = this.px, py = this.py, pz = this.pz, vx = this.pvx, vy = this.pvy, vz = this.pvz;
  var th = Math.random() * Math.PI * 2, ph = Math.acos(2 * Math.random() - 1);
  var bx = Math.sin(ph) * Math.cos(th), by = Math.cos(ph), bz = Math.sin(ph) * Math.sin(th);
  for (var p = 0; p < this._nP; p++) {
    var ex = px[p] - this.cx, ey = py[p] - this.cy, ez = pz[p] - this.cz;
    var er = Math.sqrt(ex * ex + ey * ey + ez * ez) || 1e-9;
    var pr = (ex * bx + ey * by + ez * bz) / er;
    if (pr <= 0) continue;
    var f = pr * pr * pr * (0.7 + Math.random() * 0.3);
    vx[p] += ex / er * dvx * f;
    vy[p] += ey / er * dvx * f;
    vz[p] += ez / er * dvx * f;
  }
};

// The particle fields a renderer reads each frame (scene.updateParticles,
// js/scene.js:1595-1700) in one batched readback: typed arrays over pinned
// host memory, valid until the next readRender call. `fields` lists names
// (default: all of px py pz pflag pT pLight pDepth pPh pWx pWy pWz).
var RENDER_BITS = { px: 1, py: 2, pz: 4, pflag: 8, pT: 16, pLight: 32, pDepth: 64, pPh: 128, pWx: 256,
  pWy: 512, pWz: 1024 };
FluidSolver.prototype.readRender = function (fields) {
  var mask = 0;
  (fields || Object.keys(RENDER_BITS)).forEach(function (k) { mask |= RENDER_BITS[k] || 0; });
  return this._call(function (L, h) { return L.readRender(h, mask); });
};

// addBall, solver.js:3910-3912
FluidSolver.prototype.addBall = function (x, y, z, r, rho) {
  this.balls.push({ x: x, y: y, z: z, vx: 0, vy: 0, vz: 0, r: r, rho: rho, submerged: 0 });
};

// releases the device context now instead of at garbage collection
FluidSolver.prototype.dispose = function () {
  if (this._h) { lib().destroy(this._h); this._h = null; }
};

FluidSolver.AIR = AIR; FluidSolver.FLUID = FLUID; FluidSolver.SOLID = SOLID;
FluidSolver.STAGE = STAGE;
module.exports = FluidSolver;
