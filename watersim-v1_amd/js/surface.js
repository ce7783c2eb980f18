/*
 * MarchingTetrahedra.build(dens, nx, ny, nz, dx, iso) of the reference
 * (js/surface.js:66-177, export :286-291) on the GPU: the field is uploaded,
 * triangulated by the HIP kernels and the mesh read back into module-level
 * grow-only buffers, which are reused across calls as in the reference
 * (js/surface.js:53-64, capacity doubling from 32768 vertices).
 */
'use strict';

var path = require('path');

var addon = null, ctx = null;
function handle() {
  if (!ctx) {
    var where = path.join(__dirname, '..', '_build', 'watersim_napi.node');
    try {
      addon = require(where);
    } catch (e) {
      throw new Error('watersim: the HIP addon is not built (' + where + '): ' + e.message);
    }
    // a minimal context: the field and its dimensions come with each call
    ctx = addon.create(4, 4, 4, 1.0, 1);
  }
  return ctx;
}

var cap = 0, pos = new Float32Array(0), nrm = new Float32Array(0);

function ensure(n) {
  if (n <= cap) return;
  var c = Math.max(cap, 32768);
  while (c < n) c *= 2;
  cap = c;
  pos = new Float32Array(cap * 3);
  nrm = new Float32Array(cap * 3);
}

function build(dens, nx, ny, nz, dx, iso) {
  var h = handle();
  var f = dens instanceof Float32Array ? dens : Float32Array.from(dens);
  ensure(1);
  var r = addon.mtBuild(h, f, nx, ny, nz, dx, iso, pos, nrm);
  if (r.count > cap) {
    ensure(r.count);
    r = addon.mtBuild(h, f, nx, ny, nz, dx, iso, pos, nrm);
  }
  return { count: r.count, volume: r.volume, pos: pos, nrm: nrm };
}

// ------------------------------------------------------------------ metaballs
// Metaballs.build(solver, tension, detail) and Metaballs.iso(tension),
// js/surface.js:178-284. A FluidSolver of this package builds the skin from
// its own device state; any other object with the reference solver's fields
// (nx, dx, spacing, nP, px, py, pz, pflag) is copied into a scratch context.
var MB_R_MIN = 1.6, MB_R_SPAN = 2.2, MB_ISO0 = 1.5, MB_ISO_ALPHA = 1.5;

function metaballIso(tension) {
  if (!isFinite(tension)) tension = 0.5;
  tension = Math.max(0, Math.min(1, tension));
  return MB_ISO0 * Math.pow((MB_R_MIN + MB_R_SPAN * tension) / MB_R_MIN, MB_ISO_ALPHA);
}

function buildMetaballs(solver, tension, detail) {
  // the reference's argument rules: a non-finite tension is 0.5 (after JS
  // number coercion), a non-number or non-positive detail is factor 1
  var t = isFinite(tension) ? Math.max(0, Math.min(1, tension)) : 0.5;
  var d = typeof detail === 'number' && isFinite(detail) && detail > 0 ? detail : NaN;
  var h, scratch = null;
  if (solver && solver._h && typeof solver._call === 'function') {
    solver._call(function () {});   // flush host-side edits of the particle arrays
    h = solver._h;
  } else {
    handle();
    var n = solver.nP | 0;
    scratch = addon.create(solver.nx, solver.ny || solver.nx, solver.nz || solver.nx, solver.dx, Math.max(n, 1));
    addon.setF(scratch, 'spacing', solver.spacing);
    addon.setF(scratch, 'nP', n);
    if (n > 0) {
      addon.upload(scratch, 'px', Float32Array.from(solver.px.subarray ? solver.px.subarray(0, n) : solver.px));
      addon.upload(scratch, 'py', Float32Array.from(solver.py.subarray ? solver.py.subarray(0, n) : solver.py));
      addon.upload(scratch, 'pz', Float32Array.from(solver.pz.subarray ? solver.pz.subarray(0, n) : solver.pz));
      var fl = new Uint8Array(n);
      for (var p = 0; p < n; p++) fl[p] = solver.pflag[p];
      addon.upload(scratch, 'pflag', fl);
    }
    h = scratch;
  }
  try {
    ensure(1);
    var r = addon.metaballs(h, t, d, pos, nrm);
    if (r.count > cap) {
      ensure(r.count);
      r = addon.metaballs(h, t, d, pos, nrm);
    }
    return { count: r.count, volume: r.volume, pos: pos, nrm: nrm };
  } finally {
    if (scratch) addon.destroy(scratch);
  }
}

module.exports = { build: build, Metaballs: { build: buildMetaballs, iso: metaballIso } };
