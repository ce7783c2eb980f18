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

module.exports = { build: build };
