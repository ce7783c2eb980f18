/*
 * Metaballs.build / Metaballs.iso through the Node facade on the GPU, with
 * the checks of the reference's own test (test/metaball.test.js): a 16^3
 * solver stub holding a 7^3 block of liquid particles one spacing apart
 * (spacing = dx / 4) and an 11^3 block for the coarse / fine lattices.
 */
'use strict';
var path = require('path');
var Surface = require(path.join(__dirname, '..', '..', 'watersim-v1_amd', 'js', 'surface.js'));
var MB = Surface.Metaballs;

var failures = 0;
function check(name, ok, detail) {
  console.log((ok ? '  PASS ' : '  FAIL ') + name + (detail !== undefined ? '  [' + detail + ']' : ''));
  if (!ok) failures++;
}

var n = 16, dx = 1 / n, sp = dx * 0.25;
function block(side) {
  var c = (side - 1) / 2, px = [], py = [], pz = [];
  for (var k = 0; k < side; k++) for (var j = 0; j < side; j++) for (var i = 0; i < side; i++) {
    px.push(0.5 + (i - c) * sp); py.push(0.5 + (j - c) * sp); pz.push(0.5 + (k - c) * sp);
  }
  var m = px.length;
  return { nx: n, ny: n, nz: n, dx: dx, W: 1, H: 1, D: 1, spacing: sp, nP: m,
    px: Float32Array.from(px), py: Float32Array.from(py), pz: Float32Array.from(pz), pflag: new Int8Array(m) };
}
function meshOk(m) {
  if (!(m.count > 0 && m.count % 3 === 0)) return false;
  for (var v = 0; v < m.count; v++) {
    var p = 3 * v;
    if (!isFinite(m.pos[p]) || !isFinite(m.nrm[p])) return false;
    if (Math.abs(Math.hypot(m.nrm[p], m.nrm[p + 1], m.nrm[p + 2]) - 1) > 1e-5) return false;
  }
  return true;
}

var s7 = block(7);
var low = MB.build(s7, 0);
check('the metaball mesh has whole triangles with finite unit normals and encloses volume',
  meshOk(low) && low.volume > 0, low.count + ' verts');
var out = true;
for (var v = 0; v < low.count; v++) {
  var p = 3 * v;
  if ((low.pos[p] - 0.5) * low.nrm[p] + (low.pos[p + 1] - 0.5) * low.nrm[p + 1] +
      (low.pos[p + 2] - 0.5) * low.nrm[p + 2] <= -0.5) out = false;
}
check('normals point away from the blob centre', out);
var lowCount = low.count, again = MB.build(s7, 0);
check('deterministic build that reuses the mesher scratch', again.count === lowCount && again.pos === low.pos);

var prev = -1, grows = true;
[0, 0.25, 0.5, 0.75, 1].forEach(function (t) { var m = MB.build(s7, t); grows = grows && m.volume > prev; prev = m.volume; });
check('volume grows with tension', grows);
check('the contour level rises with tension and clamps', MB.iso(0) < MB.iso(0.5) && MB.iso(0.5) < MB.iso(1) &&
  MB.iso(1) === MB.iso(2) && MB.iso(0) === MB.iso(-1));

var omitted = MB.build(s7, 0.5), oc = omitted.count, ov = omitted.volume;
var same = true;
[1, undefined, NaN, 0, -2, null, 'x'].forEach(function (bad) {
  var m = MB.build(s7, 0.5, bad);
  same = same && m.count === oc && m.volume === ov;
});
check('detail 1, missing, NaN, non-positive or non-number builds the solver lattice', same);

var s11 = block(11);
var body1 = MB.build(s11, 1, 1), c1 = body1.count, v1 = body1.volume;
var coarse = MB.build(s11, 1, 0.6), cc = coarse.count, cv = coarse.volume;
var fine = MB.build(s11, 1, 1.4);
check('a coarse lattice marches fewer vertices and keeps the volume within 15 %',
  cc > 0 && cc < c1 && Math.abs(cv - v1) / v1 < 0.15, cc + ' < ' + c1 + ', ' + (100 * cv / v1).toFixed(1) + ' %');
check('a fine lattice adds vertices and keeps the volume within 15 %',
  fine.count > c1 && Math.abs(fine.volume - v1) / v1 < 0.15 && meshOk(fine),
  fine.count + ' > ' + c1 + ', ' + (100 * fine.volume / v1).toFixed(1) + ' %');

var hi = MB.build(s7, 0.5, 999), hc = hi.count, hv = hi.volume, atMax = MB.build(s7, 0.5, 1.6);
var lo = MB.build(s7, 0.5, 0.01), lc = lo.count, lv = lo.volume, atMin = MB.build(s7, 0.5, 0.5);
check('absurd details clamp to 0.5 .. 1.6', hc === atMax.count && hv === atMax.volume &&
  lc === atMin.count && lv === atMin.volume);
prev = -1; grows = true;
[0, 0.25, 0.5, 0.75, 1].forEach(function (t) { var m = MB.build(s7, t, 1.4); grows = grows && m.volume > prev; prev = m.volume; });
check('volume grows with tension on a finer lattice too', grows);

var mixed = block(7);
for (var q = 0; q < mixed.nP; q += 3) mixed.pflag[q] = 2;
var kept = block(7), keep = [];
for (q = 0; q < kept.nP; q++) if (q % 3) keep.push(q);
var only = { nx: n, ny: n, nz: n, dx: dx, spacing: sp, nP: keep.length,
  px: Float32Array.from(keep.map(function (i) { return kept.px[i]; })),
  py: Float32Array.from(keep.map(function (i) { return kept.py[i]; })),
  pz: Float32Array.from(keep.map(function (i) { return kept.pz[i]; })), pflag: new Int8Array(keep.length) };
var a = MB.build(mixed, 0.5), ac = a.count, av = a.volume, b = MB.build(only, 0.5);
check('non-liquid particles do not enter the field', ac === b.count && av === b.volume);
var empty = MB.build({ nx: n, ny: n, nz: n, dx: dx, spacing: sp, nP: 0, px: [], py: [], pz: [], pflag: [] }, 0.5);
check('an empty solver gives an empty mesh', empty.count === 0 && empty.volume === 0);

console.log(failures ? failures + ' FAILED' : 'ALL METABALL CHECKS PASSED');
process.exitCode = failures ? 1 : 0;
