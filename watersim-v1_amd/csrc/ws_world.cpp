// Planet world construction on the host (one-time setup, not on the
// substep path): seeded value-noise terrain on a voxel lattice at dx/2
// (generateTerrain, js/solver.js:236-508), the static SOLID/AIR raster
// (_rasterizeStatic, :865-916), the radial gravity faces
// (_buildPlanetGravity, :1177-1210) and the seeded jittered water fill
// (_resetWaterSphere, :651-731). Steps and constants follow the reference so
// the same seed builds the same planet; the voxel loop runs on threads.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/watersim_world.h"

namespace {

enum : uint8_t { AIR = 0, SOLID = 2 };

// Seeded 3-D value noise: LCG-shuffled permutation, quintic fade,
// trilinear blend of lattice hashes (solver.js:236-262).
struct ValueNoise {
  uint8_t perm[512];
  explicit ValueNoise(uint32_t seed) {
    uint8_t base[256];
    for (int i = 0; i < 256; i++) base[i] = (uint8_t)i;
    uint32_t s = seed;
    for (int i = 255; i > 0; i--) {
      s = s * 1664525u + 1013904223u;
      int j = (int)((s >> 8) % (uint32_t)(i + 1));
      uint8_t t = base[i];
      base[i] = base[j];
      base[j] = t;
    }
    for (int i = 0; i < 512; i++) perm[i] = base[i & 255];
  }
  double h(int ix, int iy, int iz) const {
    return perm[(perm[(perm[ix & 255] + iy) & 255] + iz) & 255] / 255.0;
  }
  static double fade(double u) { return u * u * u * (u * (u * 6 - 15) + 10); }
  double operator()(double x, double y, double z) const {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    double u = fade(x - fx), v = fade(y - fy), w = fade(z - fz);
    double a = h(ix, iy, iz), b = h(ix + 1, iy, iz), c = h(ix, iy + 1, iz), d = h(ix + 1, iy + 1, iz);
    double e = h(ix, iy, iz + 1), f = h(ix + 1, iy, iz + 1), g = h(ix, iy + 1, iz + 1), hh = h(ix + 1, iy + 1, iz + 1);
    double x1 = a + (b - a) * u, x2 = c + (d - c) * u, x3 = e + (f - e) * u, x4 = g + (hh - g) * u;
    double y1 = x1 + (x2 - x1) * v, y2 = x3 + (x4 - x3) * v;
    return y1 + (y2 - y1) * w;
  }
};

struct Lcg {
  uint32_t s;
  double next() {
    s = s * 1664525u + 1013904223u;
    return s / 4294967296.0;
  }
};

// warped fbm continents + ridged belts + detail (solver.js:292-313)
double raw_height(const ValueNoise &nf, double bump, double dx, double dy, double dz) {
  double wx = nf(dx * 0.9 + 17.3, dy * 0.9 + 9.1, dz * 0.9 + 4.7) - 0.5;
  double wy = nf(dx * 0.9 + 71.9, dy * 0.9 + 33.7, dz * 0.9 + 21.3) - 0.5;
  double wz = nf(dx * 0.9 + 123.4, dy * 0.9 + 87.2, dz * 0.9 + 55.8) - 0.5;
  double qx = dx + 0.35 * wx, qy = dy + 0.35 * wy, qz = dz + 0.35 * wz;
  double h = 0, amp = 1, fr = 1.15, tot = 0;
  for (int o = 0; o < 4; o++) {
    h += amp * nf(qx * fr + 3.1, qy * fr + 7.7, qz * fr + 11.9);
    tot += amp;
    amp *= 0.5;
    fr *= 2.1;
  }
  h /= tot;
  double r = 0;
  amp = 1; fr = 2.6; tot = 0;
  for (int o = 0; o < 3; o++) {
    double v = nf(qx * fr + 41.2, qy * fr + 23.8, qz * fr + 66.1);
    r += amp * (1 - std::fabs(2 * v - 1));
    tot += amp;
    amp *= 0.55;
    fr *= 2.3;
  }
  r /= tot;
  double det = nf(qx * 6.5 + 91.7, qy * 6.5 + 12.4, qz * 6.5 + 78.3);
  return h + 0.45 * bump * (r - 0.5) + 0.14 * bump * (det - 0.5);
}

struct Terrain {
  double Rlo, Rhi, Rsl, span, Rsph, mW, cMin, cSpan, offs, bump;
  const ValueNoise *nf;
  double radius(double dx, double dy, double dz) const {
    if (mW <= 0) return Rsph;
    double hv = raw_height(*nf, bump, dx, dy, dz);
    double h01 = (hv - cMin) / cSpan + offs;
    if (h01 < 0) h01 = 0; else if (h01 > 1) h01 = 1;
    double rv = Rlo + std::pow(h01, 1.25) * span;
    return Rsph + (rv - Rsph) * mW;
  }
};

template <class F>
void parallel_for(int n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 32) nt = 32;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      for (int i = (int)t; i < n; i += (int)nt) f(i);
    });
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

// _buildPlanetGravity, solver.js:1177-1210: -oceanR^2 e / |e|^3 per face
// (e from the domain centre to the face centre), u | v | w back to back
int ws_planet_gravity(int nx, int ny, int nz, double dx, double oceanR, float *g) {
  if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || !g) return -1;
  const long uN = (long)(nx + 1) * ny * nz, vN = (long)nx * (ny + 1) * nz;
  const double cx = nx * dx / 2, cy = ny * dx / 2, cz = nz * dx / 2;
  const double g0 = oceanR * oceanR;
  auto r3f = [](double a, double b, double c) {
    double r3 = std::pow(a * a + b * b + c * c, 1.5);
    return r3 != 0 ? r3 : 1e-9;
  };
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      double ey = (j + 0.5) * dx - cy, ez = (k + 0.5) * dx - cz;
      for (int i = 0; i <= nx; i++) {
        double ex = i * dx - cx;
        g[((long)k * ny + j) * (nx + 1) + i] = (float)(-g0 * ex / r3f(ex, ey, ez));
      }
    }
  for (int k = 0; k < nz; k++)
    for (int j = 0; j <= ny; j++) {
      double ey = j * dx - cy, ez = (k + 0.5) * dx - cz;
      for (int i = 0; i < nx; i++) {
        double ex = (i + 0.5) * dx - cx;
        g[uN + ((long)k * (ny + 1) + j) * nx + i] = (float)(-g0 * ey / r3f(ex, ey, ez));
      }
    }
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j < ny; j++) {
      double ey = (j + 0.5) * dx - cy, ez = k * dx - cz;
      for (int i = 0; i < nx; i++) {
        double ex = (i + 0.5) * dx - cx;
        g[uN + vN + ((long)k * ny + j) * nx + i] = (float)(-g0 * ez / r3f(ex, ey, ez));
      }
    }
  return 0;
}

int ws_world_build(const ws_world_params *p, ws_world *w) {
  if (!p || !w || p->nx < 4 || p->ny < 4 || p->nz < 4 || !(p->dx > 0)) return -1;
  std::memset(w, 0, sizeof(*w));
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double dx = p->dx, W = nx * dx, H = ny * dx, D = nz * dx;
  const double cx = W / 2, cy = H / 2, cz = D / 2;
  const double R1 = p->coreR, Rsl = p->coreR + p->depth;
  const double bump = p->bumpiness;
  // relief scaled to the world (solver.js:280-288)
  double depthW = Rsl - R1 > 1e-4 ? Rsl - R1 : 1e-4;
  double Rlo = R1 * 0.55 > Rsl - 1.8889 * depthW ? R1 * 0.55 : Rsl - 1.8889 * depthW;
  double Rhi = Rsl + 0.7778 * depthW;
  double span = Rhi - Rlo;
  double mW = bump / 0.3 < 1 ? bump / 0.3 : 1;
  double Rsph = Rsl - (0.4 * depthW < 0.45 * (Rsl - Rlo) ? 0.4 * depthW : 0.45 * (Rsl - Rlo));
  ValueNoise nf(p->seed);
  // calibration over 8192 quasi-uniform directions (solver.js:317-341)
  const int CALN = 8192;
  std::vector<float> cal(CALN);
  Lcg cr{p->seed};
  double cMin = 1e9, cMax = -1e9;
  for (int ci = 0; ci < CALN; ci++) {
    double u1 = cr.next() * 2 - 1, th = cr.next() * 6.2831853;
    double sq = std::sqrt(1 - u1 * u1);
    double hv = raw_height(nf, bump, sq * std::cos(th), u1, sq * std::sin(th));
    cal[ci] = (float)hv;
    if (hv < cMin) cMin = hv; else if (hv > cMax) cMax = hv;
  }
  double cSpan = (cMax - cMin) != 0 ? (cMax - cMin) : 1;
  double hSl = (Rsl - Rlo) / span;
  double lo = -1, hi = 1, delta = 0;
  for (int it = 0; it < 48; it++) {
    delta = 0.5 * (lo + hi);
    int cnt = 0;
    for (int ci = 0; ci < CALN; ci++) {
      double h01 = (cal[ci] - cMin) / cSpan;
      double hs = h01 + delta;
      if (hs < 0) hs = 0; else if (hs > 1) hs = 1;
      if (std::pow(hs, 1.25) > hSl) cnt++;
    }
    if (cnt > 0.27 * CALN) hi = delta; else lo = delta;
  }
  Terrain T{Rlo, Rhi, Rsl, span, Rsph, mW, cMin, cSpan, delta, bump, &nf};
  // voxel lattice at dx/2 (solver.js:352-372)
  const double dv = dx / 2;
  const int n = (int)std::ceil(W / dv);
  const long n3 = (long)n * n * n, n2 = (long)n * n;
  uint8_t *solid = (uint8_t *)std::calloc(n3, 1);
  float *R = (float *)std::calloc(n3, sizeof(float));
  if (!solid || !R) { std::free(solid); std::free(R); return -2; }
  parallel_for(n, [&](int kv) {
    double pz0 = (kv + 0.5) * dv - cz;
    for (int jv = 0; jv < n; jv++) {
      double py0 = (jv + 0.5) * dv - cy;
      for (int iv = 0; iv < n; iv++) {
        double px0 = (iv + 0.5) * dv - cx;
        double rr = std::sqrt(px0 * px0 + py0 * py0 + pz0 * pz0);
        long idx = ((long)kv * n + jv) * n + iv;
        if (rr < 1e-6) { R[idx] = (float)Rlo; solid[idx] = 1; continue; }
        double rv = T.radius(px0 / rr, py0 / rr, pz0 / rr);
        R[idx] = (float)rv;
        if (rr <= rv) solid[idx] = 1;
      }
    }
  });
  // slope smoothing: 18-neighbour in-place passes, then 6-neighbour passes
  // with mean-radius preservation (solver.js:384-426); sequential because
  // each update reads already-updated neighbours
  int PASSES = n < 48 ? 1 : (n < 64 ? 2 : 3);
  double BLEND = n < 48 ? 0.15 : (n < 64 ? 0.35 : 0.5);
  for (int ps = 0; ps < PASSES; ps++)
    for (int kv = 1; kv < n - 1; kv++) {
      long zB = kv * n2, zL = (kv - 1) * n2, zU = (kv + 1) * n2;
      for (int jv = 1; jv < n - 1; jv++) {
        long yB = zB + (long)jv * n, yL = zL + (long)(jv - 1) * n, yU = zL + (long)(jv + 1) * n;
        long y2 = zU + (long)(jv - 1) * n, y3 = zU + (long)jv * n, y4 = zU + (long)(jv + 1) * n;
        long yL2 = zB + (long)(jv - 1) * n, yU2 = zB + (long)(jv + 1) * n;
        for (int iv = 1; iv < n - 1; iv++) {
          long iR = yB + iv;
          double avg = ((double)R[iR - 1] + R[iR + 1] + R[yL2 + iv] + R[yU2 + iv] + R[yL + iv] + R[y2 + iv] +
                        R[y3 + iv] + R[y4 + iv] + R[yL + iv - 1] + R[yL + iv + 1] + R[y2 + iv - 1] +
                        R[y2 + iv + 1] + R[yU + iv] + R[y3 + iv - 1] + R[y3 + iv + 1] + R[y4 + iv] +
                        R[yL2 + iv - 1] + R[yL2 + iv + 1]) *
                       (1.0 / 18);
          R[iR] = (float)(R[iR] + (avg - R[iR]) * BLEND);
        }
      }
    }
  double mean0 = 0, mean1 = 0;
  long cntR = 0;
  for (long idx = 0; idx < n3; idx += 7) { mean0 += R[idx]; cntR++; }
  mean0 /= cntR;
  for (int ps = 0; ps < PASSES; ps++)
    for (int kv = 1; kv < n - 1; kv++)
      for (int jv = 1; jv < n - 1; jv++) {
        long row = kv * n2 + (long)jv * n;
        for (int iv = 1; iv < n - 1; iv++) {
          long iS = row + iv;
          double avg2 = ((double)R[iS - 1] + R[iS + 1] + R[iS - n] + R[iS + n] + R[iS - n2] + R[iS + n2]) * (1.0 / 6);
          R[iS] = (float)(R[iS] + (avg2 - R[iS]) * BLEND);
        }
      }
  for (long idx = 0; idx < n3; idx += 7) mean1 += R[idx];
  mean1 /= cntR;
  if (mean1 > 1e-9) {
    double rScale = mean0 / mean1;
    for (long idx = 0; idx < n3; idx++) R[idx] = (float)(R[idx] * rScale);
  }
  parallel_for(n, [&](int kv) {
    double pz = (kv + 0.5) * dv - cz;
    for (int jv = 0; jv < n; jv++) {
      double py = (jv + 0.5) * dv - cy;
      for (int iv = 0; iv < n; iv++) {
        double px = (iv + 0.5) * dv - cx;
        long idx = ((long)kv * n + jv) * n + iv;
        solid[idx] = std::sqrt(px * px + py * py + pz * pz) <= R[idx] ? 1 : 0;
      }
    }
  });
  // land fraction by radial march over 4000 directions (solver.js:484-497)
  int land = 0;
  for (int ci = 0; ci < 4000; ci++) {
    double lu = cr.next() * 2 - 1, lth = cr.next() * 6.2831853;
    double lsq = std::sqrt(1 - lu * lu);
    double ldx = lsq * std::cos(lth), ldy = lu, ldz = lsq * std::sin(lth);
    for (double rr2 = Rhi + dv; rr2 > Rlo - dv; rr2 -= dv * 0.5) {
      int sx = (int)((cx + ldx * rr2) / dv), sy = (int)((cy + ldy * rr2) / dv), sz = (int)((cz + ldz * rr2) / dv);
      if (sx < 0 || sy < 0 || sz < 0 || sx >= n || sy >= n || sz >= n) continue;
      if (solid[((long)sz * n + sy) * n + sx]) {
        if (rr2 > Rsl) land++;
        break;
      }
    }
  }
  // static raster: shell + voxel at the cell centre (solver.js:872-900)
  long nC = (long)nx * ny * nz;
  uint8_t *st = (uint8_t *)std::malloc(nC);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        bool shell = j == 0 || j == ny - 1 || k == 0 || k == nz - 1 || i == 0 || i == nx - 1;
        int vx = (int)((i + 0.5) * dx / dv), vy = (int)((j + 0.5) * dx / dv), vz = (int)((k + 0.5) * dx / dv);
        vx = vx < 0 ? 0 : (vx >= n ? n - 1 : vx);
        vy = vy < 0 ? 0 : (vy >= n ? n - 1 : vy);
        vz = vz < 0 ? 0 : (vz >= n ? n - 1 : vz);
        bool rock = solid[((long)vz * n + vy) * n + vx] == 1;
        st[((long)k * ny + j) * nx + i] = (shell || rock) ? SOLID : AIR;
      }
  auto rock_cell = [&](double x, double y, double z) {
    int i = (int)(x / dx), j = (int)(y / dx), k = (int)(z / dx);
    i = i < 1 ? 1 : (i > nx - 2 ? nx - 2 : i);
    j = j < 1 ? 1 : (j > ny - 2 ? ny - 2 : j);
    k = k < 1 ? 1 : (k > nz - 2 ? nz - 2 : k);
    return st[((long)k * ny + j) * nx + i] == SOLID;
  };
  auto voxel_rock = [&](double x, double y, double z) {
    int i = (int)(x / dv), j = (int)(y / dv), k = (int)(z / dv);
    i = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    j = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
    k = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
    return solid[((long)k * n + j) * n + i] == 1;
  };
  // water fill (solver.js:651-718)
  const double R2 = Rsl;
  double Vshell = 4.0 / 3.0 * M_PI * (R2 * R2 * R2 - Rlo * Rlo * Rlo);
  int hits = 0;
  Lcg mc{p->seed ^ 0x9e3779b9u};
  for (int mi = 0; mi < 20000; mi++) {
    double u1 = mc.next() * 2 - 1, th = mc.next() * 6.2831853;
    double sq = std::sqrt(1 - u1 * u1);
    double rr = std::cbrt(Rlo * Rlo * Rlo + mc.next() * (R2 * R2 * R2 - Rlo * Rlo * Rlo));
    double qx = cx + sq * std::cos(th) * rr, qy = cy + u1 * rr, qz = cz + sq * std::sin(th) * rr;
    if (!rock_cell(qx, qy, qz)) hits++;
  }
  double Vw = Vshell * hits / 20000;
  double s = std::cbrt(Vw / p->targetParticles);
  double rHi = R2 - 0.4 * s;
  long cap = (long)p->targetParticles + 8192;
  std::vector<float> fx, fy, fz;
  fx.reserve(cap); fy.reserve(cap); fz.reserve(cap);
  Lcg fj{p->seed ^ 0x51ab3c77u};
  long np = 0;
  for (double y = cy - R2; y <= cy + R2; y += s)
    for (double z = cz - R2; z <= cz + R2; z += s)
      for (double x = cx - R2; x <= cx + R2; x += s) {
        if (np >= cap) break;
        double ex = x - cx, ey = y - cy, ez = z - cz;
        if (ex * ex + ey * ey + ez * ez > rHi * rHi) continue;
        double jx = x + (fj.next() - 0.5) * s * 0.3;
        double jy = y + (fj.next() - 0.5) * s * 0.3;
        double jz = z + (fj.next() - 0.5) * s * 0.3;
        if (rock_cell(jx, jy, jz) || voxel_rock(jx, jy, jz)) continue;
        fx.push_back((float)jx); fy.push_back((float)jy); fz.push_back((float)jz);
        np++;
      }
  // radial inverse-square gravity of unit strength (solver.js:1177-1210)
  long uN = (long)(nx + 1) * ny * nz, vN = (long)nx * (ny + 1) * nz, wN = (long)nx * ny * (nz + 1);
  float *g = (float *)std::malloc(sizeof(float) * (uN + vN + wN));
  ws_planet_gravity(nx, ny, nz, dx, R2, g);
  w->tn = n;
  w->dv = dv;
  w->solid = solid;
  w->R = R;
  w->staticTypes = st;
  w->gravity = g;
  w->nP = np;
  w->capacity = cap;
  w->px = (float *)std::malloc(sizeof(float) * (np ? np : 1));
  w->py = (float *)std::malloc(sizeof(float) * (np ? np : 1));
  w->pz = (float *)std::malloc(sizeof(float) * (np ? np : 1));
  std::memcpy(w->px, fx.data(), sizeof(float) * np);
  std::memcpy(w->py, fy.data(), sizeof(float) * np);
  std::memcpy(w->pz, fz.data(), sizeof(float) * np);
  w->spacing = s;
  w->particleVolume = s * s * s;
  w->oceanR = R2;
  w->coreR = R1;
  w->Rlo = Rlo;
  w->Rhi = Rhi;
  w->landFrac = land / 4000.0;
  return 0;
}

void ws_world_free(ws_world *w) {
  if (!w) return;
  std::free(w->solid);
  std::free(w->R);
  std::free(w->staticTypes);
  std::free(w->gravity);
  std::free(w->px);
  std::free(w->py);
  std::free(w->pz);
  std::memset(w, 0, sizeof(*w));
}

}  // extern "C"
