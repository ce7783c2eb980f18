// Balls of the planet world (js/solver.js:918-938, 1005-1026, 2294-2564):
// rigid spheres that the grid sees as moving solids, that push particles
// out of their volume, and that float, sink, roll on the rock and collide.
//
// The reference keeps each ball as an object of JavaScript numbers and does
// all of this in double precision. The state lives on the device as nine
// doubles per ball, and every ball formula is evaluated in FP64 in the
// reference's operation order; this file is compiled without FMA
// contraction (-ffp-contract=off), and FP64 division and square root are
// IEEE correctly rounded on gfx950, so the results are the reference's bits.
//
// Work split:
//   raster      one thread per cell, all balls in order (a later ball's
//               velocity wins), clearing the solid velocities in the same pass
//   push-out    one thread per particle, all balls in order
//   histogram   the waterline histogram of every ball in one particle pass
//               (integer counts, LDS then global atomics: order-free)
//   integrator  one block per ball (the balls are independent until their
//               mutual contacts): thread 0 runs the force law and the
//               integration, the block scans the terrain voxels under the
//               ball for the deepest penetration (a max reduction that keeps
//               the first voxel in scan order, as the serial scan does)
//   contacts    the ball-ball pairs in order, one thread
#include <climits>

#include <hipcub/hipcub.hpp>

#include "ws_frame.h"

namespace ws {

namespace {

constexpr double kJsPi = 3.141592653589793;  // Math.PI
constexpr int kHist = 24;            // waterline buckets (solver.js:2360)
constexpr int kHistRow = kHist + 1;  // buckets + sample count per ball

struct BallEnv {
  int nx, ny, nz, tn;
  double dx, cx, cy, cz, coreR, oceanR, domainR, gravity, dv;
};

BallEnv ball_env(ws_ctx *c) {
  BallEnv e;
  e.nx = c->nx; e.ny = c->ny; e.nz = c->nz; e.tn = c->tn;
  e.dx = c->dx; e.cx = c->cx; e.cy = c->cy; e.cz = c->cz;
  e.coreR = c->coreR; e.oceanR = c->oceanR; e.domainR = c->domainR; e.gravity = c->gravity; e.dv = c->dv;
  return e;
}

// ------------------------------------------------------------------ raster
struct RasterBox {
  int i0, i1, j0, j1, k0, k1;
  double x, y, z, rr;
  float vx, vy, vz;
};

// moving solids of _rasterizeStatic (solver.js:918-938): every cell centre
// within r + 0.4 dx of a ball (inside the ball's clamped bounding box)
// becomes SOLID with the ball's velocity; all other cells get velocity 0
__global__ __launch_bounds__(256) void k_ball_raster(int nx, int ny, int nz, double dx, int nb,
                                                     const double *__restrict__ balls, uint8_t *__restrict__ ct,
                                                     float *__restrict__ sv) {
  __shared__ RasterBox box[WS_MAX_BALLS];
  if ((int)threadIdx.x < nb) {
    const double *b = balls + 9 * threadIdx.x;
    double r = b[6];
    RasterBox B;
    B.i0 = clampi((int)floor((b[0] - r) / dx), 1, nx - 2);
    B.i1 = clampi((int)ceil((b[0] + r) / dx), 1, nx - 2);
    B.j0 = clampi((int)floor((b[1] - r) / dx), 1, ny - 2);
    B.j1 = clampi((int)ceil((b[1] + r) / dx), 1, ny - 2);
    B.k0 = clampi((int)floor((b[2] - r) / dx), 1, nz - 2);
    B.k1 = clampi((int)ceil((b[2] + r) / dx), 1, nz - 2);
    B.x = b[0]; B.y = b[1]; B.z = b[2];
    B.rr = (r + 0.4 * dx) * (r + 0.4 * dx);
    B.vx = (float)b[3]; B.vy = (float)b[4]; B.vz = (float)b[5];
    box[threadIdx.x] = B;
  }
  __syncthreads();
  long nC = (long)nx * ny * nz;
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
  bool hit = false;
  float vx = 0.f, vy = 0.f, vz = 0.f;
  for (int n = 0; n < nb; n++) {
    const RasterBox &B = box[n];
    if (i < B.i0 || i > B.i1 || j < B.j0 || j > B.j1 || k < B.k0 || k > B.k1) continue;
    double ex = (i + 0.5) * dx - B.x, ey = (j + 0.5) * dx - B.y, ez = (k + 0.5) * dx - B.z;
    if (ex * ex + ey * ey + ez * ez <= B.rr) {
      hit = true;
      vx = B.vx; vy = B.vy; vz = B.vz;
    }
  }
  if (hit) ct[c] = SOLID;
  sv[c] = vx;
  sv[nC + c] = vy;
  sv[2 * nC + c] = vz;
}

// ---------------------------------------------------------------- push-out
// _pushBallsChunk (solver.js:1009-1026): a particle (any flag) closer than
// r + 0.002 to a ball centre moves to that distance along the normal and
// loses its inward normal velocity relative to the ball; balls in order
__global__ __launch_bounds__(256) void k_push_balls(int nP, int nb, const double *__restrict__ balls,
                                                    float *__restrict__ px, float *__restrict__ py,
                                                    float *__restrict__ pz, float *__restrict__ pvx,
                                                    float *__restrict__ pvy, float *__restrict__ pvz) {
  __shared__ double sb[WS_MAX_BALLS * 9];
  for (int e = threadIdx.x; e < nb * 9; e += blockDim.x) sb[e] = balls[e];
  __syncthreads();
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  float x = px[p], y = py[p], z = pz[p];
  float vx = pvx[p], vy = pvy[p], vz = pvz[p];
  bool moved = false;
  for (int n = 0; n < nb; n++) {
    const double *b = sb + 9 * n;
    double rr = b[6] + 0.002;
    double ex = (double)x - b[0], ey = (double)y - b[1], ez = (double)z - b[2];
    double d2 = ex * ex + ey * ey + ez * ez;
    if (d2 < rr * rr) {
      double d = sqrt(d2);
      if (d == 0) d = 1e-9;
      double inv = 1 / d;
      double nxn = ex * inv, nyn = ey * inv, nzn = ez * inv;
      x = (float)(b[0] + nxn * rr);
      y = (float)(b[1] + nyn * rr);
      z = (float)(b[2] + nzn * rr);
      double rvx = (double)vx - b[3], rvy = (double)vy - b[4], rvz = (double)vz - b[5];
      double vn = rvx * nxn + rvy * nyn + rvz * nzn;
      if (vn < 0) {
        vx = (float)((double)vx - nxn * vn);
        vy = (float)((double)vy - nyn * vn);
        vz = (float)((double)vz - nzn * vn);
      }
      moved = true;
    }
  }
  if (moved) {
    px[p] = x; py[p] = y; pz[p] = z;
    pvx[p] = vx; pvy[p] = vy; pvz[p] = vz;
  }
}

// the particle bins of the marking pass, rebuilt from the pushed positions
// (the P2G kernel bins particles by the cell of their current position)
__global__ void k_rebin(Geo g, World w, int nP, const float *__restrict__ px, const float *__restrict__ py,
                        const float *__restrict__ pz, int *__restrict__ pcell, unsigned *__restrict__ prank,
                        unsigned *__restrict__ cnt) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP || pcell[p] < 0) return;
  int i = clampi(cell_floor(px[p], g.inv_dx, w.thr, g.nx), 0, g.nx - 1);
  int j = clampi(cell_floor(py[p], g.inv_dx, w.thr, g.ny), 0, g.ny - 1);
  int k = clampi(cell_floor(pz[p], g.inv_dx, w.thr, g.nz), 0, g.nz - 1);
  int c = (k * g.ny + j) * g.nx + i;
  pcell[p] = c;
  prank[p] = atomicAdd(&cnt[c], 1u);
}

__global__ void k_zero_words(unsigned *__restrict__ p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// --------------------------------------------------------------- histogram
// the waterline histogram of _updateBalls (solver.js:2357-2407, sphere
// mode): liquid particles (flag 0) in the shell 2.2 r .. 3.2 r around a
// ball count in the bucket of their radius over [coreR, coreR + hSpan)
__global__ __launch_bounds__(256) void k_ball_hist(int nP, int nb, const double *__restrict__ balls, double cx,
                                                   double cy, double cz, double hLo, double hSpan,
                                                   const float *__restrict__ px, const float *__restrict__ py,
                                                   const float *__restrict__ pz, const uint8_t *__restrict__ pflag,
                                                   unsigned *__restrict__ hist) {
  __shared__ double sb[WS_MAX_BALLS * 4];
  __shared__ unsigned sh[WS_MAX_BALLS * kHistRow];
  for (int e = threadIdx.x; e < nb * 4; e += blockDim.x) sb[e] = balls[9 * (e / 4) + (e % 4 == 3 ? 6 : e % 4)];
  for (int e = threadIdx.x; e < nb * kHistRow; e += blockDim.x) sh[e] = 0;
  __syncthreads();
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nP; p += gridDim.x * blockDim.x) {
    if (pflag[p] != 0) continue;
    double ppx = px[p], ppy = py[p], ppz = pz[p];
    double ldx = ppx - cx, ldy = ppy - cy, ldz = ppz - cz;
    double level = sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
    for (int n = 0; n < nb; n++) {
      const double *b = sb + 4 * n;
      double bdx = ppx - b[0], bdy = ppy - b[1], bdz = ppz - b[2];
      double dh2 = bdx * bdx + bdy * bdy + bdz * bdz;
      double r1 = 2.2 * b[3], r2 = 3.2 * b[3];
      if (dh2 < r1 * r1 || dh2 > r2 * r2) continue;
      double fb = floor(((level - hLo) / hSpan) * kHist);
      int bucket = fb < 0 ? 0 : (fb >= kHist ? kHist - 1 : (int)fb);
      atomicAdd(&sh[n * kHistRow + bucket], 1u);
      atomicAdd(&sh[n * kHistRow + kHist], 1u);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nb * kHistRow; e += blockDim.x)
    if (sh[e]) atomicAdd(&hist[e], sh[e]);
}

// ------------------------------------------------------------- integrator
// one face grid sampled at a double point: _sampleVel3 (solver.js:820-854)
__device__ double sample_face_d(const BallEnv &e, const float *__restrict__ A, int comp, double x, double y,
                                double z) {
  double dx = e.dx, a, b, cc, tx, ty, tz;
  int nx = e.nx, ny = e.ny, nz = e.nz;
  if (comp == 0) { a = x / dx; b = y / dx - 0.5; cc = z / dx - 0.5; }
  else if (comp == 1) { a = x / dx - 0.5; b = y / dx; cc = z / dx - 0.5; }
  else { a = x / dx - 0.5; b = y / dx - 0.5; cc = z / dx; }
  int i0 = (int)floor(a), j0 = (int)floor(b), k0 = (int)floor(cc);
  if (i0 < 0) { i0 = 0; tx = 0; } else if (i0 > nx - 1) { i0 = nx - 1; tx = 1; } else tx = a - i0;
  if (j0 < 0) { j0 = 0; ty = 0; } else if (j0 > ny - 2) { j0 = ny - 2; ty = 1; } else ty = b - j0;
  if (k0 < 0) { k0 = 0; tz = 0; } else if (k0 > nz - 2) { k0 = nz - 2; tz = 1; } else tz = cc - k0;
  long sj, sk, idx;
  if (comp == 0) { sj = nx + 1; sk = (long)ny * (nx + 1); idx = ((long)k0 * ny + j0) * (nx + 1) + i0; }
  else if (comp == 1) { sj = nx; sk = (long)(ny + 1) * nx; idx = ((long)k0 * (ny + 1) + j0) * nx + i0; }
  else { sj = nx; sk = (long)ny * nx; idx = ((long)k0 * ny + j0) * nx + i0; }
  double s0 = 1 - tx, s1 = tx, t0 = 1 - ty, t1 = ty, r0 = 1 - tz, r1 = tz;
  double v00 = A[idx] * s0 + A[idx + 1] * s1;
  double v10 = A[idx + sj] * s0 + A[idx + sj + 1] * s1;
  double v01 = A[idx + sk] * s0 + A[idx + sk + 1] * s1;
  double v11 = A[idx + sk + sj] * s0 + A[idx + sk + sj + 1] * s1;
  return (v00 * t0 + v10 * t1) * r0 + (v01 * t0 + v11 * t1) * r1;
}

// the closest point of terrain voxel (i, j, k) to the ball centre and the
// penetration depth (solver.js:2311-2323)
__device__ __forceinline__ double voxel_pen(double bx, double by, double bz, double r, int i, int j, int k,
                                            double dv, double &ddx, double &ddy, double &ddz, double &dd) {
  double qx = bx, qy = by, qz = bz;
  double vx0 = i * dv, vy0 = j * dv, vz0 = k * dv;
  if (qx < vx0) qx = vx0; else if (qx > vx0 + dv) qx = vx0 + dv;
  if (qy < vy0) qy = vy0; else if (qy > vy0 + dv) qy = vy0 + dv;
  if (qz < vz0) qz = vz0; else if (qz > vz0 + dv) qz = vz0 + dv;
  ddx = bx - qx; ddy = by - qy; ddz = bz - qz;
  dd = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
  return r - dd;
}

// _updateBalls for one ball per block (solver.js:2408-2541): the waterline
// from the histogram, the submerged cap, buoyancy and weight under the
// inverse-square law, drag against the grid velocity (clamped), added mass,
// damping and the 4 m/s clamp, the explicit step, then _ballVsTerrain
// (:2299-2342, or the bare core sphere) and the domain sphere
__global__ __launch_bounds__(256) void k_ball_update(BallEnv e, int nb, double *__restrict__ balls,
                                                     const unsigned *__restrict__ hist, float *__restrict__ surfOut,
                                                     const float *__restrict__ vel, const uint8_t *__restrict__ tsolid,
                                                     const DevScalars *__restrict__ s) {
  __shared__ double B[9];
  __shared__ int box[6];
  __shared__ double sPen[256];
  __shared__ int sIdx[256];
  const int n = blockIdx.x;
  if (n >= nb) return;
  const int t = threadIdx.x;
  if (t == 0) {
    const double rhoF = 1000, Cd = 0.5;
    double dt = s->dtD;
    double b[9];
    for (int q = 0; q < 9; q++) b[q] = balls[9 * n + q];
    // waterline (solver.js:2408-2430)
    const unsigned *h = hist + n * kHistRow;
    int total = (int)h[kHist];
    double hLo = e.coreR, hSpan = (e.oceanR - e.coreR) * 1.5;
    double wl = -INFINITY;
    if (total > 20) {
      double acc = 0, target = total * 0.10;
      for (int bkt = kHist - 1; bkt >= 0; bkt--) {
        acc += (double)h[bkt];
        if (acc >= target) { wl = hLo + (bkt + 0.5) / kHist * hSpan; break; }
      }
      if (wl == -INFINITY || total < 140) wl = e.oceanR;
      double wlMax = e.oceanR + (e.oceanR - e.coreR) * 0.3;
      if (wl > wlMax) wl = wlMax;
      if (wl < e.coreR) wl = e.coreR;
    } else {
      wl = e.oceanR;
    }
    float surfF = (float)wl;  // _surfY is a Float32Array
    surfOut[n] = surfF;
    // forces and the step (solver.js:2434-2504)
    double x = b[0], y = b[1], z = b[2], vx = b[3], vy = b[4], vz = b[5], r = b[6], rho = b[7];
    double m = rho * 4 / 3 * kJsPi * r * r * r;
    double area = kJsPi * r * r;
    double surf = (double)surfF;
    double bR = sqrt((x - e.cx) * (x - e.cx) + (y - e.cy) * (y - e.cy) + (z - e.cz) * (z - e.cz));
    double inner = bR - r;
    double hh = surf - inner;
    if (!(hh > 0)) hh = 0;
    if (hh > 2 * r) hh = 2 * r;
    double Vs = kJsPi * hh * hh * (3 * r - hh) / 3;
    b[8] = Vs / (4.0 / 3 * kJsPi * r * r * r);
    double upx = 0, upy = 1, upz = 0, gEff = e.gravity;
    if (bR > 1e-6) {
      upx = (x - e.cx) / bR; upy = (y - e.cy) / bR; upz = (z - e.cz) / bR;
      gEff = e.gravity * (e.oceanR * e.oceanR) / (bR * bR);
    }
    double Fb = (rhoF * Vs - m) * gEff;
    double fu = sample_face_d(e, vel, 0, x, y, z);
    double fv = sample_face_d(e, vel + (long)(e.nx + 1) * e.ny * e.nz, 1, x, y, z);
    double fw = sample_face_d(e, vel + (long)(e.nx + 1) * e.ny * e.nz + (long)e.nx * (e.ny + 1) * e.nz, 2, x, y, z);
    double rvx = vx - fu, rvy = vy - fv, rvz = vz - fw;
    double vr = sqrt(rvx * rvx + rvy * rvy + rvz * rvz);
    double kd = 0.5 * rhoF * Cd * area * vr;
    double Fx = -kd * rvx + upx * Fb, Fz = -kd * rvz + upz * Fb, Fyd = -kd * rvy + upy * Fb;
    {
      double fdx = Fx - upx * Fb, fdy = Fyd - upy * Fb, fdz = Fz - upz * Fb;
      double fdm = sqrt(fdx * fdx + fdy * fdy + fdz * fdz);
      double fdmMax = 0.8 * m * e.gravity;
      if (fdm > fdmMax) {
        double fdc = fdmMax / fdm;
        Fx = upx * Fb + fdx * fdc; Fyd = upy * Fb + fdy * fdc; Fz = upz * Fb + fdz * fdc;
      }
    }
    double mEff = m + 0.5 * rhoF * Vs;
    vx += (Fx / mEff) * dt;
    vz += (Fz / mEff) * dt;
    vy += (Fyd / mEff) * dt;
    vx *= 0.999; vy *= 0.999; vz *= 0.999;
    const double bvmax = 4;
    if (vx > bvmax) vx = bvmax; else if (vx < -bvmax) vx = -bvmax;
    if (vz > bvmax) vz = bvmax; else if (vz < -bvmax) vz = -bvmax;
    if (vy > bvmax) vy = bvmax; else if (vy < -bvmax) vy = -bvmax;
    x += vx * dt; y += vy * dt; z += vz * dt;
    b[0] = x; b[1] = y; b[2] = z; b[3] = vx; b[4] = vy; b[5] = vz;
    for (int q = 0; q < 9; q++) B[q] = b[q];
    if (tsolid) {
      // the voxel box of _ballVsTerrain: (coord / dv) | 0, clamped
      int i0 = (int)((x - r) / e.dv), i1 = (int)((x + r) / e.dv);
      int j0 = (int)((y - r) / e.dv), j1 = (int)((y + r) / e.dv);
      int k0 = (int)((z - r) / e.dv), k1 = (int)((z + r) / e.dv);
      if (i0 < 0) i0 = 0;
      if (j0 < 0) j0 = 0;
      if (k0 < 0) k0 = 0;
      if (i1 > e.tn - 1) i1 = e.tn - 1;
      if (j1 > e.tn - 1) j1 = e.tn - 1;
      if (k1 > e.tn - 1) k1 = e.tn - 1;
      box[0] = i0; box[1] = i1; box[2] = j0; box[3] = j1; box[4] = k0; box[5] = k1;
    }
  }
  __syncthreads();
  if (tsolid) {
    int ni = box[1] - box[0] + 1, nj = box[3] - box[2] + 1, nk = box[5] - box[4] + 1;
    long cnt = (ni > 0 && nj > 0 && nk > 0) ? (long)ni * nj * nk : 0;
    double bx = B[0], by = B[1], bz = B[2], r = B[6];
    double bestPen = 0;
    int bestIdx = INT_MAX;
    for (long q = t; q < cnt; q += blockDim.x) {
      int i = box[0] + (int)(q % ni), j = box[2] + (int)((q / ni) % nj), k = box[4] + (int)(q / ((long)ni * nj));
      if (!tsolid[((long)k * e.tn + j) * e.tn + i]) continue;
      double ddx, ddy, ddz, dd;
      double pen = voxel_pen(bx, by, bz, r, i, j, k, e.dv, ddx, ddy, ddz, dd);
      if (pen > bestPen) { bestPen = pen; bestIdx = (int)q; }
    }
    sPen[t] = bestPen;
    sIdx[t] = bestIdx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (t < w) {
        double po = sPen[t + w];
        int io = sIdx[t + w];
        if (po > sPen[t] || (po == sPen[t] && io < sIdx[t])) { sPen[t] = po; sIdx[t] = io; }
      }
      __syncthreads();
    }
    if (t == 0 && sIdx[0] != INT_MAX) {
      int q = sIdx[0];
      int i = box[0] + q % ni, j = box[2] + (q / ni) % nj, k = box[4] + q / (ni * nj);
      double ddx, ddy, ddz, dd;
      double pen = voxel_pen(bx, by, bz, r, i, j, k, e.dv, ddx, ddy, ddz, dd);
      double bpx, bpy, bpz;
      if (dd > 1e-6) {
        bpx = ddx / dd; bpy = ddy / dd; bpz = ddz / dd;
      } else {
        double rx = bx - e.cx, ry = by - e.cy, rz = bz - e.cz;
        double rl = sqrt(rx * rx + ry * ry + rz * rz);
        if (rl == 0) rl = 1e-9;
        bpx = rx / rl; bpy = ry / rl; bpz = rz / rl;
      }
      B[0] = bx + bpx * (pen + 0.002);
      B[1] = by + bpy * (pen + 0.002);
      B[2] = bz + bpz * (pen + 0.002);
      double vx = B[3], vy = B[4], vz = B[5];
      double vn = vx * bpx + vy * bpy + vz * bpz;
      if (vn < 0) {
        vx -= bpx * vn; vy -= bpy * vn; vz -= bpz * vn;
        vx *= 0.985; vy *= 0.985; vz *= 0.985;
      }
      B[3] = vx; B[4] = vy; B[5] = vz;
    }
  } else if (t == 0) {
    // no terrain: the bare core sphere (solver.js:2508-2521)
    double cbx = B[0] - e.cx, cby = B[1] - e.cy, cbz = B[2] - e.cz;
    double cbr = sqrt(cbx * cbx + cby * cby + cbz * cbz);
    if (cbr == 0) cbr = 1e-9;
    if (cbr < e.coreR + B[6]) {
      double rnx0 = cbx / cbr, rny0 = cby / cbr, rnz0 = cbz / cbr;
      double rr0 = e.coreR + B[6];
      B[0] = e.cx + rnx0 * rr0; B[1] = e.cy + rny0 * rr0; B[2] = e.cz + rnz0 * rr0;
      double vn0 = B[3] * rnx0 + B[4] * rny0 + B[5] * rnz0;
      if (vn0 < 0) {
        B[3] -= rnx0 * vn0 * 1.3; B[4] -= rny0 * vn0 * 1.3; B[5] -= rnz0 * vn0 * 1.3;
        B[3] *= 0.98; B[4] *= 0.98; B[5] *= 0.98;
      }
    }
  }
  if (t == 0) {
    // the domain sphere (solver.js:2523-2532)
    double sbx = B[0] - e.cx, sby = B[1] - e.cy, sbz = B[2] - e.cz;
    double sbr = sqrt(sbx * sbx + sby * sby + sbz * sbz);
    if (sbr == 0) sbr = 1e-9;
    double rnx = sbx / sbr, rny = sby / sbr, rnz = sbz / sbr;
    if (sbr > e.domainR - B[6]) {
      double rr1 = e.domainR - B[6];
      B[0] = e.cx + rnx * rr1; B[1] = e.cy + rny * rr1; B[2] = e.cz + rnz * rr1;
      double vn1 = B[3] * rnx + B[4] * rny + B[5] * rnz;
      if (vn1 > 0) { B[3] -= rnx * vn1 * 1.3; B[4] -= rny * vn1 * 1.3; B[5] -= rnz * vn1 * 1.3; }
    }
    for (int q = 0; q < 9; q++) balls[9 * n + q] = B[q];
  }
}

// ball-ball contacts in pair order (solver.js:2543-2563)
__global__ void k_ball_pairs(int nb, double *__restrict__ balls) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int a = 0; a < nb; a++)
    for (int b2 = a + 1; b2 < nb; b2++) {
      double *A = balls + 9 * a, *Bb = balls + 9 * b2;
      double ex = Bb[0] - A[0], ey = Bb[1] - A[1], ez = Bb[2] - A[2];
      double d2 = ex * ex + ey * ey + ez * ez, rSum = A[6] + Bb[6];
      if (d2 > rSum * rSum || d2 < 1e-12) continue;
      double d = sqrt(d2), nxn = ex / d, nyn = ey / d, nzn = ez / d;
      double overlap = rSum - d;
      double mA = A[7] * A[6] * A[6] * A[6], mB = Bb[7] * Bb[6] * Bb[6] * Bb[6];
      double tA = mB / (mA + mB), tB = mA / (mA + mB);
      A[0] -= nxn * overlap * tA; A[1] -= nyn * overlap * tA; A[2] -= nzn * overlap * tA;
      Bb[0] += nxn * overlap * tB; Bb[1] += nyn * overlap * tB; Bb[2] += nzn * overlap * tB;
      double rvn = (Bb[3] - A[3]) * nxn + (Bb[4] - A[4]) * nyn + (Bb[5] - A[5]) * nzn;
      if (rvn < 0) {
        double jimp = -(1 + 0.3) * rvn / (1 / mA + 1 / mB);
        A[3] -= nxn * jimp / mA; A[4] -= nyn * jimp / mA; A[5] -= nzn * jimp / mA;
        Bb[3] += nxn * jimp / mB; Bb[4] += nyn * jimp / mB; Bb[5] += nzn * jimp / mB;
      }
    }
}

}  // namespace

// ------------------------------------------------------------------- host
int balls_raster(ws_ctx *c) {
  if (!c->nBalls && !c->hadSolidVelocity) return WS_OK;
  hipLaunchKernelGGL(k_ball_raster, dim3(blocks_for(c->nC, 256)), dim3(256), 0, c->stream, c->nx, c->ny, c->nz, c->dx,
                     c->nBalls, c->balls, c->cellType, c->sv);
  WS_LAUNCHED();
  c->hadSolidVelocity = c->nBalls > 0;
  return WS_OK;
}

int balls_push(ws_ctx *c, bool rebin) {
  if (!c->nBalls || c->nP <= 0) return WS_OK;
  hipLaunchKernelGGL(k_push_balls, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->nP, c->nBalls, c->balls,
                     c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz);
  WS_LAUNCHED();
  if (!rebin) return WS_OK;
  hipLaunchKernelGGL(k_zero_words, dim3(blocks_for(c->nC + 1, 256)), dim3(256), 0, c->stream, c->cellCount,
                     c->nC + 1);
  hipLaunchKernelGGL(k_rebin, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->geo, make_world(c), c->nP,
                     c->px, c->py, c->pz, c->pcell, c->prank, c->cellCount);
  WS_LAUNCHED();
  if (int r = scan_excl(c, c->cellCount, c->cellStart, (long)c->nC + 1)) return r;
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->nP, c->pcell, c->prank,
                     c->cellStart, c->sortedIdx);
  WS_LAUNCHED();
  return WS_OK;
}

int balls_update(ws_ctx *c) {
  if (!c->nBalls) return WS_OK;
  if (c->haveTerrain && !c->haveSolid) {
    c->err = "balls on a terrain world need the voxel terrain (terrainSolid) for their rock contacts";
    return WS_ESTATE;
  }
  int nb = c->nBalls;
  hipLaunchKernelGGL(k_zero_words, dim3(blocks_for((long)nb * kHistRow, 256)), dim3(256), 0, c->stream, c->ballHist,
                     (long)nb * kHistRow);
  if (c->nP > 0) {
    int g = blocks_for(c->nP, 256);
    if (g > 1024) g = 1024;
    double hSpan = (c->oceanR - c->coreR) * 1.5;
    hipLaunchKernelGGL(k_ball_hist, dim3(g), dim3(256), 0, c->stream, c->nP, nb, c->balls, c->cx, c->cy, c->cz,
                       c->coreR, hSpan, c->px, c->py, c->pz, c->pflag, c->ballHist);
  }
  hipLaunchKernelGGL(k_ball_update, dim3(nb), dim3(256), 0, c->stream, ball_env(c), nb, c->balls, c->ballHist,
                     c->ballSurf, c->vel, c->haveSolid ? c->tsolid : nullptr, c->ds);
  if (nb > 1) hipLaunchKernelGGL(k_ball_pairs, dim3(1), dim3(64), 0, c->stream, nb, c->balls);
  WS_LAUNCHED();
  return WS_OK;
}

void balls_release(ws_ctx *c) {
  (void)hipFree(c->balls);
  (void)hipFree(c->sv);
  (void)hipFree(c->ballHist);
  (void)hipFree(c->ballSurf);
  c->balls = nullptr;
  c->sv = c->ballSurf = nullptr;
  c->ballHist = nullptr;
  c->nBalls = 0;
}

}  // namespace ws

using namespace ws;

extern "C" {

int ws_set_balls(ws_ctx *c, int n, const double *state) {
  if (!c || n < 0 || (n > 0 && !state)) return WS_EINVAL;
  if (n > WS_MAX_BALLS) {
    c->err = "at most " + std::to_string(WS_MAX_BALLS) + " balls";
    return WS_EINVAL;
  }
  if (n > 0 && !c->balls) {
    WS_HIP(hipMalloc(&c->balls, sizeof(double) * 9 * WS_MAX_BALLS));
    WS_HIP(hipMalloc(&c->sv, sizeof(float) * 3 * c->nC));
    WS_HIP(hipMalloc(&c->ballHist, sizeof(unsigned) * kHistRow * WS_MAX_BALLS));
    WS_HIP(hipMalloc(&c->ballSurf, sizeof(float) * WS_MAX_BALLS));
    WS_HIP(hipMemsetAsync(c->sv, 0, sizeof(float) * 3 * c->nC, c->stream));
    WS_HIP(hipMemsetAsync(c->ballHist, 0, sizeof(unsigned) * kHistRow * WS_MAX_BALLS, c->stream));
    WS_HIP(hipMemsetAsync(c->ballSurf, 0, sizeof(float) * WS_MAX_BALLS, c->stream));
  }
  if (n > 0) WS_HIP(hipMemcpyAsync(c->balls, state, sizeof(double) * 9 * n, hipMemcpyHostToDevice, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  c->nBalls = n;
  return WS_OK;
}

int ws_get_balls(ws_ctx *c, int n, double *state) {
  if (!c || n < 0 || n > c->nBalls || (n > 0 && !state)) return WS_EINVAL;
  if (n == 0) return WS_OK;
  WS_HIP(hipMemcpyAsync(state, c->balls, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

}  // extern "C"
