// Particle-side kernels of the substep: G2P (js/solver.js:1646-1756) fused
// with advection (js/solver.js:1769-2292) for grid-coupled fluid, spray, rain
// and ice, and the three-phase vapor advection whose ceiling and floor
// bounces draw from the reference's sequential LCG streams (solver.js:2061,
// 2084) by jump-ahead to each particle's offset in the stream.
#include "ws_kernels.h"

namespace ws {

__device__ __forceinline__ void sample3(const Geo &g, const float *__restrict__ vel, float x, float y, float z,
                                        float out[3]) {
  Stencil s;
  face_stencil(g, 0, x, y, z, s);
  out[0] = stencil_eval(s, vel);
  face_stencil(g, 1, x, y, z, s);
  out[1] = stencil_eval(s, vel + g.uN);
  face_stencil(g, 2, x, y, z, s);
  out[2] = stencil_eval(s, vel + g.uN + g.vN);
}

// sampleVel (solver.js:811-854) at n caller points: the sampler G2P uses,
// so a sample at a particle's position equals the G2P cache entry
__global__ void k_sample_vel(Geo g, long n, const float *__restrict__ xyz, const float *__restrict__ vel,
                             float *__restrict__ out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v[3];
  sample3(g, vel, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], v);
  out[3 * i] = v[0];
  out[3 * i + 1] = v[1];
  out[3 * i + 2] = v[2];
}

__device__ __forceinline__ float rlen(float x, float y, float z) {
  float r = sqrtf(x * x + y * y + z * z);
  return r != 0.f ? r : 1e-9f;  // "|| 1e-9" in the reference
}

// _onRockBelow, solver.js:581-587
__device__ __forceinline__ bool on_rock_below(const World &w, const Geo &g, float x, float y, float z) {
  if (!w.R) return false;
  float ex = x - g.cx, ey = y - g.cy, ez = z - g.cz;
  float el = rlen(ex, ey, ez);
  float s = g.dx * 0.9f / el;
  return rock_cell(w, g, x - ex * s, y - ey * s, z - ez * s);
}

// G2P of one grid-coupled particle (flag 0), solver.js:1664-1753. Returns
// true if it was demoted to spray. adv receives the new-grid sample.
__device__ bool g2p_one(const Geo &g, const World &w, const Par &P, const float *__restrict__ vel,
                        const float *__restrict__ velO, const uint8_t *__restrict__ ct, float dt, float x, float y,
                        float z, float &vx, float &vy, float &vz, uint8_t &cool, uint8_t &flag, float adv[3]) {
  if (cool > 0) {
    cool--;
    vy = vy - P.gravity * dt;
    sample3(g, vel, x, y, z, adv);
    return false;
  }
  Stencil s;
  face_stencil(g, 0, x, y, z, s);
  float ou = stencil_eval(s, velO), nu = stencil_eval(s, vel);
  face_stencil(g, 1, x, y, z, s);
  float ov = stencil_eval(s, velO + g.uN), nv = stencil_eval(s, vel + g.uN);
  face_stencil(g, 2, x, y, z, s);
  float ow = stencil_eval(s, velO + g.uN + g.vN), nw = stencil_eval(s, vel + g.uN + g.vN);
  adv[0] = nu; adv[1] = nv; adv[2] = nw;
  float a = 1.f - P.pic;
  vx = (vx + nu - ou) * a + nu * P.pic;
  vy = (vy + nv - ov) * a + nv * P.pic;
  vz = (vz + nw - ow) * a + nw * P.pic;
  float pv2 = vx * vx + vy * vy + vz * vz;
  if (pv2 > P.maxSpeed * P.maxSpeed) {
    float psc = P.maxSpeed / sqrtf(pv2);
    vx *= psc; vy *= psc; vz *= psc;
  }
  int ci = clampi(cell_floor(x, g.inv_dx, w.thr, g.nx), 0, g.nx - 1);
  int cj = clampi(cell_floor(y, g.inv_dx, w.thr, g.ny), 0, g.ny - 1);
  int ck = clampi(cell_floor(z, g.inv_dx, w.thr, g.nz), 0, g.nz - 1);
  int c = (ck * g.ny + cj) * g.nx + ci, nxy = g.nx * g.ny;
  if (ct[c] == AIR) {
    bool nb = (ci > 0 && ct[c - 1] == FLUID) || (ci < g.nx - 1 && ct[c + 1] == FLUID) ||
              (cj > 0 && ct[c - g.nx] == FLUID) || (cj < g.ny - 1 && ct[c + g.nx] == FLUID) ||
              (ck > 0 && ct[c - nxy] == FLUID) || (ck < g.nz - 1 && ct[c + nxy] == FLUID);
    if (!nb && !on_rock_below(w, g, x, y, z)) { flag = 1; return true; }
  }
  return false;
}

__device__ __forceinline__ int cell_of(const Geo &g, const World &w, float x, float y, float z) {
  int ci = clampi(cell_floor(x, g.inv_dx, w.thr, g.nx), 0, g.nx - 1);
  int cj = clampi(cell_floor(y, g.inv_dx, w.thr, g.ny), 0, g.ny - 1);
  int ck = clampi(cell_floor(z, g.inv_dx, w.thr, g.nz), 0, g.nz - 1);
  return (ck * g.ny + cj) * g.nx + ci;
}

// Advection of one particle with flag 0, 1, 4 or 5 (sphere world),
// solver.js:1797-1976 and :2176-2290. adv is the G2P sample at (x,y,z).
__device__ void advect_one(const Geo &g, const World &w, const Par &P, const float *__restrict__ vel,
                           const uint8_t *__restrict__ ct, float dt, float &x, float &y, float &z, float &vx,
                           float &vy, float &vz, uint8_t &flag, uint8_t &cool, float &air, const float adv[3]) {
  const float cx = g.cx, cy = g.cy, cz = g.cz, dx = g.dx;
  const float domHi = P.domainR - dx * 0.02f;
  const bool hasTerr = w.R != nullptr;
  if (flag == 5) {  // ice, solver.js:1797-1864
    if (air == 1.f || !hasTerr) { vx = vy = vz = 0.f; return; }
    float gI = P.gravity != 0.f ? P.gravity : 9.81f;
    float exI = x - cx, eyI = y - cy, ezI = z - cz, elI = rlen(exI, eyI, ezI);
    float axI = -gI * 0.6f * exI / elI, ayI = -gI * 0.6f * eyI / elI, azI = -gI * 0.6f * ezI / elI;
    vx = vx + (axI - 12.f * vx) * dt;
    vy = vy + (ayI - 12.f * vy) * dt;
    vz = vz + (azI - 12.f * vz) * dt;
    x += vx * dt; y += vy * dt; z += vz * dt;
    float rrI = sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    float top = P.oceanR + P.atmosphereH;
    if (rrI > top) {
      float s = top / rrI;
      x = cx + (x - cx) * s; y = cy + (y - cy) * s; z = cz + (z - cz) * s;
      vx *= 0.5f; vy *= 0.5f; vz *= 0.5f;
    }
    if (rock_cell(w, g, x, y, z)) { air = 1.f; vx = vy = vz = 0.f; return; }
    int ci = (int)floorf(x * g.inv_dx), cj = (int)floorf(y * g.inv_dx), ck = (int)floorf(z * g.inv_dx);
    if (ci >= 0 && cj >= 0 && ck >= 0 && ci < g.nx && cj < g.ny && ck < g.nz &&
        ct[(ck * g.ny + cj) * g.nx + ci] == FLUID) {
      float exF = x - cx, eyF = y - cy, ezF = z - cz, elF = rlen(exF, eyF, ezF), st = dx * 0.5f;
      for (int mf = 0; mf < 12; mf++) {
        float qx = x + exF / elF * st, qy = y + eyF / elF * st, qz = z + ezF / elF * st;
        int qi = (int)floorf(qx * g.inv_dx), qj = (int)floorf(qy * g.inv_dx), qk = (int)floorf(qz * g.inv_dx);
        if (qi < 0 || qj < 0 || qk < 0 || qi >= g.nx || qj >= g.ny || qk >= g.nz) break;
        if (ct[(qk * g.ny + qj) * g.nx + qi] != FLUID) break;
        x = qx; y = qy; z = qz;
      }
      air = 1.f; vx = vy = vz = 0.f;
    }
    return;
  }
  if (flag == 1 || flag == 4) {  // ballistic spray / rain, solver.js:1865-1976
    float gx = x - cx, gy = y - cy, gz = z - cz;
    float gr2 = gx * gx + gy * gy + gz * gz;
    float ga = P.gravity * P.oceanR * P.oceanR * dt / (gr2 * sqrtf(gr2));
    vx = vx - gx * ga; vy = vy - gy * ga; vz = vz - gz * ga;
    float dr = fmaxf(1.f - 2.5f * dt, 0.f);
    vx *= dr; vy *= dr; vz *= dr;
    float dsp = sqrtf(vx * vx + vy * vy + vz * vz);
    if (dsp > P.maxSpeed) { float s = P.maxSpeed / dsp; vx *= s; vy *= s; vz *= s; }
    x += vx * dt; y += vy * dt; z += vz * dt;
    float bx = x - cx, by = y - cy, bz = z - cz, br = rlen(bx, by, bz);
    float skyHi = P.oceanR + P.atmosphereH;
    if (br > skyHi) {
      float s0 = skyHi / br;
      x = cx + bx * s0; y = cy + by * s0; z = cz + bz * s0;
      float vr0 = (vx * bx + vy * by + vz * bz) / br;
      if (vr0 > 0.f) {
        vx -= bx / br * vr0 * 1.15f; vy -= by / br * vr0 * 1.15f; vz -= bz / br * vr0 * 1.15f;
      }
    } else if (br <= domHi) {
      int tries = 0;
      while (hasTerr && rock_cell(w, g, x, y, z) && tries < 8) {
        float rW = sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
        float st = fminf(dx * 0.5f, domHi - rW);
        if (st <= 1e-5f) { tries = 8; break; }
        x += bx / br * st; y += by / br * st; z += bz / br * st;
        tries++;
      }
      if (!hasTerr && br < P.coreR + dx * 1.02f) {
        float s1 = (P.coreR + dx * 1.02f) / br;
        x = cx + bx * s1; y = cy + by * s1; z = cz + bz * s1;
        tries = 1;
      }
      if (tries > 0) {
        if (hasTerr) {
          float exS = x - cx, eyS = y - cy, ezS = z - cz, erS = rlen(exS, eyS, ezS);
          float want = terrain_radius(w, x, y, z) + dx * 0.06f;
          if (erS > want) { x = cx + exS / erS * want; y = cy + eyS / erS * want; z = cz + ezS / erS * want; }
          int kS = 0;
          while (rock_cell(w, g, x, y, z) && kS < 4) {
            x += exS / erS * dx * 0.5f; y += eyS / erS * dx * 0.5f; z += ezS / erS * dx * 0.5f;
            kS++;
          }
        }
        float vr1 = (vx * bx + vy * by + vz * bz) / br;
        vx -= bx / br * vr1; vy -= by / br * vr1; vz -= bz / br * vr1;
        vx *= 0.3f; vy *= 0.3f; vz *= 0.3f;
        flag = 0;
      }
    }
    if (ct[cell_of(g, w, x, y, z)] == FLUID) {
      vx *= 0.45f; vy *= 0.45f; vz *= 0.45f;
      if (vy > 0.5f) vy = 0.5f;
      cool = 2;
      flag = 0;
    }
    return;
  }
  if (flag != 0) return;  // vapor: advect_vapor_*
  // grid-coupled fluid, solver.js:2176-2290
  const float xLo = dx * 1.02f, xHi = P.W - dx * 1.02f;
  float t0x = adv[0], t0y = adv[1], t0z = adv[2];
  float sp2 = t0x * t0x + t0y * t0y + t0z * t0z;
  float v2x = t0x, v2y = t0y, v2z = t0z;
  if (sp2 > 0.25f) {
    float mx = fminf(fmaxf(x + 0.5f * dt * t0x, xLo), xHi);
    float my = fminf(fmaxf(y + 0.5f * dt * t0y, xLo), P.H - dx * 1.02f);
    float mz = fminf(fmaxf(z + 0.5f * dt * t0z, xLo), P.D - dx * 1.02f);
    float t2[3];
    sample3(g, vel, mx, my, mz, t2);
    v2x = t2[0]; v2y = t2[1]; v2z = t2[2];
  }
  sp2 = v2x * v2x + v2y * v2y + v2z * v2z;
  if (sp2 > P.maxSpeed * P.maxSpeed) {
    float sc = P.maxSpeed / sqrtf(sp2);
    v2x *= sc; v2y *= sc; v2z *= sc;
    vx *= sc; vy *= sc; vz *= sc;
  }
  x += v2x * dt; y += v2y * dt; z += v2z * dt;
  float nx2 = x - cx, ny2 = y - cy, nz2 = z - cz, nr = rlen(nx2, ny2, nz2);
  if (nr > domHi) {
    float s2 = domHi / nr;
    x = cx + nx2 * s2; y = cy + ny2 * s2; z = cz + nz2 * s2;
    vy = -fabsf(vy) * 0.1f - 0.3f;
    flag = 1;
  }
  if (nr < domHi) {
    int tries = 0;
    while (hasTerr && rock_cell(w, g, x, y, z) && tries < 8) {
      float rW = sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
      float st = fminf(dx * 0.5f, domHi - rW);
      if (st <= 1e-5f) { tries = 8; break; }
      x += nx2 / nr * st; y += ny2 / nr * st; z += nz2 / nr * st;
      tries++;
    }
    if (tries > 0) {
      if (hasTerr) {
        float exS = x - cx, eyS = y - cy, ezS = z - cz, erS = rlen(exS, eyS, ezS);
        float want = terrain_radius(w, x, y, z) + dx * 0.06f;
        if (erS > want) { x = cx + exS / erS * want; y = cy + eyS / erS * want; z = cz + ezS / erS * want; }
        int kS = 0;
        while (rock_cell(w, g, x, y, z) && kS < 4) {
          x += exS / erS * dx * 0.5f; y += eyS / erS * dx * 0.5f; z += ezS / erS * dx * 0.5f;
          kS++;
        }
      }
      float vr2 = (vx * nx2 + vy * ny2 + vz * nz2) / nr;
      if (vr2 < 0.f) { vx -= nx2 / nr * vr2; vy -= ny2 / nr * vr2; vz -= nz2 / nr * vr2; }
    }
    if (hasTerr && (tries > 0 || rock_cell(w, g, x - nx2 / nr * dx, y - ny2 / nr * dx, z - nz2 / nr * dx))) {
      float fr0 = fmaxf(1.f - 0.6f * dt, 0.f);
      vx *= fr0; vy *= fr0; vz *= fr0;
    }
  }
  float sx = x - cx, sy = y - cy, sz = z - cz, sr = rlen(sx, sy, sz);
  float vrad = (vx * sx + vy * sy + vz * sz) / sr;
  float da = fminf(1.f * dt, 0.5f);
  vx -= sx / sr * vrad * da; vy -= sy / sr * vrad * da; vz -= sz / sr * vrad * da;
  float dg = fmaxf(1.f - 0.42f * dt, 0.f);
  vx *= dg; vy *= dg; vz *= dg;
}

// mode: 1 = G2P only (writes adv), 2 = advect only (reads adv), 3 = fused
__device__ __forceinline__ void k_g2p_advect_body(long bx, long nbx, Geo g, World w, Par P, int nP, int mode,
                                                  const float *__restrict__ vel, const float *__restrict__ velO,
                                                  const uint8_t *__restrict__ ct, float *__restrict__ px,
                                                  float *__restrict__ py, float *__restrict__ pz,
                                                  float *__restrict__ pvx, float *__restrict__ pvy,
                                                  float *__restrict__ pvz, uint8_t *__restrict__ pflag,
                                                  uint8_t *__restrict__ pcool, float *__restrict__ pAir,
                                                  float *__restrict__ advU, float *__restrict__ advV,
                                                  float *__restrict__ advW, DevScalars *s) {
  int p = bx * blockDim.x + threadIdx.x;
  unsigned airborne = 0;
  if (p < nP) {
    float dt = s->dt;
    uint8_t flag = pflag[p], cool = pcool[p];
    float x = px[p], y = py[p], z = pz[p], vx = pvx[p], vy = pvy[p], vz = pvz[p];
    float adv[3] = {0.f, 0.f, 0.f};
    bool changed = false;
    if (mode & 1) {
      if (flag != 0) airborne = 1;
      else {
        float dtt = dt != 0.f ? dt : (s->dtLast != 0.f ? s->dtLast : 0.01f);
        if (g2p_one(g, w, P, vel, velO, ct, dtt, x, y, z, vx, vy, vz, cool, flag, adv)) airborne = 1;
        changed = true;
        if (mode == 1) { advU[p] = adv[0]; advV[p] = adv[1]; advW[p] = adv[2]; }
      }
    } else if (flag == 0) {
      adv[0] = advU[p]; adv[1] = advV[p]; adv[2] = advW[p];
    }
    float air = pAir[p];
    if ((mode & 2) && flag != 2 && flag != 3) {
      advect_one(g, w, P, vel, ct, dt, x, y, z, vx, vy, vz, flag, cool, air, adv);
      changed = true;
      pAir[p] = air;
      px[p] = x; py[p] = y; pz[p] = z;
    }
    if (changed) {
      pvx[p] = vx; pvy[p] = vy; pvz[p] = vz;
      pflag[p] = flag;
      pcool[p] = cool;
    }
  }
  if (mode & 1) {
    unsigned long long m = __ballot(airborne);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&s->airborne, (unsigned)__popcll(m));
  }
}
__global__ __launch_bounds__(256) void k_g2p_advect(Geo g, World w, Par P, int nP, int mode,
                                                    const float *__restrict__ vel, const float *__restrict__ velO,
                                                    const uint8_t *__restrict__ ct, float *__restrict__ px,
                                                    float *__restrict__ py, float *__restrict__ pz,
                                                    float *__restrict__ pvx, float *__restrict__ pvy,
                                                    float *__restrict__ pvz, uint8_t *__restrict__ pflag,
                                                    uint8_t *__restrict__ pcool, float *__restrict__ pAir,
                                                    float *__restrict__ advU, float *__restrict__ advV,
                                                    float *__restrict__ advW, DevScalars *s) {
  k_g2p_advect_body(blockIdx.x, gridDim.x, g, w, P, nP, mode, vel, velO, ct, px, py, pz, pvx, pvy, pvz, pflag, pcool,
                    pAir, advU, advV, advW, s);
}
WS_PLANETS_OPS(k_g2p_advect_ops, k_g2p_advect, k_g2p_advect_body)

// ---------------------------------------------------------------- vapor
// Vapor and cloud parcels, solver.js:1978-2174. The reference draws from two
// LCG streams in particle order (ceiling bounce _cfS, floor bounce _bfS);
// each phase ends where a draw may happen, an exclusive scan of the draw
// flags gives every parcel its position in the stream, and the next phase
// jumps the LCG ahead to it. scratch per particle: vC (3), draw threshold.
__device__ __forceinline__ uint32_t lcg_jump(uint32_t s, uint32_t n) {
  uint32_t A = 1u, C = 0u, a = 1664525u, c = 1013904223u;
  while (n) {
    if (n & 1u) { A = a * A; C = a * C + c; }
    c = a * c + c;
    a = a * a;
    n >>= 1;
  }
  return A * s + C;
}

__global__ void k_vapor_a(Geo g, World w, Par P, int nP, float *__restrict__ px, float *__restrict__ py,
                          float *__restrict__ pz, float *__restrict__ pvx, float *__restrict__ pvy,
                          float *__restrict__ pvz, const uint8_t *__restrict__ pflag, float *__restrict__ pAir,
                          float *__restrict__ scr, unsigned *__restrict__ need, const DevScalars *s) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  uint8_t fl = pflag[p];
  if (fl != 2 && fl != 3) { need[p] = 0; return; }
  float dt = s->dt, cx = g.cx, cy = g.cy, cz = g.cz;
  float x = px[p], y = py[p], z = pz[p], vx = pvx[p], vy = pvy[p], vz = pvz[p];
  pAir[p] = pAir[p] + dt;
  float vEx = x - cx, vEy = y - cy, vEz = z - cz, vEr = rlen(vEx, vEy, vEz);
  float ux = vEx / vEr, uy = vEy / vEr, uz = vEz / vEr;
  float exposure = ux * P.sunX + uy * P.sunY + uz * P.sunZ;
  float jet = fminf(P.oceanR * 0.58f, P.maxSpeed * 0.45f);
  float ret = fminf(P.oceanR * 0.32f, P.maxSpeed * 0.25f);
  float vr = vx * ux + vy * uy + vz * uz;
  float radial = -P.gravity * 0.06f;
  float wx = -uz * jet - (P.sunX - exposure * ux) * ret;
  float wy = -(P.sunY - exposure * uy) * ret;
  float wz = ux * jet - (P.sunZ - exposure * uz) * ret;
  float relax = fminf(dt * 1.3f, 1.f);
  float nvx = vx + (wx - (vx - vr * ux)) * relax + ux * radial * dt;
  float nvy = vy + (wy - (vy - vr * uy)) * relax + uy * radial * dt;
  float nvz = vz + (wz - (vz - vr * uz)) * relax + uz * radial * dt;
  vx = nvx; vy = nvy; vz = nvz;
  float svv = vx * vx + vy * vy + vz * vz;
  if (svv > P.maxSpeed * P.maxSpeed) { float q = P.maxSpeed / sqrtf(svv); vx *= q; vy *= q; vz *= q; }
  x += vx * dt; y += vy * dt; z += vz * dt;
  float vCx = x - cx, vCy = y - cy, vCz = z - cz, vCr = rlen(vCx, vCy, vCz);
  float ceilR = P.oceanR + P.atmosphereH - 1e-3f * g.dx;
  if (w.R) {
    float RtE = terrain_radius(w, x, y, z);
    if (RtE + 0.02f > ceilR) ceilR = RtE + 0.02f;
  }
  if (vCr > ceilR) {
    float sE = ceilR / vCr;
    x = cx + vCx * sE; y = cy + vCy * sE; z = cz + vCz * sE;
    float vrE = (vx * vCx + vy * vCy + vz * vCz) / vCr;
    if (vrE > 0.f) { vx -= vCx / vCr * vrE * 2.f; vy -= vCy / vCr * vrE * 2.f; vz -= vCz / vCr * vrE * 2.f; }
  }
  float vrC2 = (vx * vCx + vy * vCy + vz * vCz) / vCr;
  unsigned nd = 0;
  float thresh = 0.f;
  if (vrC2 > 0.f && vCr > P.oceanR) {
    float hh = fminf((vCr - P.oceanR) / fmaxf(P.atmosphereH, 1e-4f), 1.f);
    float prf = P.ceilReflect == 1 ? hh * hh : P.ceilReflect == 2 ? (expf(3.f * hh) - 1.f) * 0.0523957f : hh;
    if (prf > 0.f) { nd = 1; thresh = prf * dt * 4.f; }
  }
  px[p] = x; py[p] = y; pz[p] = z;
  pvx[p] = vx; pvy[p] = vy; pvz[p] = vz;
  long C4 = (long)nP;
  scr[p] = vCx; scr[C4 + p] = vCy; scr[2 * C4 + p] = vCz; scr[3 * C4 + p] = thresh;
  need[p] = nd;
}

__global__ void k_vapor_b(Geo g, Par P, int nP, const float *__restrict__ px, const float *__restrict__ py,
                          const float *__restrict__ pz, float *__restrict__ pvx, float *__restrict__ pvy,
                          float *__restrict__ pvz, const uint8_t *__restrict__ pflag, const float *__restrict__ scr,
                          const unsigned *__restrict__ needA, const unsigned *__restrict__ offA,
                          unsigned *__restrict__ needB, const DevScalars *s, const uint32_t *rng) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  uint8_t fl = pflag[p];
  if (fl != 2 && fl != 3) { needB[p] = 0; return; }
  long C4 = (long)nP;
  float vCx = scr[p], vCy = scr[C4 + p], vCz = scr[2 * C4 + p], thresh = scr[3 * C4 + p];
  float vCr = rlen(vCx, vCy, vCz);
  float vx = pvx[p], vy = pvy[p], vz = pvz[p];
  if (needA[p]) {
    uint32_t st = lcg_jump(rng[0], offA[p] + 1u);
    if ((float)st * 2.3283064365386963e-10f < thresh) {
      float vrC2 = (vx * vCx + vy * vCy + vz * vCz) / vCr;
      vx -= vCx / vCr * vrC2 * 2.f; vy -= vCy / vCr * vrC2 * 2.f; vz -= vCz / vCr * vrC2 * 2.f;
    }
  }
  float x = px[p], y = py[p], z = pz[p];
  float fCx = x - g.cx, fCy = y - g.cy, fCz = z - g.cz, fCr = rlen(fCx, fCy, fCz);
  float floorR = P.oceanR + 0.15f * g.dx;
  unsigned nd = 0;
  if (fCr < floorR && (vx * fCx + vy * fCy + vz * fCz) / fCr < 0.f) nd = 1;
  pvx[p] = vx; pvy[p] = vy; pvz[p] = vz;
  needB[p] = nd;
}

__global__ void k_vapor_c(Geo g, World w, Par P, int nP, float *__restrict__ px, float *__restrict__ py,
                          float *__restrict__ pz, float *__restrict__ pvx, float *__restrict__ pvy,
                          float *__restrict__ pvz, uint8_t *__restrict__ pflag, float *__restrict__ pAir,
                          const float *__restrict__ pT, const uint8_t *__restrict__ ct, const float *__restrict__ scr,
                          const unsigned *__restrict__ needB, const unsigned *__restrict__ offB,
                          const DevScalars *s, const uint32_t *rng) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  uint8_t fl = pflag[p];
  if (fl != 2 && fl != 3) return;
  float cx = g.cx, cy = g.cy, cz = g.cz, dx = g.dx;
  long C4 = (long)nP;
  float vCx = scr[p], vCy = scr[C4 + p], vCz = scr[2 * C4 + p], vCr = rlen(vCx, vCy, vCz);
  float x = px[p], y = py[p], z = pz[p], vx = pvx[p], vy = pvy[p], vz = pvz[p];
  float floorR = P.oceanR + 0.15f * dx;
  if (needB[p]) {
    float fCx = x - cx, fCy = y - cy, fCz = z - cz, fCr = rlen(fCx, fCy, fCz);
    float vrF = (vx * fCx + vy * fCy + vz * fCz) / fCr;
    uint32_t st = lcg_jump(rng[1], offB[p] + 1u);
    float sc = ((float)st * 2.3283064365386963e-10f < 0.07f) ? floorR / fCr : (floorR + 0.02f * dx) / fCr;
    x = cx + fCx * sc; y = cy + fCy * sc; z = cz + fCz * sc;
    vx -= fCx / fCr * vrF * 2.f; vy -= fCy / fCr * vrF * 2.f; vz -= fCz / fCr * vrF * 2.f;
  }
  float domHi = P.domainR - dx * 0.02f;
  int tries = 0;
  float rNow = sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
  while (w.R && rNow < P.domainR && rock_cell(w, g, x, y, z) && tries < 8) {
    float rW = sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    float st = fminf(dx * 0.5f, domHi - rW);
    if (st <= 1e-5f) break;
    x += vCx / vCr * st; y += vCy / vCr * st; z += vCz / vCr * st;
    tries++;
  }
  if (tries > 0) {
    float vrT = (vx * vCx + vy * vCy + vz * vCz) / vCr;
    if (vrT < 0.f) { float bT = 1.5f * vrT / vCr; vx -= vCx * bT; vy -= vCy * bT; vz -= vCz * bT; }
    vx *= 0.92f; vy *= 0.92f; vz *= 0.92f;
  }
  bool back = ct[cell_of(g, w, x, y, z)] == FLUID;
  if (back) {
    float bx = x - cx, by = y - cy, bz = z - cz, br = rlen(bx, by, bz);
    if (br > P.oceanR + 0.15f * dx) back = false;
    else if ((vx * bx + vy * by + vz * bz) / br > 0.02f) back = false;
  }
  px[p] = x; py[p] = y; pz[p] = z;
  float air = pAir[p];
  if (back) {
    vx = vy = vz = 0.f; air = 0.f; fl = 0;
  } else if (fl == 2 && ((pT[p] <= P.rainTe && air >= P.rainLiftAge) || air > 40.f)) {
    vx = vy = vz = 0.f; air = 0.f; fl = 4;
  }
  pvx[p] = vx; pvy[p] = vy; pvz[p] = vz;
  pAir[p] = air;
  pflag[p] = fl;
}

// stream state after n draws (one thread)
__global__ void k_rng_advance(uint32_t *rng, int which, const unsigned *need, const unsigned *off, int nP) {
  if (threadIdx.x != 0 || nP <= 0) return;
  unsigned total = off[nP - 1] + need[nP - 1];
  rng[which] = lcg_jump(rng[which], total);
}

}  // namespace ws
