// Grid-side kernels of the substep (js/solver.js:865-1636): rasterize + mark,
// particle binning, deterministic P2G, boundary conditions, viscosity,
// gravity, pressure assembly, red-black SOR, projection, vorticity
// confinement, grid clamp and the extrapolation flood.
#include "ws_kernels.h"

namespace ws {

// ---------------------------------------------------------------- scalars
// dt and the gravity ramp of this substep (solver.js:2656-2662); mode 1 takes
// dt from the CFL law of _stepSync (solver.js:2689-2690) on the device.
__device__ __forceinline__ void k_begin_substep_body(long bx, long nbx, DevScalars *s, double dt_host, int use_law,
                                                     double cfl_dx, double dt_base, float gravity) {
  if (threadIdx.x != 0) return;
  // the length is a double, as in the reference; the float copy feeds the
  // float kernels
  double dt = dt_host;
  if (use_law) {
    double um = (double)__uint_as_float(s->umax_bits);
    double law = cfl_dx / (um > 0.08 ? um : 0.08);
    dt = law < dt_base ? law : dt_base;
    if (dt > 0.01) dt = 0.01;
    if (dt < 1e-5) dt = 1e-5;
  }
  double gsd = s->gRamp / 1.5;
  if (gsd > 1) gsd = 1;
  s->gRamp += dt;
  s->gs = gsd <= 0 ? 0.f : (float)(gsd * (double)gravity);
  s->dt = (float)dt;
  s->dtD = dt;
  s->simTime += dt;
}
__global__ void k_begin_substep(DevScalars *s, double dt_host, int use_law, double cfl_dx, double dt_base,
                                float gravity) {
  k_begin_substep_body(blockIdx.x, gridDim.x, s, dt_host, use_law, cfl_dx, dt_base, gravity);
}
WS_PLANETS_OPS(k_begin_substep_ops, k_begin_substep, k_begin_substep_body)

__device__ __forceinline__ void k_end_substep_body(long bx, long nbx, DevScalars *s) {
  if (threadIdx.x == 0) s->dtLast = s->dt;
}
__global__ void k_end_substep(DevScalars *s) {
  k_end_substep_body(blockIdx.x, gridDim.x, s);
}
WS_PLANETS_OPS(k_end_substep_ops, k_end_substep, k_end_substep_body)

// ------------------------------------------------------------- rasterize
// _rasterizeStatic (cached copy, solver.js:868-870) and zeroed bin counts
__device__ __forceinline__ void k_init_cells_body(long bx, long nbx, const uint8_t *__restrict__ st,
                                                  uint8_t *__restrict__ ct, unsigned *__restrict__ cnt, int nC) {
  int c = bx * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  ct[c] = st[c];
  cnt[c] = 0;
}
__global__ void k_init_cells(const uint8_t *__restrict__ st, uint8_t *__restrict__ ct, unsigned *__restrict__ cnt,
                             int nC) {
  k_init_cells_body(blockIdx.x, gridDim.x, st, ct, cnt, nC);
}
WS_PLANETS_OPS(k_init_cells_ops, k_init_cells, k_init_cells_body)

// "sea in the sky" test of _markFluidCells (solver.js:964-975) for a float
// position: float filter, float-float when the float margin is inside its
// error band. Returns 1 = demote.
__device__ int sky_demote(const World &w, const Geo &g, float x, float y, float z, unsigned *refined) {
  float rx = x - g.cx, ry = y - g.cy, rz = z - g.cz;
  float r2 = rx * rx + ry * ry + rz * rz;
  float lim = w.lim_df.hi;
  int c1;
  if (r2 > lim * (1.f + 1e-4f)) c1 = 1;
  else if (r2 < lim * (1.f - 1e-4f)) c1 = 0;
  else c1 = -1;
  df r2d;
  bool have_r2d = false;
  if (c1 < 0) {
    df ex = df_sub(df_make(x), w.cxd), ey = df_sub(df_make(y), w.cyd), ez = df_sub(df_make(z), w.czd);
    r2d = df_add(df_add(df_mul(ex, ex), df_mul(ey, ey)), df_mul(ez, ez));
    have_r2d = true;
    c1 = df_gt0(df_sub(r2d, w.lim_df)) ? 1 : 0;
    atomicAdd(refined, 1u);
  }
  if (!c1) return 0;
  float m = sqrtf(r2) - terrain_radius(w, x, y, z) - w.skyAlt;
  int c2;
  if (m > 2e-5f) c2 = 1;
  else if (m < -2e-5f) c2 = 0;
  else {
    if (!have_r2d) {
      df ex = df_sub(df_make(x), w.cxd), ey = df_sub(df_make(y), w.cyd), ez = df_sub(df_make(z), w.czd);
      r2d = df_add(df_add(df_mul(ex, ex), df_mul(ey, ey)), df_mul(ez, ez));
    }
    df md = df_sub(df_sub(df_sqrt(r2d), terrain_radius_df(w, x, y, z)), w.skyAlt_df);
    c2 = df_gt0(md) ? 1 : 0;
    atomicAdd(refined, 1u);
  }
  if (!c2) return 0;
  return rock_cell(w, g, x, y, z) ? 0 : 1;
}

// ceiling guard on the planet (solver.js:979-991): airborne > 0.3 over ground
__device__ int ceiling_demote(const World &w, const Geo &g, float x, float y, float z, unsigned *refined) {
  float rx = x - g.cx, ry = y - g.cy, rz = z - g.cz;
  float m = sqrtf(rx * rx + ry * ry + rz * rz) - terrain_radius(w, x, y, z) - 0.3f;
  if (m > 2e-5f) return 1;
  if (m < -2e-5f) return 0;
  df ex = df_sub(df_make(x), w.cxd), ey = df_sub(df_make(y), w.cyd), ez = df_sub(df_make(z), w.czd);
  df r2d = df_add(df_add(df_mul(ex, ex), df_mul(ey, ey)), df_mul(ez, ez));
  df md = df_sub(df_sub(df_sqrt(r2d), terrain_radius_df(w, x, y, z)), w.p3_df);
  atomicAdd(refined, 1u);
  return df_gt0(md) ? 1 : 0;
}

// _markFluidCells, solver.js:945-1000, plus the counting-sort histogram of
// the particles that stay grid-coupled (their bin is the cell they mark).
__device__ __forceinline__ void k_mark_body(long bx, long nbx, Geo g, World w, int nP, const float *__restrict__ px,
                                            const float *__restrict__ py, const float *__restrict__ pz,
                                            uint8_t *__restrict__ pflag, uint8_t *__restrict__ ct,
                                            unsigned *__restrict__ cnt, int *__restrict__ pcell,
                                            unsigned *__restrict__ prank, DevScalars *s) {
  int p = bx * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  if (pflag[p] != 0) { pcell[p] = -1; return; }
  float x = px[p], y = py[p], z = pz[p];
  if (sky_demote(w, g, x, y, z, &s->df_refined)) { pflag[p] = 1; pcell[p] = -1; return; }
  int i = clampi(cell_floor(x, g.inv_dx, w.thr, g.nx), 0, g.nx - 1);
  int j = clampi(cell_floor(y, g.inv_dx, w.thr, g.ny), 0, g.ny - 1);
  int k = clampi(cell_floor(z, g.inv_dx, w.thr, g.nz), 0, g.nz - 1);
  if (j >= g.ny - 2 && ceiling_demote(w, g, x, y, z, &s->df_refined)) { pflag[p] = 1; pcell[p] = -1; return; }
  int c = (k * g.ny + j) * g.nx + i;
  // AIR -> FLUID on the current types (static rock and ball cells stay
  // SOLID); concurrent markers of one cell all store the same byte
  if (ct[c] == AIR) ct[c] = FLUID;
  pcell[p] = c;
  prank[p] = atomicAdd(&cnt[c], 1u);
}
__global__ void k_mark(Geo g, World w, int nP, const float *__restrict__ px, const float *__restrict__ py,
                       const float *__restrict__ pz, uint8_t *__restrict__ pflag, uint8_t *__restrict__ ct,
                       unsigned *__restrict__ cnt, int *__restrict__ pcell, unsigned *__restrict__ prank,
                       DevScalars *s) {
  k_mark_body(blockIdx.x, gridDim.x, g, w, nP, px, py, pz, pflag, ct, cnt, pcell, prank, s);
}
WS_PLANETS_OPS(k_mark_ops, k_mark, k_mark_body)

__device__ __forceinline__ void k_scatter_body(long bx, long nbx, int nP, const int *__restrict__ pcell,
                                               const unsigned *__restrict__ prank, const unsigned *__restrict__ start,
                                               unsigned *__restrict__ sorted) {
  int p = bx * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  int c = pcell[p];
  if (c >= 0) sorted[start[c] + prank[p]] = (unsigned)p;
}
__global__ void k_scatter(int nP, const int *__restrict__ pcell, const unsigned *__restrict__ prank,
                          const unsigned *__restrict__ start, unsigned *__restrict__ sorted) {
  k_scatter_body(blockIdx.x, gridDim.x, nP, pcell, prank, start, sorted);
}
WS_PLANETS_OPS(k_scatter_ops, k_scatter, k_scatter_body)

// -------------------------------------------------------------------- P2G
// _p2gInto + _p2gNormalizeFace, solver.js:1035-1120. One workgroup per cell
// tile owns the tile's faces; it visits the particles binned in the tile and
// its one-cell halo (their trilinear footprint reaches one cell) and
// accumulates momentum and weight in LDS as 64-bit fixed point. Integer
// addition is exact and order independent, so the result is deterministic
// without float atomics; the reference's sequential f32 sums differ from the
// exact sums by float rounding only.
//
// The kernel is bound by the vector ALU (three components x eight corners
// per visit, two fixed-point conversions per owned corner) and by the latency
// of each visit's chain (find the row, read the particle index, read the
// particle). So:
// - a tile whose region holds no particle (two loads of the bin starts)
//   writes its zero faces and leaves;
// - each particle row of the region is cut into three segments (the -x halo
//   cell, the tile's cells, the +x halo cell), and the segments are visited
//   grouped by which side of the tile they lie on along x, y and z: the
//   lanes of a wavefront then own the same corners, and a corner no lane
//   owns is skipped (a halo particle owns half of its corners or fewer along
//   each halo axis);
// - per particle the two offsets of each axis (0 along the component's own
//   axis, 0.5 along the others) are computed once, and per component a
//   two-bit mask per axis says which of the stencil's two faces the tile
//   owns;
// - the next visit's particle loads and the one after's index load are in
//   flight while a visit is accumulated.
__device__ __forceinline__ long long fix(float v) { return (long long)__builtin_rintf(v * kFixScaleF); }

template <int TX, int TY, int TZ>
__device__ __forceinline__ void k_p2g_tile_body(long bx, long, Geo g, int tilesX, int tilesY,
                                                const unsigned *__restrict__ start, const unsigned *__restrict__ sorted,
                                                const float *__restrict__ px, const float *__restrict__ py,
                                                const float *__restrict__ pz, const float *__restrict__ pvx,
                                                const float *__restrict__ pvy, const float *__restrict__ pvz,
                                                float *__restrict__ vel, float *__restrict__ wgt) {
  // owned faces of one component (+1 plane along its axis for the last tile)
  constexpr int BOX = cmax(cmax((TX + 1) * TY * TZ, TX * (TY + 1) * TZ), TX * TY * (TZ + 1));
  constexpr int NSEG = 3 * (TY + 2) * (TZ + 2);
  static_assert(NSEG <= 512, "two segments per thread");
  __shared__ unsigned long long accM[3][BOX];
  __shared__ unsigned long long accW[3][BOX];
  __shared__ unsigned segBeg[NSEG], segPre[NSEG + 1];
  __shared__ unsigned clsCount[27], clsStart[28];
  __shared__ unsigned wsum[4];

  const int b = (int)bx;
  const int tx = b % tilesX, ty = (b / tilesX) % tilesY, tz = b / (tilesX * tilesY);
  const int i0 = tx * TX, j0 = ty * TY, k0 = tz * TZ;
  // particle rows of the tile + halo: cells [i0-1, i0+T] x [j0-1, j0+T] x [k0-1, k0+T]
  const int ci0 = max(i0 - 1, 0), ci1 = min(i0 + TX, g.nx - 1);
  const int cj0 = max(j0 - 1, 0), cj1 = min(j0 + TY, g.ny - 1);
  const int ck0 = max(k0 - 1, 0), ck1 = min(k0 + TZ, g.nz - 1);
  // owned face ranges: [i0, hi) per component and axis
  const int iHiC = min(i0 + TX, g.nx), jHiC = min(j0 + TY, g.ny), kHiC = min(k0 + TZ, g.nz);
  const int uHi = i0 + TX >= g.nx ? g.nx + 1 : i0 + TX, vHi = j0 + TY >= g.ny ? g.ny + 1 : j0 + TY,
            wHi = k0 + TZ >= g.nz ? g.nz + 1 : k0 + TZ;
  // no particle binned anywhere from the region's first cell to its last
  const bool empty = start[(ck0 * g.ny + cj0) * g.nx + ci0] == start[(ck1 * g.ny + cj1) * g.nx + ci1 + 1];
  if (!empty) {
    for (int e = threadIdx.x; e < 3 * BOX; e += blockDim.x) {
      (&accM[0][0])[e] = 0ull;
      (&accW[0][0])[e] = 0ull;
    }
    if (threadIdx.x < 27) clsCount[threadIdx.x] = 0;
    const int nrj = cj1 - cj0 + 1, nseg = 3 * nrj * (ck1 - ck0 + 1);
    __syncthreads();
    // segment s: row s / 3, part s % 3; its class (side along x, y, z: 0
    // below the tile, 1 inside, 2 above) and its slot among its class
    unsigned sb[2], sl[2], sc[2], ss[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int s = threadIdx.x + 256 * h;
      sl[h] = 0;
      sb[h] = sc[h] = ss[h] = 0;
      if (s < nseg) {
        const int r = s / 3, part = s % 3;
        const int cj = cj0 + r % nrj, ck = ck0 + r / nrj;
        const int row = (ck * g.ny + cj) * g.nx;
        const int a = part == 0 ? ci0 : (part == 1 ? i0 : iHiC), e = part == 0 ? i0 : (part == 1 ? iHiC : ci1 + 1);
        if (e > a) {
          sb[h] = start[row + a];
          sl[h] = start[row + e] - sb[h];
        }
        const int jc = cj < j0 ? 0 : (cj >= j0 + TY ? 2 : 1), kc = ck < k0 ? 0 : (ck >= k0 + TZ ? 2 : 1);
        sc[h] = part * 9 + jc * 3 + kc;
        if (sl[h]) ss[h] = atomicAdd(&clsCount[sc[h]], 1u);
      }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      unsigned v = threadIdx.x < 27 ? clsCount[threadIdx.x] : 0u;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(v, o);
        if ((int)threadIdx.x >= o) v += t;
      }
      if (threadIdx.x < 27) clsStart[threadIdx.x + 1] = v;
      if (threadIdx.x == 0) clsStart[0] = 0;
    }
    __syncthreads();
    const int nzs = clsStart[27];  // non-empty segments, in class order
#pragma unroll
    for (int h = 0; h < 2; h++)
      if (sl[h]) {
        const unsigned pos = clsStart[sc[h]] + ss[h];
        segBeg[pos] = sb[h];
        segPre[pos + 1] = sl[h];
      }
    __syncthreads();
    {  // inclusive scan of the segment lengths, two per thread
      const int e0 = 2 * threadIdx.x + 1, e1 = e0 + 1;
      const unsigned a0 = e0 <= nzs ? segPre[e0] : 0u, a1 = e1 <= nzs ? segPre[e1] : 0u;
      unsigned v = a0 + a1;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(v, o);
        if ((int)(threadIdx.x & 63) >= o) v += t;
      }
      if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = v;
      __syncthreads();
      for (int w = 0; w < (int)(threadIdx.x >> 6); w++) v += wsum[w];
      if (e0 <= nzs) segPre[e0] = v - a1;
      if (e1 <= nzs) segPre[e1] = v;
      if (threadIdx.x == 0) segPre[0] = 0;
    }
    __syncthreads();
    const unsigned total = segPre[nzs];
    auto locate = [&](unsigned v) -> unsigned {
      int lo = 0, hi = nzs;  // last segment starting at or before v
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (segPre[mid] <= v) lo = mid; else hi = mid;
      }
      return sorted[segBeg[lo] + (v - segPre[lo])];
    };
    // The binned particles are in cell order, so neighbouring entries share
    // their faces. A wavefront takes every 8th entry of a 256-entry chunk
    // (thread 32a + b takes entry 8b + a) so that its lanes' LDS atomics fall
    // on different faces instead of queueing on the same few.
    const unsigned perm = (threadIdx.x & 31u) * 8u + (threadIdx.x >> 5);
    unsigned vA = perm;
    const unsigned pA = vA < total ? locate(vA) : 0u;
    unsigned pB = vA + 256u < total ? locate(vA + 256u) : 0u;
    float xA = px[pA], yA = py[pA], zA = pz[pA], uA = pvx[pA], vA_ = pvy[pA], wA_ = pvz[pA];
    while (vA < total) {
      // entry 0 stands in for a visit past the end (its loads are harmless)
      const float xB = px[pB], yB = py[pB], zB = pz[pB], uB = pvx[pB], vB_ = pvy[pB], wB_ = pvz[pB];
      const unsigned pC = vA + 512u < total ? locate(vA + 512u) : 0u;
      const float vc[3] = {uA, vA_, wA_};
      // per axis, offset 0 (the component along the axis) and 0.5 (the others)
      int fx[2], fy[2], fz[2];
      float tx_[2], ty_[2], tz_[2];
#pragma unroll
      for (int o = 0; o < 2; o++) {
        const float off = o ? 0.5f : 0.f;
        int f = (int)floorf(xA * g.inv_dx - off);
        if (f < 0) { f = 0; tx_[o] = 0.f; } else if (f > g.nx - 1) { f = g.nx - 1; tx_[o] = 1.f; } else tx_[o] = axis_frac(g, xA, off, f);
        fx[o] = f;
        f = (int)floorf(yA * g.inv_dx - off);
        if (f < 0) { f = 0; ty_[o] = 0.f; } else if (f > g.ny - 2) { f = g.ny - 2; ty_[o] = 1.f; } else ty_[o] = axis_frac(g, yA, off, f);
        fy[o] = f;
        f = (int)floorf(zA * g.inv_dx - off);
        if (f < 0) { f = 0; tz_[o] = 0.f; } else if (f > g.nz - 2) { f = g.nz - 2; tz_[o] = 1.f; } else tz_[o] = axis_frac(g, zA, off, f);
        fz[o] = f;
      }
#pragma unroll
      for (int comp = 0; comp < 3; comp++) {
        const int ox = comp == 0 ? 0 : 1, oy = comp == 1 ? 0 : 1, oz = comp == 2 ? 0 : 1;
        const int fi = fx[ox], fj = fy[oy], fk = fz[oz];
        const int hx = comp == 0 ? uHi : iHiC, hy = comp == 1 ? vHi : jHiC, hz = comp == 2 ? wHi : kHiC;
        const float s0 = 1.f - tx_[ox], s1 = tx_[ox], t0 = 1.f - ty_[oy], t1 = ty_[oy], r0 = 1.f - tz_[oz], r1 = tz_[oz];
        const float wa = s0 * t0, wb = s1 * t0, wc = s0 * t1, wd = s1 * t1;
        if (comp != 0 && fi == g.nx - 1) {
          // v/w stencils clamped at i0 = nx-1 read the next row (the
          // reference's idx+1 at the +x edge, solver.js:765,772)
          const int d0 = g.nx, d1 = comp == 1 ? g.ny + 1 : g.ny;
#pragma unroll
          for (int e = 0; e < 8; e++) {
            const int di = e & 1, dj = (e >> 1) & 1, dk = (e >> 2) & 1;
            const float wt = (di ? (dj ? wd : wb) : (dj ? wc : wa)) * (dk ? r1 : r0);
            int ii = fi + di, jj = fj + dj, kk = fk + dk;
            if (ii >= d0) { ii -= d0; jj += 1; if (jj >= d1) { jj = 0; kk += 1; } }
            if (ii < i0 || ii >= hx || jj < j0 || jj >= hy || kk < k0 || kk >= hz) continue;
            const int l = ((kk - k0) * (TY + (comp == 1)) + (jj - j0)) * (TX + (comp == 0)) + (ii - i0);
            atomicAdd(&accM[comp][l], (unsigned long long)fix(wt * vc[comp]));
            atomicAdd(&accW[comp][l], (unsigned long long)fix(wt));
          }
          continue;
        }
        const unsigned mx = (unsigned)(fi >= i0 && fi < hx) | (unsigned)(fi + 1 >= i0 && fi + 1 < hx) << 1;
        const unsigned my = (unsigned)(fj >= j0 && fj < hy) | (unsigned)(fj + 1 >= j0 && fj + 1 < hy) << 1;
        const unsigned mz = (unsigned)(fk >= k0 && fk < hz) | (unsigned)(fk + 1 >= k0 && fk + 1 < hz) << 1;
        if (!(mx && my && mz)) continue;
        const int SX = TX + (comp == 0), SY = TY + (comp == 1);
        const int lb = ((fk - k0) * SY + (fj - j0)) * SX + (fi - i0);
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int di = e & 1, dj = (e >> 1) & 1, dk = (e >> 2) & 1;
          if (!((mx >> di) & (my >> dj) & (mz >> dk) & 1u)) continue;
          const float wt = (di ? (dj ? wd : wb) : (dj ? wc : wa)) * (dk ? r1 : r0);
          const int l = lb + (dk * SY + dj) * SX + di;
          atomicAdd(&accM[comp][l], (unsigned long long)fix(wt * vc[comp]));
          atomicAdd(&accW[comp][l], (unsigned long long)fix(wt));
        }
      }
      xA = xB; yA = yB; zA = zB; uA = uB; vA_ = vB_; wA_ = wB_;
      pB = pC;
      vA += 256u;
    }
    __syncthreads();
  }
  // normalise and write the owned faces
  for (int comp = 0; comp < 3; comp++) {
    const int d0 = comp == 0 ? g.nx + 1 : g.nx, d1 = comp == 1 ? g.ny + 1 : g.ny;
    const int oi1 = comp == 0 ? uHi : iHiC, oj1 = comp == 1 ? vHi : jHiC, ok1 = comp == 2 ? wHi : kHiC;
    const int ni = oi1 - i0, nj = oj1 - j0, nk = ok1 - k0;
    const long base = comp == 0 ? 0 : (comp == 1 ? g.uN : (long)g.uN + g.vN);
    for (int e = threadIdx.x; e < ni * nj * nk; e += blockDim.x) {
      const int li = e % ni, lj = (e / ni) % nj, lk = e / (ni * nj);
      const long f = base + ((long)(k0 + lk) * d1 + (j0 + lj)) * d0 + (i0 + li);
      float W = 0.f, M = 0.f;
      if (!empty) {
        const int l = (lk * (TY + (comp == 1)) + lj) * (TX + (comp == 0)) + li;
        W = (float)(long long)accW[comp][l] * kFixInvF;
        M = (float)(long long)accM[comp][l] * kFixInvF;
      }
      vel[f] = W > 1e-8f ? M / W : M;
      wgt[f] = W;
    }
  }
}
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_p2g_tile(Geo g, int tilesX, int tilesY, const unsigned *__restrict__ start,
                                                  const unsigned *__restrict__ sorted, const float *__restrict__ px,
                                                  const float *__restrict__ py, const float *__restrict__ pz,
                                                  const float *__restrict__ pvx, const float *__restrict__ pvy,
                                                  const float *__restrict__ pvz, float *__restrict__ vel,
                                                  float *__restrict__ wgt) {
  k_p2g_tile_body<TX, TY, TZ>(blockIdx.x, gridDim.x, g, tilesX, tilesY, start, sorted, px, py, pz, pvx, pvy, pvz, vel,
                              wgt);
}
#define WS_P2G_INST(X, Y, Z)                                                                                   \
  template __global__ void k_p2g_tile<X, Y, Z>(Geo, int, int, const unsigned *, const unsigned *, const float *, \
                                               const float *, const float *, const float *, const float *,      \
                                               const float *, float *, float *);
WS_P2G_INST(8, 8, 8)
WS_P2G_INST(4, 4, 4)
WS_P2G_INST(16, 8, 8)
WS_P2G_INST(8, 8, 4)
#undef WS_P2G_INST
WS_PLANETS_OPS(k_p2g_tile_ops0, (k_p2g_tile<P2G_BINS[0][0], P2G_BINS[0][1], P2G_BINS[0][2]>),
               (k_p2g_tile_body<P2G_BINS[0][0], P2G_BINS[0][1], P2G_BINS[0][2]>))
WS_PLANETS_OPS(k_p2g_tile_ops1, (k_p2g_tile<P2G_BINS[1][0], P2G_BINS[1][1], P2G_BINS[1][2]>),
               (k_p2g_tile_body<P2G_BINS[1][0], P2G_BINS[1][1], P2G_BINS[1][2]>))
WS_PLANETS_OPS(k_p2g_tile_ops2, (k_p2g_tile<P2G_BINS[2][0], P2G_BINS[2][1], P2G_BINS[2][2]>),
               (k_p2g_tile_body<P2G_BINS[2][0], P2G_BINS[2][1], P2G_BINS[2][2]>))
WS_PLANETS_OPS(k_p2g_tile_ops3, (k_p2g_tile<P2G_BINS[3][0], P2G_BINS[3][1], P2G_BINS[3][2]>),
               (k_p2g_tile_body<P2G_BINS[3][0], P2G_BINS[3][1], P2G_BINS[3][2]>))

// ------------------------------------------------- boundary conditions
// Face classification shared by BC, gravity, projection and extrapolation.
struct FaceInfo {
  int comp, i, j, k;
  int cl, cr;  // adjacent cells (-1 if outside)
  bool plane;  // domain boundary plane
};

// (face indices are below 2^31, so the decoding divides in 32-bit unsigned
// arithmetic: a 64-bit division is a long software routine on the GPU)
__device__ __forceinline__ bool face_of(const Geo &g, long f, FaceInfo &fi) {
  const unsigned nx = (unsigned)g.nx, ny = (unsigned)g.ny;
  if (f < g.uN) {
    const unsigned u = (unsigned)f, r = u / (nx + 1u);
    fi.comp = 0;
    fi.i = (int)(u - r * (nx + 1u));
    fi.j = (int)(r % ny);
    fi.k = (int)(r / ny);
    fi.plane = fi.i == 0 || fi.i == g.nx;
    int c = (fi.k * g.ny + fi.j) * g.nx + fi.i;
    fi.cl = fi.i > 0 ? c - 1 : -1;
    fi.cr = fi.i < g.nx ? c : -1;
  } else if (f < (long)g.uN + g.vN) {
    const unsigned u = (unsigned)(f - g.uN), r = u / nx;
    fi.comp = 1;
    fi.i = (int)(u - r * nx);
    fi.j = (int)(r % (ny + 1u));
    fi.k = (int)(r / (ny + 1u));
    fi.plane = fi.j == 0 || fi.j == g.ny;
    int c = (fi.k * g.ny + fi.j) * g.nx + fi.i;
    fi.cl = fi.j > 0 ? c - g.nx : -1;
    fi.cr = fi.j < g.ny ? c : -1;
  } else {
    f -= (long)g.uN + g.vN;
    if (f >= g.wN) return false;
    const unsigned u = (unsigned)f, r = u / nx;
    fi.comp = 2;
    fi.i = (int)(u - r * nx);
    fi.j = (int)(r % ny);
    fi.k = (int)(r / ny);
    fi.plane = fi.k == 0 || fi.k == g.nz;
    int c = (fi.k * g.ny + fi.j) * g.nx + fi.i;
    fi.cl = fi.k > 0 ? c - g.nx * g.ny : -1;
    fi.cr = fi.k < g.nz ? c : -1;
  }
  return true;
}

// _applyBC (solver.js:1130-1171; solid faces take the solid's velocity), FLIP snapshot
// (:2579) and the extrapolation validity mark (:1578-1611: weighted, a
// domain plane or solid-adjacent). With fuse_gravity the radial gravity of
// _applyPlanetGravitySlice (:1229-1256) and the second _applyBC follow in
// the same pass (used when viscosity is off).
// the boundary value of a face next to a solid: the solid cell's velocity
// component (the left / lower cell first), 0 on the domain planes and when
// no moving solid exists (sv == nullptr)
__device__ __forceinline__ float solid_face_value(const Geo &g, const FaceInfo &fi, uint8_t tl,
                                                  const float *__restrict__ sv) {
  if (fi.plane || !sv) return 0.f;
  const float *a = sv + (long)fi.comp * g.nC;
  return tl == SOLID ? a[fi.cl] : a[fi.cr];
}

__device__ __forceinline__ void k_bc_snapshot_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ ct,
                                                   float *__restrict__ vel, float *__restrict__ velO,
                                                   const float *__restrict__ wgt, uint8_t *__restrict__ valid,
                                                   int fuse_gravity, const float *__restrict__ gvec,
                                                   const DevScalars *__restrict__ s, int write_valid,
                                                   const float *__restrict__ sv) {
  long F = (long)g.uN + g.vN + g.wN;
  long f = (long)bx * blockDim.x + threadIdx.x;
  if (f >= F) return;
  FaceInfo fi;
  face_of(g, f, fi);
  uint8_t tl = fi.cl >= 0 ? ct[fi.cl] : SOLID, tr = fi.cr >= 0 ? ct[fi.cr] : SOLID;
  bool bc = fi.plane || tl == SOLID || tr == SOLID;
  float u = bc ? solid_face_value(g, fi, tl, sv) : vel[f];
  velO[f] = u;
  if (write_valid) valid[f] = (wgt[f] > 1e-8f || bc) ? 1 : 0;
  if (fuse_gravity && !bc && s->gs > 0.f && (tl == FLUID || tr == FLUID)) u = u + gvec[f] * s->dt * s->gs;
  vel[f] = u;
}
__global__ void k_bc_snapshot(Geo g, const uint8_t *__restrict__ ct, float *__restrict__ vel,
                              float *__restrict__ velO, const float *__restrict__ wgt,
                              uint8_t *__restrict__ valid, int fuse_gravity, const float *__restrict__ gvec,
                              const DevScalars *__restrict__ s, int write_valid, const float *__restrict__ sv) {
  k_bc_snapshot_body(blockIdx.x, gridDim.x, g, ct, vel, velO, wgt, valid, fuse_gravity, gvec, s, write_valid, sv);
}
WS_PLANETS_OPS(k_bc_snapshot_ops, k_bc_snapshot, k_bc_snapshot_body)

// extrapolation validity alone (solver.js:1578-1611), for the stand-alone
// _extrapolate stage; the substep computes it inside k_bc_snapshot
__global__ void k_valid(Geo g, const uint8_t *__restrict__ ct, const float *__restrict__ wgt,
                        uint8_t *__restrict__ valid) {
  long F = (long)g.uN + g.vN + g.wN;
  long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  FaceInfo fi;
  face_of(g, f, fi);
  uint8_t tl = fi.cl >= 0 ? ct[fi.cl] : SOLID, tr = fi.cr >= 0 ? ct[fi.cr] : SOLID;
  valid[f] = (wgt[f] > 1e-8f || fi.plane || tl == SOLID || tr == SOLID) ? 1 : 0;
}

// radial gravity + _applyBC after viscosity
__device__ __forceinline__ void k_gravity_bc_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ ct,
                                                  float *__restrict__ vel, const float *__restrict__ gvec,
                                                  const DevScalars *__restrict__ s, const float *__restrict__ sv) {
  long F = (long)g.uN + g.vN + g.wN;
  long f = (long)bx * blockDim.x + threadIdx.x;
  if (f >= F) return;
  FaceInfo fi;
  face_of(g, f, fi);
  uint8_t tl = fi.cl >= 0 ? ct[fi.cl] : SOLID, tr = fi.cr >= 0 ? ct[fi.cr] : SOLID;
  bool bc = fi.plane || tl == SOLID || tr == SOLID;
  if (bc) { vel[f] = solid_face_value(g, fi, tl, sv); return; }
  if (s->gs > 0.f && (tl == FLUID || tr == FLUID)) vel[f] = vel[f] + gvec[f] * s->dt * s->gs;
}
__global__ void k_gravity_bc(Geo g, const uint8_t *__restrict__ ct, float *__restrict__ vel,
                             const float *__restrict__ gvec, const DevScalars *__restrict__ s,
                             const float *__restrict__ sv) {
  k_gravity_bc_body(blockIdx.x, gridDim.x, g, ct, vel, gvec, s, sv);
}
WS_PLANETS_OPS(k_gravity_bc_ops, k_gravity_bc, k_gravity_bc_body)

// ------------------------------------------------------------ viscosity
// _viscosity, solver.js:1259-1304, is one in-place lexicographic sweep over
// the fluid-fluid faces: a face reads the already-updated -i/-j/-k
// neighbours. That is a lower-triangular solve, evaluated here exactly (to
// float rounding) by fixed-point iteration: x1 = Jacobi from the old field,
// x(m+1) replaces the lower neighbours by x(m). The error shrinks by 3*coef
// per iteration, so a few iterations reach float resolution (the host picks
// the count from coef).
__device__ __forceinline__ void k_visc_iter_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ ct,
                                                 const float *__restrict__ old, const float *__restrict__ cur,
                                                 float *__restrict__ out, float nu_dx2,
                                                 const DevScalars *__restrict__ s, int first) {
  const float coef = fminf(nu_dx2 * s->dt, 0.24f);  // solver.js:1264
  long F = (long)g.uN + g.vN + g.wN;
  long f = (long)bx * blockDim.x + threadIdx.x;
  if (f >= F) return;
  FaceInfo fi;
  face_of(g, f, fi);
  int d0 = fi.comp == 0 ? g.nx + 1 : g.nx, d1 = fi.comp == 1 ? g.ny + 1 : g.ny, d2 = fi.comp == 2 ? g.nz + 1 : g.nz;
  long s1 = d0, s2 = (long)d0 * d1;
  int nxy = g.nx * g.ny;
  // Every load of the face is issued before any is used (indices that do
  // not exist are replaced by the face itself or cell 0), so they are in
  // flight together; the selections below reproduce the branches exactly.
  const bool ffValid = fi.cl >= 0 && fi.cr >= 0;
  const int cl = ffValid ? fi.cl : 0, cr = ffValid ? fi.cr : 0;
  const bool hasI0 = fi.i - 1 >= 0, hasI1 = fi.i + 1 < d0, hasJ0 = fi.j - 1 >= 0, hasJ1 = fi.j + 1 < d1;
  const bool hasK0 = fi.k - 1 >= 0, hasK1 = fi.k + 1 < d2;
  // the lower neighbour along dir is also a fluid-fluid face: its cells are
  // this face's cells shifted by -1 along dir, except that the lower
  // neighbour along the face normal at index 1 is a domain plane
  const bool lowI = ffValid && hasI0 && !(fi.comp == 0 && fi.i - 1 == 0);
  const bool lowJ = ffValid && hasJ0 && !(fi.comp == 1 && fi.j - 1 == 0);
  const bool lowK = ffValid && hasK0 && !(fi.comp == 2 && fi.k - 1 == 0);
  const float u0 = old[f];
  const uint8_t tl = ct[cl], tr = ct[cr];
  const uint8_t tlI = ct[lowI ? cl - 1 : 0], trI = ct[lowI ? cr - 1 : 0];
  const uint8_t tlJ = ct[lowJ ? cl - g.nx : 0], trJ = ct[lowJ ? cr - g.nx : 0];
  const uint8_t tlK = ct[lowK ? cl - nxy : 0], trK = ct[lowK ? cr - nxy : 0];
  const float oI0 = old[hasI0 ? f - 1 : f], oI1 = old[hasI1 ? f + 1 : f];
  const float oJ0 = old[hasJ0 ? f - s1 : f], oJ1 = old[hasJ1 ? f + s1 : f];
  const float oK0 = old[hasK0 ? f - s2 : f], oK1 = old[hasK1 ? f + s2 : f];
  const float cI0 = cur[hasI0 ? f - 1 : f], cJ0 = cur[hasJ0 ? f - s1 : f], cK0 = cur[hasK0 ? f - s2 : f];
  if (!ffValid || tl != FLUID || tr != FLUID) { out[f] = u0; return; }
  const bool ffI = lowI && tlI == FLUID && trI == FLUID;
  const bool ffJ = lowJ && tlJ == FLUID && trJ == FLUID;
  const bool ffK = lowK && tlK == FLUID && trK == FLUID;
  float val = u0 * 6.f;
  val -= hasI0 ? ((!first && ffI) ? cI0 : oI0) : u0;
  val -= hasI1 ? oI1 : u0;
  val -= hasJ0 ? ((!first && ffJ) ? cJ0 : oJ0) : u0;
  val -= hasJ1 ? oJ1 : u0;
  val -= hasK0 ? ((!first && ffK) ? cK0 : oK0) : u0;
  val -= hasK1 ? oK1 : u0;
  out[f] = u0 + coef * val;
}
__global__ void k_visc_iter(Geo g, const uint8_t *__restrict__ ct, const float *__restrict__ old,
                            const float *__restrict__ cur, float *__restrict__ out, float nu_dx2,
                            const DevScalars *__restrict__ s, int first) {
  k_visc_iter_body(blockIdx.x, gridDim.x, g, ct, old, cur, out, nu_dx2, s, first);
}
WS_PLANETS_OPS(k_visc_iter_ops, k_visc_iter, k_visc_iter_body)

// ------------------------------------------------------ pressure system
// _pressureSolve assembly, solver.js:1322-1394: rhs = -dx^2 div, diag =
// fluid + air neighbours, relaxation weight omega/diag (stored float as the
// reference's pW), fluid-neighbour mask; non-fluid cells hold q = 0.
__device__ __forceinline__ void k_assemble_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ ct,
                                                const float *__restrict__ vel, float *__restrict__ q,
                                                float *__restrict__ q2, float *__restrict__ rhs, float *__restrict__ pw,
                                                uint8_t *__restrict__ nbm, uint8_t *__restrict__ dg, float omega) {
  int c = bx * blockDim.x + threadIdx.x;
  if (c >= g.nC) return;
  if (ct[c] != FLUID) { q[c] = 0.f; q2[c] = 0.f; pw[c] = 0.f; nbm[c] = 0; dg[c] = 0; return; }
  int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny);
  int nxy = g.nx * g.ny;
  int nb[6] = {i > 0 ? c - 1 : -1, i < g.nx - 1 ? c + 1 : -1, j > 0 ? c - g.nx : -1,
               j < g.ny - 1 ? c + g.nx : -1, k > 0 ? c - nxy : -1, k < g.nz - 1 ? c + nxy : -1};
  int Nf = 0, Na = 0;
  unsigned m = 0x40;  // bit 6: fluid cell
  for (int d = 0; d < 6; d++) {
    if (nb[d] < 0) continue;
    uint8_t t = ct[nb[d]];
    if (t == FLUID) { Nf++; m |= 1u << d; }
    else if (t == AIR) Na++;
  }
  const float *u = vel, *v = vel + g.uN, *w = vel + g.uN + g.vN;
  int iu = (k * g.ny + j) * (g.nx + 1) + i, iv = (k * (g.ny + 1) + j) * g.nx + i;
  float dsum = u[iu + 1] - u[iu] + v[iv + g.nx] - v[iv] + w[c + nxy] - w[c];
  float div = dsum / g.dx;
  rhs[c] = -g.dx * g.dx * div;
  int diag = Nf + Na;
  pw[c] = diag > 0 ? omega / (float)diag : 0.f;
  nbm[c] = (uint8_t)m;
  dg[c] = (uint8_t)diag;
}
__global__ void k_assemble(Geo g, const uint8_t *__restrict__ ct, const float *__restrict__ vel,
                           float *__restrict__ q, float *__restrict__ q2, float *__restrict__ rhs,
                           float *__restrict__ pw, uint8_t *__restrict__ nbm, uint8_t *__restrict__ dg,
                           float omega) {
  k_assemble_body(blockIdx.x, gridDim.x, g, ct, vel, q, q2, rhs, pw, nbm, dg, omega);
}
WS_PLANETS_OPS(k_assemble_ops, k_assemble, k_assemble_body)

// One colour of red-black SOR (the relaxation of solver.js:1402): cells with
// (i+j+k)&1 == color. Each thread updates one cell of that colour.
__global__ __launch_bounds__(256) void k_sor(Geo g, float *__restrict__ q, const float *__restrict__ rhs,
                                             const float *__restrict__ pw, const uint8_t *__restrict__ nbm,
                                             int color, float om1) {
  int half = (g.nx + 1) >> 1;
  int b = xcd_block(blockIdx.x, gridDim.x);
  int t = b * blockDim.x + threadIdx.x;
  int rowi = t / half, h = t % half;
  if (rowi >= g.ny * g.nz) return;
  int j = rowi % g.ny, k = rowi / g.ny;
  int i = 2 * h + ((j + k + color) & 1);
  if (i >= g.nx) return;
  int c = rowi * g.nx + i;
  unsigned m = nbm[c];
  if (!(m & 0x40)) return;
  const float rh = rhs[c], w = pw[c], qc = q[c];
  const float sum = nb_sum(q, c, m, g.nx, g.nx * g.ny);
  q[c] = sor_relax(qc, sum, rh, w, om1);
}

// One full red-black sweep per launch, tiled through LDS. A block owns a
// 32x8x8 cell tile: it loads q with a two-cell halo and the coefficients with
// a one-cell halo, relaxes the red cells of the tile and its one-cell ring
// (the ring values equal what the neighbouring blocks compute for them: same
// inputs, same arithmetic), then the black cells of the tile, and writes the
// tile to the other q buffer. Double buffering keeps every block reading the
// previous sweep. Tiles without a fluid cell return at once.
__global__ __launch_bounds__(256) void k_sor_tile(Geo g, int tilesX, int tilesY, const uint8_t *__restrict__ active,
                                                  const float *__restrict__ qin, float *__restrict__ qout,
                                                  const float *__restrict__ rhs, const float *__restrict__ pw,
                                                  const uint8_t *__restrict__ nbm, float om1) {
  constexpr int TX = SOR_TX, TY = SOR_TY, TZ = SOR_TZ;
  constexpr int QX = TX + 4, QY = TY + 4, QZ = TZ + 4, QN = QX * QY * QZ;
  constexpr int CX = TX + 2, CY = TY + 2, CZ = TZ + 2, CN = CX * CY * CZ;
  __shared__ float sq[QN];
  __shared__ float srhs[CN];
  __shared__ float spw[CN];
  __shared__ uint8_t snb[CN];
  int b = xcd_block(blockIdx.x, gridDim.x);
  if (!active[b]) return;
  int tx = b % tilesX, ty = (b / tilesX) % tilesY, tz = b / (tilesX * tilesY);
  int i0 = tx * TX, j0 = ty * TY, k0 = tz * TZ;
  const int nxy = g.nx * g.ny;
  constexpr int B = 8;
  for (int e0 = 0; e0 < QN; e0 += 256 * B) {
    float v[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      int e = e0 + u * 256 + threadIdx.x;
      int li = e % QX, lj = (e / QX) % QY, lk = e / (QX * QY);
      int i = i0 - 2 + li, j = j0 - 2 + lj, k = k0 - 2 + lk;
      bool in = e < QN && i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
      v[u] = in ? qin[in ? k * nxy + j * g.nx + i : 0] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      int e = e0 + u * 256 + threadIdx.x;
      if (e < QN) sq[e] = v[u];
    }
  }
  for (int e0 = 0; e0 < CN; e0 += 256 * B) {
    float r[B], w[B];
    uint8_t m[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      int e = e0 + u * 256 + threadIdx.x;
      int li = e % CX, lj = (e / CX) % CY, lk = e / (CX * CY);
      int i = i0 - 1 + li, j = j0 - 1 + lj, k = k0 - 1 + lk;
      bool in = e < CN && i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
      int c = in ? k * nxy + j * g.nx + i : 0;
      r[u] = in ? rhs[c] : 0.f;
      w[u] = in ? pw[c] : 0.f;
      m[u] = in ? nbm[c] : 0;
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      int e = e0 + u * 256 + threadIdx.x;
      if (e < CN) { srhs[e] = r[u]; spw[e] = w[u]; snb[e] = m[u]; }
    }
  }
  __syncthreads();
  // red cells ((i+j+k) even) of the tile and its ring, then black cells of the tile
  for (int color = 0; color < 2; color++) {
    int lo = color == 0 ? 0 : 1, hiX = color == 0 ? CX : CX - 1, hiY = color == 0 ? CY : CY - 1;
    int hiZ = color == 0 ? CZ : CZ - 1;
    int nX = hiX - lo, nY = hiY - lo, nZ = hiZ - lo;
    for (int e = threadIdx.x; e < nX * nY * nZ; e += 256) {
      int ci = lo + e % nX, cj = lo + (e / nX) % nY, ck = lo + e / (nX * nY);
      int i = i0 - 1 + ci, j = j0 - 1 + cj, k = k0 - 1 + ck;
      if (((i + j + k) & 1) != color) continue;
      int cc = (ck * CY + cj) * CX + ci;
      unsigned msk = snb[cc];
      if (!(msk & 0x40)) continue;
      int qi = ((ck + 1) * QY + (cj + 1)) * QX + (ci + 1);
      float sum = 0.f;
      if (msk & 1) sum += sq[qi - 1];
      if (msk & 2) sum += sq[qi + 1];
      if (msk & 4) sum += sq[qi - QX];
      if (msk & 8) sum += sq[qi + QX];
      if (msk & 16) sum += sq[qi - QX * QY];
      if (msk & 32) sum += sq[qi + QX * QY];
      sq[qi] = sor_relax(sq[qi], sum, srhs[cc], spw[cc], om1);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < TX * TY * TZ; e += 256) {
    int li = e % TX, lj = (e / TX) % TY, lk = e / (TX * TY);
    int i = i0 + li, j = j0 + lj, k = k0 + lk;
    if (i < g.nx && j < g.ny && k < g.nz) qout[k * nxy + j * g.nx + i] = sq[((lk + 2) * QY + (lj + 2)) * QX + (li + 2)];
  }
}

// ---- fluid-cell lists for the per-colour sweeps
// A planet's fluid is a thin shell (about an eighth of the cells at 128^3),
// so the sweeps run over compact lists of the fluid cells of each colour
// instead of the whole grid. Cells keep their index order within a list;
// the relaxation of one colour reads only the other colour, so the result
// equals the whole-grid per-colour sweep bit for bit.

// scan keys: 1 for a red fluid cell, 2^32 for a black one (one 64-bit
// exclusive scan then gives both list positions)
__device__ __forceinline__ void k_sor_count_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ nbm,
                                                 unsigned long long *__restrict__ key) {
  int c = bx * blockDim.x + threadIdx.x;
  if (c > g.nC) return;
  unsigned long long v = 0;
  if (c < g.nC && (nbm[c] & 0x40)) {
    int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny);
    v = ((i + j + k) & 1) ? (1ull << 32) : 1ull;
  }
  key[c] = v;
}
__global__ void k_sor_count(Geo g, const uint8_t *__restrict__ nbm, unsigned long long *__restrict__ key) {
  k_sor_count_body(blockIdx.x, gridDim.x, g, nbm, key);
}
WS_PLANETS_OPS(k_sor_count_ops, k_sor_count, k_sor_count_body)

__device__ __forceinline__ void k_sor_compact_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ nbm,
                                                   const uint8_t *__restrict__ dg, const float *__restrict__ rhs,
                                                   const float *__restrict__ pw,
                                                   const unsigned long long *__restrict__ off, long half,
                                                   unsigned *__restrict__ code, float *__restrict__ lrhs,
                                                   float *__restrict__ lpw, uint8_t *__restrict__ ldg, DevScalars *ds) {
  int c = bx * blockDim.x + threadIdx.x;
  if (c > g.nC) return;
  if (c == g.nC) {
    unsigned long long t = off[c];
    ds->nRed = (unsigned)(t & 0xffffffffull);
    ds->nBlack = (unsigned)(t >> 32);
    return;
  }
  unsigned m = nbm[c];
  if (!(m & 0x40)) return;
  int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny);
  unsigned long long o = off[c];
  long pos = ((i + j + k) & 1) ? half + (long)(o >> 32) : (long)(o & 0xffffffffull);
  code[pos] = (unsigned)c | ((m & 63u) << 26);
  lrhs[pos] = rhs[c];
  lpw[pos] = pw[c];
  ldg[pos] = dg[c];
}
__global__ void k_sor_compact(Geo g, const uint8_t *__restrict__ nbm, const uint8_t *__restrict__ dg,
                              const float *__restrict__ rhs, const float *__restrict__ pw,
                              const unsigned long long *__restrict__ off, long half, unsigned *__restrict__ code,
                              float *__restrict__ lrhs, float *__restrict__ lpw, uint8_t *__restrict__ ldg,
                              DevScalars *ds) {
  k_sor_compact_body(blockIdx.x, gridDim.x, g, nbm, dg, rhs, pw, off, half, code, lrhs, lpw, ldg, ds);
}
WS_PLANETS_OPS(k_sor_compact_ops, k_sor_compact, k_sor_compact_body)

// one colour: q[c] = q[c](1 - omega) + omega/diag (rhs + sum of fluid
// neighbours), the arithmetic of k_sor
__global__ __launch_bounds__(256) void k_sor_list(Geo g, int n, const unsigned *__restrict__ code,
                                                  const float *__restrict__ lrhs, const float *__restrict__ lpw,
                                                  float *__restrict__ q, float om1) {
  int b = xcd_block(blockIdx.x, gridDim.x);
  int t = b * blockDim.x + threadIdx.x;
  if (t >= n) return;
  // the code, rhs and weight in one round trip, then the cell and its six
  // neighbours in one more
  const unsigned e = code[t];
  const float rh = lrhs[t], w = lpw[t];
  const int c = (int)(e & 0x3ffffffu);
  const unsigned m = e >> 26;
  const float qc = q[c];
  const float sum = nb_sum(q, c, m, g.nx, g.nx * g.ny);
  q[c] = sor_relax(qc, sum, rh, w, om1);
}

// ||rhs - A q||^2 and ||rhs||^2 over the two lists (red, then black), FP64
// block partials over a fixed grid, as k_residual
__global__ __launch_bounds__(256) void k_residual_list(Geo g, const DevScalars *__restrict__ ds, long half,
                                                       const unsigned *__restrict__ code,
                                                       const float *__restrict__ lrhs,
                                                       const uint8_t *__restrict__ ldg, const float *__restrict__ q,
                                                       double *__restrict__ partials) {
  __shared__ double sr[256], sb[256];
  long nR = ds->nRed, n = nR + ds->nBlack;
  double rr = 0, bb = 0;
  int nxy = g.nx * g.ny;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    long pos = t < nR ? t : half + (t - nR);
    unsigned e = code[pos];
    int c = (int)(e & 0x3ffffffu);
    unsigned m = e >> 26;
    const float sum = nb_sum(q, c, m, g.nx, nxy);
    double rh = (double)lrhs[pos];
    double r = rh - ((double)ldg[pos] * (double)q[c] - (double)sum);
    rr += r * r;
    bb += rh * rh;
  }
  sr[threadIdx.x] = rr;
  sb[threadIdx.x] = bb;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { sr[threadIdx.x] += sr[threadIdx.x + o]; sb[threadIdx.x] += sb[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { partials[2 * blockIdx.x] = sr[0]; partials[2 * blockIdx.x + 1] = sb[0]; }
}

// per SOR tile: does it hold a fluid cell (the tile sweep skips the rest)
__global__ void k_sor_tiles_mark(Geo g, int tilesX, int tilesY, int nTiles, const uint8_t *__restrict__ nbm,
                                 uint8_t *__restrict__ active) {
  int b = blockIdx.x;
  if (b >= nTiles) return;
  int tx = b % tilesX, ty = (b / tilesX) % tilesY, tz = b / (tilesX * tilesY);
  int i0 = tx * SOR_TX, j0 = ty * SOR_TY, k0 = tz * SOR_TZ;
  int any = 0;
  for (int e = threadIdx.x; e < SOR_TX * SOR_TY * SOR_TZ; e += blockDim.x) {
    int i = i0 + e % SOR_TX, j = j0 + (e / SOR_TX) % SOR_TY, k = k0 + e / (SOR_TX * SOR_TY);
    if (i < g.nx && j < g.ny && k < g.nz && (nbm[(k * g.ny + j) * g.nx + i] & 0x40)) any = 1;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) active[b] = (uint8_t)any;
}

// S red-black sweeps per launch over a 32x16x16 cell tile (temporal
// blocking). The workgroup loads q with a halo of 4 cells into LDS and runs
// the colour phases on shrinking regions: the red phase of sweep t covers
// the tile plus a ring of 2(S-1-t)+1 cells, the black phase a ring of
// 2(S-1-t). Every ring cell is recomputed from exactly the inputs its owning
// tile uses, so the tile ends with the values of S global sweeps, bit for
// bit, and is written to the other q buffer (every workgroup reads the
// previous state). Tiles without a fluid cell return at once.
// The fluid cells of each phase are listed per tile once per solve
// (k_sor_tb_build) with their LDS index, neighbour mask, diagonal count and
// rhs, so a launch streams its lists: each thread takes list entries
// t, t + 512, ...; a phase's entries are loaded during the phase before,
// and a phase does all of a thread's LDS reads before its writes (cells of
// one colour never neighbour each other).
namespace {
constexpr int TB_QX = SOR_TB_X + 2 * TB_H, TB_QY = SOR_TB_Y + 2 * TB_H, TB_QZ = SOR_TB_Z + 2 * TB_H;
constexpr int TB_QN = TB_QX * TB_QY * TB_QZ;
static_assert(TB_QN < (1 << 15), "LDS index in 15 bits");

template <int PH>
struct TbPhase {
  static constexpr int NT = SOR_TB_THREADS, U = (tb_items(PH) + NT - 1) / NT;
  unsigned pk[U];  // LDS index (15 bits) | neighbour mask << 15 (6) | diagonal count << 21 (3)
  float rh[U];
  int n;
  __device__ __forceinline__ void load(const unsigned *__restrict__ lpk, const float *__restrict__ lr,
                                       const int *__restrict__ cnt, int tile) {
    n = cnt[4 * tile + PH];
    const long base = (long)tile * TB_CAP + tb_off(PH);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = u * NT + (int)threadIdx.x;
      const long at = base + (e < n ? e : 0);
      pk[u] = lpk[at];
      rh[u] = lr[at];
    }
  }
  __device__ __forceinline__ void relax(float *sq, const float *spw, float om1) {
    float nv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (u * NT >= n) break;  // uniform: no entry of this slot in the workgroup
      const int e = u * NT + (int)threadIdx.x;
      const unsigned w = pk[u];
      // an empty slot reads a harmless interior cell
      const int qi = e < n ? (int)(w & 0x7fffu) : (TB_H * TB_QY + TB_H) * TB_QX + TB_H;
      const float c0 = sq[qi];
      const float a = sq[qi - 1], b = sq[qi + 1], cc = sq[qi - TB_QX], dd = sq[qi + TB_QX];
      const float ee = sq[qi - TB_QX * TB_QY], ff = sq[qi + TB_QX * TB_QY];
      const float pw = spw[(w >> 21) & 7u];
      // the sum of k_sor_list (an absent neighbour adds +0: the sum never
      // holds -0, so the bits are those of skipping it)
      const unsigned m = w >> 15;
      float sum = 0.f;
      sum += (m & 1u) ? a : 0.f;
      sum += (m & 2u) ? b : 0.f;
      sum += (m & 4u) ? cc : 0.f;
      sum += (m & 8u) ? dd : 0.f;
      sum += (m & 16u) ? ee : 0.f;
      sum += (m & 32u) ? ff : 0.f;
      nv[u] = sor_relax(c0, sum, rh[u], pw, om1);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (u * NT >= n) break;
      if (u * NT + (int)threadIdx.x < n) sq[pk[u] & 0x7fffu] = nv[u];
    }
  }
};
}  // namespace

template <int S>
__global__ __launch_bounds__(SOR_TB_THREADS) void k_sor_tb(Geo g, int tilesX, int tilesY,
                                                           const uint8_t *__restrict__ active,
                                                           const float *__restrict__ qin, float *__restrict__ qout,
                                                           const unsigned *__restrict__ lpk,
                                                           const float *__restrict__ lr, const int *__restrict__ cnt,
                                                           float omega, float om1) {
  static_assert(S == 1 || S == 2, "one or two sweeps per launch");
  constexpr int NT = SOR_TB_THREADS, TX = SOR_TB_X, TY = SOR_TB_Y, TZ = SOR_TB_Z, H = TB_H;
  constexpr int QX = TB_QX, QY = TB_QY, QN = TB_QN;
  __shared__ float sq[QN];
  __shared__ float spw[8];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  if (!active[b]) return;
  const int tx = b % tilesX, ty = (b / tilesX) % tilesY, tz = b / (tilesX * tilesY);
  const int i0 = tx * TX, j0 = ty * TY, k0 = tz * TZ;
  const int nxy = g.nx * g.ny;
  // the relaxation weights omega / diag as k_assemble rounds them
  if (threadIdx.x < 8) spw[threadIdx.x] = threadIdx.x >= 1 && threadIdx.x <= 6 ? omega / (float)threadIdx.x : 0.f;
  // a one-sweep launch runs the last two phases of a two-sweep one
  TbPhase<S == 2 ? 0 : 2> pa;
  TbPhase<S == 2 ? 1 : 3> pb;
  TbPhase<2> pc;
  TbPhase<3> pd;
  {
    constexpr int BQ = (QN + NT - 1) / NT;
    float v[BQ];
#pragma unroll
    for (int u = 0; u < BQ; u++) {
      const int e = u * NT + (int)threadIdx.x;
      const int li = e % QX, lj = (e / QX) % QY, lk = e / (QX * QY);
      const int i = i0 - H + li, j = j0 - H + lj, k = k0 - H + lk;
      const bool in = e < QN && i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
      const float x = qin[in ? k * nxy + j * g.nx + i : 0];
      v[u] = in ? x : 0.f;
    }
    pa.load(lpk, lr, cnt, b);
#pragma unroll
    for (int u = 0; u < BQ; u++) {
      const int e = u * NT + (int)threadIdx.x;
      if (e < QN) sq[e] = v[u];
    }
  }
  __syncthreads();
  // each phase's list entries are loaded before the previous phase's
  // relaxation, which hides them
  pb.load(lpk, lr, cnt, b);
  pa.relax(sq, spw, om1);
  __syncthreads();
  if constexpr (S == 2) pc.load(lpk, lr, cnt, b);
  pb.relax(sq, spw, om1);
  __syncthreads();
  if constexpr (S == 2) {
    pd.load(lpk, lr, cnt, b);
    pc.relax(sq, spw, om1);
    __syncthreads();
    pd.relax(sq, spw, om1);
    __syncthreads();
  }
  for (int e = threadIdx.x; e < TX * TY * TZ; e += NT) {
    const int li = e % TX, lj = (e / TX) % TY, lk = e / (TX * TY);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    if (i < g.nx && j < g.ny && k < g.nz)
      qout[k * nxy + j * g.nx + i] = sq[((lk + H) * QY + (lj + H)) * QX + (li + H)];
  }
}
template __global__ void k_sor_tb<1>(Geo, int, int, const uint8_t *, const float *, float *, const unsigned *,
                                     const float *, const int *, float, float);
template __global__ void k_sor_tb<2>(Geo, int, int, const uint8_t *, const float *, float *, const unsigned *,
                                     const float *, const int *, float, float);

// The per-tile phase lists of the temporal-blocking sweep, built once per
// solve: phase ph takes the fluid cells of its colour in the tile plus a
// ring of 3 - ph cells, in row order (a stream compaction: a ballot per
// wavefront, the wavefront totals scanned in LDS).
__global__ __launch_bounds__(SOR_TB_THREADS) void k_sor_tb_build(Geo g, int tilesX, int tilesY,
                                                                 const uint8_t *__restrict__ active,
                                                                 const uint8_t *__restrict__ nbm,
                                                                 const uint8_t *__restrict__ dg,
                                                                 const float *__restrict__ rhs,
                                                                 unsigned *__restrict__ lpk, float *__restrict__ lr,
                                                                 int *__restrict__ cnt) {
  constexpr int NT = SOR_TB_THREADS, NW = NT / 64, H = TB_H;
  __shared__ int wsum[NW];
  __shared__ int run;
  const int b = blockIdx.x;
  if (!active[b]) {
    if (threadIdx.x < 4) cnt[4 * b + threadIdx.x] = 0;
    return;
  }
  const int tx = b % tilesX, ty = (b / tilesX) % tilesY, tz = b / (tilesX * tilesY);
  const int i0 = tx * SOR_TB_X, j0 = ty * SOR_TB_Y, k0 = tz * SOR_TB_Z;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int ph = 0; ph < 4; ph++) {
    const int ring = tb_ring(ph), color = ph & 1;
    const int nX = SOR_TB_X + 2 * ring, nY = SOR_TB_Y + 2 * ring, nZ = SOR_TB_Z + 2 * ring;
    const int hx = (nX + 1) / 2, items = hx * nY * nZ;
    const long base = (long)b * TB_CAP + tb_off(ph);
    if (threadIdx.x == 0) run = 0;
    __syncthreads();
    for (int e0 = 0; e0 < items; e0 += NT) {
      const int e = e0 + threadIdx.x;
      const int x2 = e % hx, row = e / hx;
      const int ly = row % nY - ring, lz = row / nY - ring;
      const int gj = j0 + ly, gk = k0 + lz;
      const int first = ((i0 - ring + gj + gk) & 1) == color ? 0 : 1;
      const int lx = -ring + first + 2 * x2;
      const int gi = i0 + lx;
      const bool in = e < items && lx < SOR_TB_X + ring && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny &&
                      gk >= 0 && gk < g.nz;
      const int c = in ? (gk * g.ny + gj) * g.nx + gi : 0;
      const unsigned m = in ? nbm[c] : 0u;
      const bool fl = (m & 0x40u) != 0;
      const unsigned long long bal = __ballot(fl);
      const int before = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) wsum[wv] = __popcll(bal);
      __syncthreads();
      int off = run;
      for (int w = 0; w < wv; w++) off += wsum[w];
      if (fl) {
        const long at = base + off + before;
        lpk[at] = (unsigned)(((lz + H) * TB_QY + (ly + H)) * TB_QX + (lx + H)) | ((m & 63u) << 15) |
                  ((unsigned)dg[c] << 21);
        lr[at] = rhs[c];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < NW; w++) t += wsum[w];
        run += t;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) cnt[4 * b + ph] = run;
    __syncthreads();
  }
}

// per temporal-blocking tile: does it hold a fluid cell
__global__ void k_sor_tb_mark(Geo g, int tilesX, int tilesY, int nTiles, const uint8_t *__restrict__ nbm,
                              uint8_t *__restrict__ active) {
  int b = blockIdx.x;
  if (b >= nTiles) return;
  int tx = b % tilesX, ty = (b / tilesX) % tilesY, tz = b / (tilesX * tilesY);
  int i0 = tx * SOR_TB_X, j0 = ty * SOR_TB_Y, k0 = tz * SOR_TB_Z;
  int any = 0;
  for (int e = threadIdx.x; e < SOR_TB_X * SOR_TB_Y * SOR_TB_Z; e += blockDim.x) {
    int i = i0 + e % SOR_TB_X, j = j0 + (e / SOR_TB_X) % SOR_TB_Y, k = k0 + e / (SOR_TB_X * SOR_TB_Y);
    if (i < g.nx && j < g.ny && k < g.nz && (nbm[(k * g.ny + j) * g.nx + i] & 0x40)) any = 1;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) active[b] = (uint8_t)any;
}

// ||rhs - A q||^2 and ||rhs||^2 over fluid cells, FP64 block partials
__global__ __launch_bounds__(256) void k_residual(Geo g, const float *__restrict__ q, const float *__restrict__ rhs,
                                                  const uint8_t *__restrict__ nbm, const uint8_t *__restrict__ dg,
                                                  double *__restrict__ partials) {
  __shared__ double sr[256], sb[256];
  double rr = 0, bb = 0;
  int nxy = g.nx * g.ny;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nC; c += gridDim.x * blockDim.x) {
    unsigned m = nbm[c];
    if (!(m & 0x40)) continue;
    const float sum = nb_sum(q, c, m, g.nx, nxy);
    double r = (double)rhs[c] - ((double)dg[c] * (double)q[c] - (double)sum);
    rr += r * r;
    bb += (double)rhs[c] * (double)rhs[c];
  }
  sr[threadIdx.x] = rr;
  sb[threadIdx.x] = bb;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { sr[threadIdx.x] += sr[threadIdx.x + o]; sb[threadIdx.x] += sb[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { partials[2 * blockIdx.x] = sr[0]; partials[2 * blockIdx.x + 1] = sb[0]; }
}

// fixed-order sum of the block partials (deterministic)
// With hs set, the scalars are also published to the pinned host mirror
// hs (as k_publish_scalars does): the other words copied from s, the two
// sums from LDS (each word written once), then the sequence word.
__global__ void k_residual_final(const double *__restrict__ partials, int n, DevScalars *s, DevScalars *hs,
                                 unsigned *seqWord, unsigned seq) {
  __shared__ double sr[256], sb[256];
  double rr = 0, bb = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) { rr += partials[2 * i]; bb += partials[2 * i + 1]; }
  sr[threadIdx.x] = rr;
  sb[threadIdx.x] = bb;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { sr[threadIdx.x] += sr[threadIdx.x + o]; sb[threadIdx.x] += sb[threadIdx.x + o]; }
    __syncthreads();
  }
  if (!hs) {
    if (threadIdx.x == 0) { s->rr = sr[0]; s->bb = sb[0]; }
    return;
  }
  constexpr int W = (int)(sizeof(DevScalars) / 4);
  constexpr int RR = (int)(offsetof(DevScalars, rr) / 4), BB = (int)(offsetof(DevScalars, bb) / 4);
  const unsigned *src = (const unsigned *)s;
  unsigned *dst = (unsigned *)hs;
  for (int i = threadIdx.x; i < W; i += blockDim.x)
    if ((i < RR || i >= RR + 2) && (i < BB || i >= BB + 2)) dst[i] = src[i];
  if (threadIdx.x == 0) {
    s->rr = sr[0];
    s->bb = sb[0];
    hs->rr = sr[0];
    hs->bb = sb[0];
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(seqWord, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ------------------------------------------------ on-device pressure solve
// The whole to-residual solve of st_pressure_solve on the device: the
// red-black list sweeps (the arithmetic of k_sor_list), the FP64 residual
// of k_residual_list, and the host loop's stopping rule (a first batch of
// three quarters of the previous solve's sweeps, then checks spaced by the
// observed rate, stop on the tolerance, the sweep cap or twelve checks
// without a 0.5 % gain). A
// solve takes G workgroups (G = 1: k_sor_solve, one context; k_sor_solve_batch
// runs many planets' solves in one launch, G per planet). Workgroup w of G
// relaxes the w-th slice of each colour list; the G workgroups of a planet
// meet at a counter barrier after every half-sweep (agent-scope release /
// acquire, a bounded number of polls; a planet whose barrier times out is
// flagged and its solve abandoned).
// Cells of one colour never neighbour each other, so a thread loads the
// inputs of kSorU cells before it writes any of them; the values are those
// k_sor_list computes, whatever G and the workgroup size.
// The residual is summed in a fixed order that depends on neither: chunks of
// kSorChunk list entries, each summed by 256 lanes (lane l takes entries l,
// l + 256, ...; a wavefront butterfly, then the four wavefronts in order),
// the chunk sums added in chunk order.
constexpr int kSorU = 8;
constexpr int kSorChunk = 2048;
constexpr int kSorPolls = 1 << 20;

// q as the solve reads it. L (the planet's workgroups share one XCD's L2):
// an agent-scope load, which skips the compute unit's L1 (it may hold q of
// an earlier half-sweep) and reads that L2, where every workgroup of the
// planet has stored its values.
template <bool L>
__device__ __forceinline__ float ldq(const float *p) {
  if constexpr (L) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}

// nb_sum with the loads of ldq<L> (same order of additions, same bits)
template <bool L>
__device__ __forceinline__ float nb_sum_l(const float *q, int c, unsigned m, int nx, int nxy) {
  if constexpr (!L) return nb_sum(q, c, m, nx, nxy);
  const float a = ldq<L>(q + ((m & 1u) ? c - 1 : c)), b = ldq<L>(q + ((m & 2u) ? c + 1 : c));
  const float cc = ldq<L>(q + ((m & 4u) ? c - nx : c)), d = ldq<L>(q + ((m & 8u) ? c + nx : c));
  const float e = ldq<L>(q + ((m & 16u) ? c - nxy : c)), f = ldq<L>(q + ((m & 32u) ? c + nxy : c));
  float sum = 0.f;
  sum += (m & 1u) ? a : 0.f;
  sum += (m & 2u) ? b : 0.f;
  sum += (m & 4u) ? cc : 0.f;
  sum += (m & 8u) ? d : 0.f;
  sum += (m & 16u) ? e : 0.f;
  sum += (m & 32u) ? f : 0.f;
  return sum;
}

// the G workgroups of a planet wait for each other (barrier number k).
// Across XCDs (L false) the arrival is an agent-scope release and the wait
// an acquire: the XCD's L2 writes its dirty lines back and drops its copies.
// Within one XCD (L true) every thread waits for its own stores to reach the
// shared L2 and the counter is relaxed; the loads after the barrier read
// that L2 (ldq<true>), so nothing is written back or dropped.
template <bool L>
__device__ __forceinline__ bool sor_barrier(const SorJob &J, unsigned k) {
  if constexpr (L) __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  __shared__ int ok;
  if (threadIdx.x == 0) {
    ok = 1;
    if (J.G > 1) {
      constexpr int order = L ? __ATOMIC_RELAXED : __ATOMIC_RELEASE;
      constexpr int wait = L ? __ATOMIC_RELAXED : __ATOMIC_ACQUIRE;
      __hip_atomic_fetch_add(J.arrive, 1u, order, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned want = k * (unsigned)J.G;
      int polls = 0;
      while (__hip_atomic_load(J.arrive, wait, __HIP_MEMORY_SCOPE_AGENT) < want) {
        if (__hip_atomic_load(J.arrive + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++polls > kSorPolls) {
          __hip_atomic_store(J.arrive + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
  }
  __syncthreads();
  return ok != 0;
}

template <int NT, bool L>
__device__ __forceinline__ void sor_half(const SorJob &J, long base, long lo, long hi, int nxy) {
  const int nx = J.g.nx;
  for (long t0 = lo + threadIdx.x; t0 < hi; t0 += (long)NT * kSorU) {
    unsigned e[kSorU];
    float rh[kSorU], pw[kSorU], sum[kSorU], qc[kSorU];
    // every load unconditional, so that a thread's loads are in flight
    // together: a slot past hi reads entry lo (its result is not written)
#pragma unroll
    for (int u = 0; u < kSorU; u++) {
      const long t = t0 + (long)u * NT;
      const long at = base + (t < hi ? t : lo);
      e[u] = J.code[at];
      rh[u] = J.lrhs[at];
      pw[u] = J.lpw[at];
    }
#pragma unroll
    for (int u = 0; u < kSorU; u++) {
      const int c = (int)(e[u] & 0x3ffffffu);
      sum[u] = nb_sum_l<L>(J.q, c, e[u] >> 26, nx, nxy);
      qc[u] = ldq<L>(J.q + c);
    }
#pragma unroll
    for (int u = 0; u < kSorU; u++)
      if (t0 + (long)u * NT < hi) J.q[e[u] & 0x3ffffffu] = sor_relax(qc[u], sum[u], rh[u], pw[u], J.om1);
  }
}

// one sweep (red, then black) by workgroup w of the planet's G
template <int NT, bool L>
__device__ __forceinline__ bool sor_sweep(const SorJob &J, int w, long nR, long nB, int nxy, unsigned &bar) {
  sor_half<NT, L>(J, 0, nR * w / J.G, nR * (w + 1) / J.G, nxy);
  if (!sor_barrier<L>(J, ++bar)) return false;
  sor_half<NT, L>(J, J.half, nB * w / J.G, nB * (w + 1) / J.G, nxy);
  return sor_barrier<L>(J, ++bar);
}

// sqrt(rr) / sqrt(bb) over both lists; every thread of every workgroup of
// the planet returns the same value (false: a barrier timed out)
template <int NT, bool L>
__device__ bool sor_residual(const SorJob &J, int w, long nR, long nB, int nxy, unsigned &bar, double *res) {
  __shared__ double wr[NT / 64], wb[NT / 64], tot;
  const int nx = J.g.nx;
  const long n = nR + nB;
  const long nChunks = (n + kSorChunk - 1) / kSorChunk;
  const int grp = threadIdx.x >> 8, lane = threadIdx.x & 255, wave = threadIdx.x >> 6;
  for (long ch0 = (long)w * (NT / 256); ch0 < nChunks; ch0 += (long)J.G * (NT / 256)) {
    const long ch = ch0 + grp;
    double rr = 0, bb = 0;
    if (ch < nChunks) {
      const long end = min(n, (ch + 1) * kSorChunk);
      for (long t = ch * kSorChunk + lane; t < end; t += 256) {
        long pos = t < nR ? t : J.half + (t - nR);
        unsigned e = J.code[pos];
        int c = (int)(e & 0x3ffffffu);
        const float a = nb_sum_l<L>(J.q, c, e >> 26, nx, nxy);
        double rh = (double)J.lrhs[pos];
        double r = rh - ((double)J.ldg[pos] * (double)ldq<L>(J.q + c) - (double)a);
        rr += r * r;
        bb += rh * rh;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      rr += __shfl_xor(rr, o);
      bb += __shfl_xor(bb, o);
    }
    if ((threadIdx.x & 63) == 0) {
      wr[wave] = rr;
      wb[wave] = bb;
    }
    __syncthreads();
    if (lane == 0 && ch < nChunks) {
      const int w0 = grp * 4;
      J.chunk[2 * ch] = ((wr[w0] + wr[w0 + 1]) + wr[w0 + 2]) + wr[w0 + 3];
      J.chunk[2 * ch + 1] = ((wb[w0] + wb[w0 + 1]) + wb[w0 + 2]) + wb[w0 + 3];
    }
    __syncthreads();
  }
  if (!sor_barrier<L>(J, ++bar)) return false;
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (long ch = 0; ch < nChunks; ch++) {
      a += __hip_atomic_load(J.chunk + 2 * ch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      b += __hip_atomic_load(J.chunk + 2 * ch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    tot = b > 0 ? sqrt(a) / sqrt(b) : 0.0;
  }
  __syncthreads();
  *res = tot;
  // the chunk sums are rewritten by the next evaluation only after every
  // workgroup has read them
  return sor_barrier<L>(J, ++bar);
}

// the solve by workgroup w of G; bar: barriers the workgroups have already
// passed in this launch
template <int NT, bool L>
__device__ void sor_solve_part(const SorJob &J, int w, unsigned bar) {
  const long nR = J.ds->nRed, nB = J.ds->nBlack;
  const int nxy = J.g.nx * J.g.ny;
  unsigned done = 0, checks = 0;
  double res = 0;
  bool ok = true;
  if (J.tol <= 0) {
    for (int s = 0; s < J.fixedIters && ok; s++) ok = sor_sweep<NT, L>(J, w, nR, nB, nxy, bar);
    done = (unsigned)max(J.fixedIters, 0);
    if (ok && J.measure) { ok = sor_residual<NT, L>(J, w, nR, nB, nxy, bar, &res); checks++; }
  } else {
    const unsigned last = J.ds->sorIters;  // the previous solve's sweeps (rewritten at the end)
    ok = sor_residual<NT, L>(J, w, nR, nB, nxy, bar, &res);
    checks++;
    double best = res, prev = res;
    int flat = 0, k = J.every;
    if (J.adaptive && last > 0) k = max(k, (int)(0.75 * last));
    while (ok && res > J.tol && (int)done < J.maxIters && flat < 12) {
      if ((int)done + k > J.maxIters) k = J.maxIters - (int)done;
      for (int s = 0; s < k && ok; s++) ok = sor_sweep<NT, L>(J, w, nR, nB, nxy, bar);
      done += (unsigned)k;
      if (!ok) break;
      ok = sor_residual<NT, L>(J, w, nR, nB, nxy, bar, &res);
      checks++;
      if (res < 0.995 * best) { best = res; flat = 0; } else flat++;
      int kn = J.every;
      if (J.adaptive && res < prev && res > J.tol) {
        double rate = pow(res / prev, 1.0 / k);  // per sweep
        double need = log(J.tol / res) / log(rate);
        kn = need < 4 ? 4 : (need > 64 ? 64 : (int)ceil(need));
      }
      prev = res;
      k = kn;
    }
  }
  if (w == 0 && threadIdx.x == 0) {
    J.ds->sorIters = done;
    J.ds->sorChecksLast = checks;
    J.ds->sorRes = ok ? res : -1.0;
    J.ds->sorSweeps += done;
    J.ds->sorChecks += checks;
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_sor_solve(SorJob J) {
  sor_solve_part<NT, false>(J, 0, 0);
}

// block b works on planet blk[b].x as its workgroup blk[b].y (x < 0: an
// empty slot of the layout). A planet whose workgroups were laid out on one
// XCD (J.place) checks that they run there: each workgroup writes its XCD id
// (hardware register XCC_ID) before the first barrier and reads the others'
// after it, so all G decide alike; they then take the L2-local barriers, and
// the others the agent-scope ones. Both give the same bits.
template <int NT>
__global__ __launch_bounds__(NT) void k_sor_solve_batch(const SorJob *__restrict__ jobs, const int2 *__restrict__ blk) {
  const int2 pw = blk[blockIdx.x];
  if (pw.x < 0) return;
  const SorJob J = jobs[pw.x];
  if (!J.place || J.G < 2) {
    sor_solve_part<NT, false>(J, pw.y, 0);
    return;
  }
  __shared__ int local;
  if (threadIdx.x == 0) {
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);  // HW_REG_XCC_ID, bits 3:0
    __hip_atomic_store(J.place + pw.y, xcc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const bool ok = sor_barrier<false>(J, 1);
  if (threadIdx.x == 0) {
    int same = 1;
    const unsigned x0 = __hip_atomic_load(J.place, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int v = 1; v < J.G; v++)
      same &= __hip_atomic_load(J.place + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == x0;
    local = ok && same;
    if (pw.y == 0) __hip_atomic_store(J.place + J.G, (unsigned)local, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (local) sor_solve_part<NT, true>(J, pw.y, 1);
  else sor_solve_part<NT, false>(J, pw.y, 1);
}

// one context's solve over gridDim.x workgroups (J.G == gridDim.x)
template <int NT>
__global__ __launch_bounds__(NT) void k_sor_solve_multi(SorJob J) {
  sor_solve_part<NT, false>(J, blockIdx.x, 0);
}

// the barrier counters of solves about to start (flags stay set)
__global__ void k_sor_arm(int n, unsigned *__restrict__ arrive) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) arrive[2 * p] = 0u;
}

// ------------------------------- neighbour-synchronised pressure solve
// The solve of sor_solve_part (same sweeps, same residual order, same
// stopping rule, so the same bits and sweep counts) on G resident
// workgroups that do not meet after every half-sweep. Workgroup w owns the
// list entries of one contiguous range of cell indices (the red list cut
// into G equal parts; the black entries of the same cells). A cell's
// neighbours lie within nx*ny cells of it, so before half-sweep h a
// workgroup waits only for the workgroups whose ranges come within that
// distance of its own to have published half-sweep h - 1 (the values it is
// about to read are then final, and nobody still reads the values it is about
// to overwrite). Values of q are read and written with agent-scope relaxed
// atomics (coherent between compute units and XCDs without L2 write-back or
// invalidation); a workgroup waits for its stores to complete before it
// publishes its progress. Residual checks remain full barriers. Waits poll a
// bounded number of times; a timed-out wait fails the solve.
constexpr int kFlowMaxG = 256;

__device__ __forceinline__ float q_load(const float *q, int c) {
  return __hip_atomic_load(q + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NT>
__device__ __forceinline__ void flow_half(const SorJob &J, long base, long lo, long hi, int nxy) {
  const int nx = J.g.nx;
  // the neighbour loads carry no branches (an absent neighbour loads the
  // cell itself and is not added), so all of a thread's loads are in flight
  // at once; the sum keeps k_sor_list's order
  const int step[6] = {-1, 1, -nx, nx, -nxy, nxy};
  for (long t0 = lo + threadIdx.x; t0 < hi; t0 += (long)NT * kSorU) {
    unsigned e[kSorU];
    float rh[kSorU], pw[kSorU], nb[kSorU][6], qc[kSorU];
#pragma unroll
    for (int u = 0; u < kSorU; u++) {
      long t = t0 + (long)u * NT;
      bool in = t < hi;
      e[u] = in ? J.code[base + t] : 0u;
      rh[u] = in ? J.lrhs[base + t] : 0.f;
      pw[u] = in ? J.lpw[base + t] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kSorU; u++) {
      // entries past hi hold code 0: cell 0, no neighbours
      const int c = (int)(e[u] & 0x3ffffffu);
      const unsigned m = e[u] >> 26;
#pragma unroll
      for (int d = 0; d < 6; d++) nb[u][d] = q_load(J.q, (m >> d) & 1 ? c + step[d] : c);
      qc[u] = q_load(J.q, c);
    }
#pragma unroll
    for (int u = 0; u < kSorU; u++) {
      const unsigned m = e[u] >> 26;
      float a = 0.f;
#pragma unroll
      for (int d = 0; d < 6; d++) a = (m >> d) & 1 ? a + nb[u][d] : a;
      if (t0 + (long)u * NT < hi)
        __hip_atomic_store(J.q + (e[u] & 0x3ffffffu), sor_relax(qc[u], a, rh[u], pw[u], J.om1), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// publish half-sweep h of workgroup w, then wait until workgroups dlo..dhi
// have published it too
__device__ __forceinline__ bool flow_sync(const SorJob &J, int w, int dlo, int dhi, unsigned h) {
  __shared__ int flow_ok;
  // this thread's q stores complete (visible at agent scope) before the
  // workgroup publishes
  __builtin_amdgcn_s_waitcnt(0);
  if (threadIdx.x == 0) flow_ok = 1;
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(J.flow + (long)w * kFlowStride, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < 64) {
    bool good = true;
    for (int v = dlo + (int)threadIdx.x; v <= dhi && good; v += 64) {
      int polls = 0;
      while (__hip_atomic_load(J.flow + (long)v * kFlowStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < h) {
        if (++polls > kSorPolls || __hip_atomic_load(J.arrive + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          good = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (!good) {
      __hip_atomic_store(J.arrive + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      flow_ok = 0;
    }
  }
  __syncthreads();
  return flow_ok != 0;
}

template <int NT>
__device__ void sor_flow_part(SorJob J, int w, int G) {
  J.G = G;
  const long nR = J.ds->nRed, nB = J.ds->nBlack;
  const int nxy = J.g.nx * J.g.ny;
  __shared__ int sclo[kFlowMaxG + 1];
  __shared__ int sdep[2];
  for (int v = threadIdx.x; v <= G; v += NT) {
    long r = nR * v / G;
    sclo[v] = v == 0 ? 0 : (r < nR ? (int)(J.code[r] & 0x3ffffffu) : J.g.nC);
  }
  __syncthreads();
  const int clo = sclo[w], chi = sclo[w + 1];
  if (threadIdx.x == 0) {
    int dlo = w, dhi = w;
    while (dlo > 0 && sclo[dlo] > clo - nxy) dlo--;      // ranges ending after clo - nxy
    while (dhi < G - 1 && sclo[dhi + 1] < chi + nxy) dhi++;  // ranges starting before chi + nxy
    sdep[0] = dlo;
    sdep[1] = dhi;
  }
  __syncthreads();
  const int dlo = sdep[0], dhi = sdep[1];
  const long rlo = nR * w / G, rhi = nR * (w + 1) / G;
  const long blo = (long)(J.off[clo] >> 32), bhi = (long)(J.off[chi] >> 32);
  unsigned done = 0, checks = 0, bar = 0, h = 0;
  double res = 0;
  bool ok = true;
  auto sweep = [&]() {
    flow_half<NT>(J, 0, rlo, rhi, nxy);
    if (!(ok = flow_sync(J, w, dlo, dhi, ++h))) return;
    flow_half<NT>(J, J.half, blo, bhi, nxy);
    ok = flow_sync(J, w, dlo, dhi, ++h);
  };
  auto check = [&]() {
    if ((ok = sor_barrier<false>(J, ++bar))) ok = sor_residual<NT, false>(J, w, nR, nB, nxy, bar, &res);
    checks++;
  };
  if (J.tol <= 0) {
    for (int s = 0; s < J.fixedIters && ok; s++) sweep();
    done = (unsigned)max(J.fixedIters, 0);
    if (ok && J.measure) check();
  } else {
    const unsigned last = J.ds->sorIters;
    check();
    double best = res, prev = res;
    int flat = 0, k = J.every;
    if (J.adaptive && last > 0) k = max(k, (int)(0.75 * last));
    while (ok && res > J.tol && (int)done < J.maxIters && flat < 12) {
      if ((int)done + k > J.maxIters) k = J.maxIters - (int)done;
      for (int s = 0; s < k && ok; s++) sweep();
      done += (unsigned)k;
      if (!ok) break;
      check();
      if (res < 0.995 * best) { best = res; flat = 0; } else flat++;
      int kn = J.every;
      if (J.adaptive && res < prev && res > J.tol) {
        double rate = pow(res / prev, 1.0 / k);
        double need = log(J.tol / res) / log(rate);
        kn = need < 4 ? 4 : (need > 64 ? 64 : (int)ceil(need));
      }
      prev = res;
      k = kn;
    }
  }
  // (ds->sorIters was read by every workgroup before the first check's
  // barrier, so workgroup 0 may rewrite it now)
  if (w == 0 && threadIdx.x == 0) {
    J.ds->sorIters = done;
    J.ds->sorChecksLast = checks;
    J.ds->sorRes = ok ? res : -1.0;
    J.ds->sorSweeps += done;
    J.ds->sorChecks += checks;
  }
}

// G workgroups of one solve. With X < 8 only the blocks of the first X XCDs
// (blocks are dealt to the 8 XCDs in turn) take part: gridDim.x = 8 G / X.
template <int NT>
__global__ __launch_bounds__(NT) void k_sor_flow(SorJob J, int X) {
  const int b = blockIdx.x;
  if (b % 8 >= X) return;
  const int w = (b / 8) * X + b % 8;
  sor_flow_part<NT>(J, w, (int)gridDim.x / 8 * X + min((int)gridDim.x % 8, X));
}

// the progress counters and barrier of a flow solve about to start
__global__ void k_flow_arm(int G, unsigned *__restrict__ flow, unsigned *__restrict__ arrive) {
  for (int v = threadIdx.x; v < G; v += blockDim.x) flow[(long)v * kFlowStride] = 0u;
  if (threadIdx.x == 0) arrive[0] = 0u;
}

template __global__ void k_sor_flow<256>(SorJob, int);
template __global__ void k_sor_flow<512>(SorJob, int);
template __global__ void k_sor_flow<1024>(SorJob, int);
template __global__ void k_sor_solve_multi<256>(SorJob);
template __global__ void k_sor_solve_multi<512>(SorJob);
template __global__ void k_sor_solve_multi<1024>(SorJob);
template __global__ void k_sor_solve<256>(SorJob);
template __global__ void k_sor_solve<512>(SorJob);
template __global__ void k_sor_solve<1024>(SorJob);
template __global__ void k_sor_solve_batch<256>(const SorJob *, const int2 *);
template __global__ void k_sor_solve_batch<512>(const SorJob *, const int2 *);
template __global__ void k_sor_solve_batch<1024>(const SorJob *, const int2 *);

// ---------------------------------------------------------- projection
// _projectSlice, solver.js:1417-1443, with the grid clamp and umax of
// _clampUmSlice (solver.js:2629-2653) fused when confinement is off.
// Block max of |u| into one atomic per block (the max of non-negative
// floats is the max of their bit patterns, so the result is exact and order
// independent; NaN never wins). Kernels using it run a bounded grid-stride
// loop so the number of atomics on the single counter stays small.
__device__ __forceinline__ void umax_reduce(unsigned bits, unsigned *umax_bits) {
  __shared__ unsigned wmax[16];
  for (int o = 32; o > 0; o >>= 1) bits = max(bits, (unsigned)__shfl_xor((int)bits, o));
  int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wmax[wv] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned m = 0;
    for (int i = 0; i < nw; i++) m = max(m, wmax[i]);
    if (m) atomicMax(umax_bits, m);
  }
}

__device__ __forceinline__ unsigned abs_bits(float u) {
  float av = fabsf(u);
  return (av == av) ? __float_as_uint(av) : 0u;
}

__device__ __forceinline__ void k_project_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ ct,
                                               const float *__restrict__ phi, float *__restrict__ vel, int do_clamp,
                                               float vcap, DevScalars *s) {
  long F = (long)g.uN + g.vN + g.wN;
  unsigned bits = 0;
  for (long f = (long)bx * blockDim.x + threadIdx.x; f < F; f += (long)nbx * blockDim.x) {
    FaceInfo fi;
    face_of(g, f, fi);
    float u = vel[f];
    if (!fi.plane) {
      uint8_t tl = ct[fi.cl], tr = ct[fi.cr];
      if (tl != SOLID && tr != SOLID) {
        float pL = tl == FLUID ? phi[fi.cl] : 0.f, pR = tr == FLUID ? phi[fi.cr] : 0.f;
        u = u - (pR - pL) * g.inv_dx;
      }
    }
    if (do_clamp) {
      if (u > vcap) u = vcap; else if (u < -vcap) u = -vcap;
      bits = max(bits, abs_bits(u));
    }
    vel[f] = u;
  }
  if (do_clamp) umax_reduce(bits, &s->umax_bits);
}
__global__ __launch_bounds__(256) void k_project(Geo g, const uint8_t *__restrict__ ct, const float *__restrict__ phi,
                                                 float *__restrict__ vel, int do_clamp, float vcap, DevScalars *s) {
  k_project_body(blockIdx.x, gridDim.x, g, ct, phi, vel, do_clamp, vcap, s);
}
WS_PLANETS_OPS(k_project_ops, k_project, k_project_body)

__device__ __forceinline__ void k_clamp_body(long bx, long nbx, long F, float *__restrict__ vel, float vcap,
                                             DevScalars *s) {
  unsigned bits = 0;
  for (long f = (long)bx * blockDim.x + threadIdx.x; f < F; f += (long)nbx * blockDim.x) {
    float u = vel[f];
    if (u > vcap) u = vcap; else if (u < -vcap) u = -vcap;
    vel[f] = u;
    bits = max(bits, abs_bits(u));
  }
  umax_reduce(bits, &s->umax_bits);
}
__global__ __launch_bounds__(256) void k_clamp(long F, float *__restrict__ vel, float vcap, DevScalars *s) {
  k_clamp_body(blockIdx.x, gridDim.x, F, vel, vcap, s);
}
WS_PLANETS_OPS(k_clamp_ops, k_clamp, k_clamp_body)

// ------------------------------------------------ vorticity confinement
// _vorticityConfinement, solver.js:1454-1559, in three cell passes and one
// face pass. scratch layout: vc[3C] centred velocity, om[4C] (x,y,z,|w|),
// fc[3C] velocity deltas of interior fluid cells.
__device__ __forceinline__ void k_vort_center_body(long bx, long nbx, Geo g, const float *__restrict__ vel,
                                                   float *__restrict__ vc) {
  int c = bx * blockDim.x + threadIdx.x;
  if (c >= g.nC) return;
  int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny);
  const float *u = vel, *v = vel + g.uN, *w = vel + g.uN + g.vN;
  int ru = (k * g.ny + j) * (g.nx + 1) + i, rv = (k * (g.ny + 1) + j) * g.nx + i;
  vc[c] = 0.5f * (u[ru] + u[ru + 1]);
  vc[g.nC + c] = 0.5f * (v[rv] + v[rv + g.nx]);
  vc[2 * g.nC + c] = 0.5f * (w[c] + w[c + g.nx * g.ny]);
}
__global__ void k_vort_center(Geo g, const float *__restrict__ vel, float *__restrict__ vc) {
  k_vort_center_body(blockIdx.x, gridDim.x, g, vel, vc);
}
WS_PLANETS_OPS(k_vort_center_ops, k_vort_center, k_vort_center_body)

__device__ __forceinline__ void k_vort_omega_body(long bx, long nbx, Geo g, const float *__restrict__ vc,
                                                  float *__restrict__ om, float inv2) {
  int c = bx * blockDim.x + threadIdx.x;
  if (c >= g.nC) return;
  int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny), nxy = g.nx * g.ny;
  int im = i > 0 ? c - 1 : c, ip = i < g.nx - 1 ? c + 1 : c;
  int jm = j > 0 ? c - g.nx : c, jp = j < g.ny - 1 ? c + g.nx : c;
  int km = k > 0 ? c - nxy : c, kp = k < g.nz - 1 ? c + nxy : c;
  const float *cx = vc, *cy = vc + g.nC, *cz = vc + 2 * g.nC;
  float wx = (cz[jp] - cz[jm] - (cy[kp] - cy[km])) * inv2;
  float wy = (cx[kp] - cx[km] - (cz[ip] - cz[im])) * inv2;
  float wz = (cy[ip] - cy[im] - (cx[jp] - cx[jm])) * inv2;
  om[c] = wx;
  om[g.nC + c] = wy;
  om[2 * g.nC + c] = wz;
  om[3 * g.nC + c] = sqrtf(wx * wx + wy * wy + wz * wz);
}
__global__ void k_vort_omega(Geo g, const float *__restrict__ vc, float *__restrict__ om, float inv2) {
  k_vort_omega_body(blockIdx.x, gridDim.x, g, vc, om, inv2);
}
WS_PLANETS_OPS(k_vort_omega_ops, k_vort_omega, k_vort_omega_body)

__device__ __forceinline__ void k_vort_force_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ ct,
                                                  const float *__restrict__ om, float *__restrict__ fc, float eps,
                                                  const DevScalars *__restrict__ s) {
  int c = bx * blockDim.x + threadIdx.x;
  if (c >= g.nC) return;
  if (ct[c] != FLUID) return;
  int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny), nxy = g.nx * g.ny;
  const float *ox = om, *oy = om + g.nC, *oz = om + 2 * g.nC, *mg = om + 3 * g.nC;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  bool interior = !((i > 0 && ct[c - 1] != FLUID) || (i < g.nx - 1 && ct[c + 1] != FLUID) ||
                    (j > 0 && ct[c - g.nx] != FLUID) || (j < g.ny - 1 && ct[c + g.nx] != FLUID) ||
                    (k > 0 && ct[c - nxy] != FLUID) || (k < g.nz - 1 && ct[c + nxy] != FLUID));
  if (mg[c] >= 0.35f && interior) {
    float gx = mg[i < g.nx - 1 ? c + 1 : c] - mg[i > 0 ? c - 1 : c];
    float gy = mg[j < g.ny - 1 ? c + g.nx : c] - mg[j > 0 ? c - g.nx : c];
    float gz = mg[k < g.nz - 1 ? c + nxy : c] - mg[k > 0 ? c - nxy : c];
    float g2 = gx * gx + gy * gy + gz * gz;
    if (!(g2 < 1e-12f)) {
      float inv = 1.f / sqrtf(g2);
      float Nx = gx * inv, Ny = gy * inv, Nz = gz * inv;
      fx = eps * (Ny * oz[c] - Nz * oy[c]);
      fy = eps * (Nz * ox[c] - Nx * oz[c]);
      fz = eps * (Nx * oy[c] - Ny * ox[c]);
      float f2 = fx * fx + fy * fy + fz * fz;
      if (f2 > 36.f) { float fs = 6.f / sqrtf(f2); fx *= fs; fy *= fs; fz *= fs; }
      float dt = s->dt;
      fx *= dt; fy *= dt; fz *= dt;
    }
  }
  fc[c] = fx;
  fc[g.nC + c] = fy;
  fc[2 * g.nC + c] = fz;
}
__global__ void k_vort_force(Geo g, const uint8_t *__restrict__ ct, const float *__restrict__ om,
                             float *__restrict__ fc, float eps, const DevScalars *__restrict__ s) {
  k_vort_force_body(blockIdx.x, gridDim.x, g, ct, om, fc, eps, s);
}
WS_PLANETS_OPS(k_vort_force_ops, k_vort_force, k_vort_force_body)

__device__ __forceinline__ void k_vort_splat_body(long bx, long nbx, Geo g, const uint8_t *__restrict__ ct,
                                                  const float *__restrict__ fc, float *__restrict__ vel) {
  long F = (long)g.uN + g.vN + g.wN;
  long f = (long)bx * blockDim.x + threadIdx.x;
  if (f >= F) return;
  FaceInfo fi;
  face_of(g, f, fi);
  if (fi.cl < 0 || fi.cr < 0 || ct[fi.cl] != FLUID || ct[fi.cr] != FLUID) return;
  const float *fcc = fc + (long)fi.comp * g.nC;
  vel[f] = vel[f] + 0.5f * (fcc[fi.cl] + fcc[fi.cr]);
}
__global__ void k_vort_splat(Geo g, const uint8_t *__restrict__ ct, const float *__restrict__ fc,
                             float *__restrict__ vel) {
  k_vort_splat_body(blockIdx.x, gridDim.x, g, ct, fc, vel);
}
WS_PLANETS_OPS(k_vort_splat_ops, k_vort_splat, k_vort_splat_body)

// _extrapolate's flood (solver.js:1613-1634): in lexicographic order, an
// invalid face takes the mean of its valid neighbours, where a face filled
// earlier in the same pass counts as valid. A face therefore depends on the
// final state of its -i/-j/-k neighbours and on the original state of its
// +i/+j/+k neighbours. Tiles of 64 x 8 x 8 faces are processed one tile
// diagonal (I+J+K) per launch; the tile and a one-face halo (the -i/-j/-k
// side already final from earlier diagonals, the +i/+j/+k side still
// original) are loaded into LDS.
//
// Validity does not depend on the values: a face ends valid iff some face
// y <= x (componentwise, inside the tile, or a final halo face just below
// it) was valid or had a +i/+j/+k neighbour that was valid originally. So
// the final flags of the whole tile are a 3-D prefix OR, computed up front
// from 64-bit row masks (a ballot and a lowest-set-bit fill along i, an OR
// down j in registers and down k through LDS). With every flag known, the
// value of a filled face is an affine function of its -i, -j and -k
// neighbours' final values, x = r (x(-i) + x(-j) + x(-k) [those valid] +
// the sum of the rest), r the reciprocal of the count; every coefficient
// is computed before the recurrence starts.
//
// The recurrence runs over rows along i by diagonals lj + lk, one wavefront
// per k-plane and one lane per face: at step s wavefront lk takes row
// lj = s - lk, whose -j row it finished itself one step earlier (kept in a
// register) and whose -k row the wavefront below finished one step earlier
// (handed over in LDS). Along the row the -i dependence is an inclusive scan
// of the affine maps x -> A x + B across the wavefront (DPP row shifts and
// row broadcasts); the multiplicative half of that scan is value-free and
// precomputed, so a step is two FMAs, six DPP FMAs, an LDS read and write,
// and a barrier. The flags are those of the serial flood exactly; the values
// differ from its sequential double sums by rounding (a few float ulps).
__device__ __forceinline__ float ex_recip(int c) {
  // 1/c, correctly rounded, for the neighbour counts 1..6
  float r = 1.f;
  r = c == 2 ? 0.5f : r;
  r = c == 3 ? 1.f / 3.f : r;
  r = c == 4 ? 0.25f : r;
  r = c == 5 ? 0.2f : r;
  r = c == 6 ? 1.f / 6.f : r;
  return r;
}

// lane-shifted copy of v (DPP): lanes with no source lane get `old`
#define EX_DPP(v, CTRL, RMASK, OLD) \
  __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(OLD), __float_as_int(v), CTRL, RMASK, 0xf, false))
// B += a * (B of the source lane), one DPP FMA; a lane with no source lane,
// or in a row the mask leaves out, is not written (keeps B). The s_nop
// gives the two wait states a DPP read needs after B's last VALU write.
#define EX_FMAC_DPP(B, a, CTRL) \
  asm("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 " CTRL " bank_mask:0xf" : "+v"(B) : "v"(B), "v"(a))

__device__ __forceinline__ void extrap_tile(const Geo &g, int4 t, float *__restrict__ vel,
                                            uint8_t *__restrict__ valid) {
  constexpr int TI = EX_TI, TJ = EX_TJ, TK = EX_TK;
  static_assert(TI == 64 && EX_THREADS == 64 * TK, "one lane per face of a row, one wavefront per k-plane");
  constexpr int HI = TI + 2, HJ = TJ + 2, HK = TK + 2, HN = HI * HJ * HK;
  __shared__ float sv[HK][HJ][HI];
  __shared__ uint8_t sf[HK][HJ][HI];
  __shared__ unsigned long long vmj[TK][TJ];  // row masks, prefix OR along i and j
  __shared__ unsigned long long vmk[TK][TJ];  // ... and along k: the final flags
  __shared__ float xfin[TK][2][64];           // a finished row, for the wavefront above
  const int comp = t.x, I = t.y, J = t.z, K = t.w;
  const int d0 = comp == 0 ? g.nx + 1 : g.nx, d1 = comp == 1 ? g.ny + 1 : g.ny, d2 = comp == 2 ? g.nz + 1 : g.nz;
  const long base = comp == 0 ? 0 : (comp == 1 ? g.uN : (long)g.uN + g.vN);
  float *Av = vel + base;
  uint8_t *Vf = valid + base;
  const int i0 = I * TI - 1, j0 = J * TJ - 1, k0 = K * TK - 1;
  {
    // every load of the tile first (one round trip), then the LDS stores
    constexpr int NB = (HN + EX_THREADS - 1) / EX_THREADS;
    float val[NB];
    int fl[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int e = b * EX_THREADS + threadIdx.x;
      const int li = e % HI, lj = (e / HI) % HJ, lk = e / (HI * HJ);
      const int i = i0 + li, j = j0 + lj, k = k0 + lk;
      const bool in = e < HN && i >= 0 && i < d0 && j >= 0 && j < d1 && k >= 0 && k < d2;
      // unconditional loads (face 0 stands in for a halo face outside the
      // grid) so that the compiler issues them together
      const long f = in ? ((long)k * d1 + j) * d0 + i : 0;
      const float v = Av[f];
      const int q = Vf[f];
      val[b] = in ? v : 0.f;
      fl[b] = in ? q : 0;
    }
    int any = 0, open = 0;
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int e = b * EX_THREADS + threadIdx.x;
      if (e < HN) {
        const int li = e % HI, lj = (e / HI) % HJ, lk = e / (HI * HJ);
        const int i = i0 + li, j = j0 + lj, k = k0 + lk;
        const bool in = i >= 0 && i < d0 && j >= 0 && j < d1 && k >= 0 && k < d2;
        const bool own = li >= 1 && li <= TI && lj >= 1 && lj <= TJ && lk >= 1 && lk <= TK;
        const int f = fl[b] ? 1 : 0;
        (&sv[0][0][0])[e] = val[b];
        (&sf[0][0][0])[e] = (uint8_t)f;
        any |= f;
        open |= (in && own && !f) ? 1 : 0;
      }
    }
    // nothing to fill (every owned face valid), or nothing to fill from (no
    // valid face in the tile or its halo): the flood leaves the tile as it is
    any = __syncthreads_or(any);
    open = __syncthreads_or(open);
    if (!any || !open) return;
  }
  const int lane = threadIdx.x & 63;
  const int lk = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // this wavefront's k-plane
  const int i = i0 + 1 + lane, k = k0 + 1 + lk;
  const int si = lane + 1, sk = lk + 1;
  const int nI = d0 - (i0 + 1);  // faces of the tile's rows inside the grid, >= 1
  const unsigned long long inMask = nI >= 64 ? ~0ull : ((1ull << nI) - 1ull);

  // (1) the final flags. A seed is a face inside the grid that is valid or
  // has a valid +i/+j/+k neighbour, or lies next to a valid -i/-j/-k halo face
  {
    unsigned long long pj = 0;
#pragma unroll
    for (int lj = 0; lj < TJ; lj++) {
      const int sj = lj + 1;
      const bool in = k < d2 && j0 + 1 + lj < d1 && lane < nI;
      int s = sf[sk][sj][si] | sf[sk][sj][si + 1] | sf[sk][sj + 1][si] | sf[sk + 1][sj][si];
      s |= lane == 0 ? sf[sk][sj][0] : 0;
      s |= lj == 0 ? sf[sk][0][si] : 0;
      s |= lk == 0 ? sf[0][sj][si] : 0;
      const unsigned long long m = __ballot(in && s);
      pj |= m ? (~0ull << __builtin_ctzll(m)) : 0ull;
      if (lane == 0) vmj[lk][lj] = pj;
    }
  }
  __syncthreads();
  if (threadIdx.x < TK * TJ) {
    const int kk = threadIdx.x / TJ, jj = threadIdx.x % TJ;
    unsigned long long a = 0;
    for (int q = 0; q <= kk; q++) a |= vmj[q][jj];
    const bool rin = k0 + 1 + kk < d2 && j0 + 1 + jj < d1;
    vmk[kk][jj] = rin ? (a & inMask) : 0ull;
  }
  __syncthreads();

  // (2) per face: x = A x(-i) + Pj x(-j) + Pk x(-k) + Q, and the scan
  // multipliers of A. Faces that stay as they are (valid, or left invalid)
  // have A = Pj = Pk = 0 and Q their value (0 for a face left invalid, so
  // no value of an invalid face reaches a product)
  float Q[TJ], Pj[TJ], Pk[TJ], As[TJ][6];
  unsigned fillBits = 0;
  {
    unsigned long long prevRow = 0;
#pragma unroll
    for (int lj = 0; lj < TJ; lj++) {
      const int sj = lj + 1;
      const unsigned long long Vm = vmk[lk][lj];
      const unsigned long long Vkm = lk > 0 ? vmk[lk > 0 ? lk - 1 : 0][lj] : 0ull;
      const bool V = (Vm >> lane) & 1ull;
      const int f0 = sf[sk][sj][si];
      const int fb = sf[sk][sj][si + 1], fd = sf[sk][sj + 1][si], ff = sf[sk + 1][sj][si];
      const float v0 = sv[sk][sj][si], vb = sv[sk][sj][si + 1], vd = sv[sk][sj + 1][si], vf = sv[sk + 1][sj][si];
      const int fh = sf[sk][sj][0], fjh = sf[sk][0][si], fkh = sf[0][sj][si];
      const float vh = sv[sk][sj][0], vjh = sv[sk][0][si], vkh = sv[0][sj][si];
      // the -i / -j / -k neighbours' final flags
      const int ua = lane > 0 ? (int)((Vm >> (lane > 0 ? lane - 1 : 0)) & 1ull) : fh;
      const int uc = lj > 0 ? (int)((prevRow >> lane) & 1ull) : fjh;
      const int ue = lk > 0 ? (int)((Vkm >> lane) & 1ull) : fkh;
      const int cnt = ua + fb + uc + fd + ue + ff;
      const bool fill = V && !f0;
      const float r = ex_recip(cnt);
      // the neighbours whose values are known now, in the reference's order
      // (-i, +i, -j, +j, -k, +k); the halo faces among them are final
      float R = 0.f;
      R = (lane == 0 && ua) ? R + vh : R;
      R = fb ? R + vb : R;
      R = (lj == 0 && uc) ? R + vjh : R;
      R = fd ? R + vd : R;
      R = (lk == 0 && ue) ? R + vkh : R;
      R = ff ? R + vf : R;
      Q[lj] = fill ? R * r : (V ? v0 : 0.f);
      Pj[lj] = (fill && lj > 0 && uc) ? r : 0.f;
      Pk[lj] = (fill && lk > 0 && ue) ? r : 0.f;
      float A = (fill && lane > 0 && ua) ? r : 0.f;
      As[lj][0] = A;
      A *= EX_DPP(A, 0x111, 0xf, 1.f);  // row_shr:1
      As[lj][1] = A;
      A *= EX_DPP(A, 0x112, 0xf, 1.f);  // row_shr:2
      As[lj][2] = A;
      A *= EX_DPP(A, 0x114, 0xf, 1.f);  // row_shr:4
      As[lj][3] = A;
      A *= EX_DPP(A, 0x118, 0xf, 1.f);  // row_shr:8
      As[lj][4] = A;
      A *= EX_DPP(A, 0x142, 0xa, 1.f);  // row_bcast:15 into rows 1 and 3
      As[lj][5] = A;
      if (fill) {
        fillBits |= 1u << lj;
        Vf[((long)k * d1 + (j0 + 1 + lj)) * d0 + i] = 1;
      }
      prevRow = Vm;
    }
  }

  // (3) the recurrence: TJ + TK - 1 steps, one barrier each; wavefront lk
  // is busy at steps lk .. lk + TJ - 1
  for (int q = 0; q < lk; q++) __syncthreads();
  float xj = 0.f;
#pragma unroll
  for (int lj = 0; lj < TJ; lj++) {
    const float xk = lk > 0 ? xfin[lk > 0 ? lk - 1 : 0][lj & 1][lane] : 0.f;
    float B = fmaf(Pk[lj], xk, fmaf(Pj[lj], xj, Q[lj]));
    EX_FMAC_DPP(B, As[lj][0], "row_shr:1 row_mask:0xf");
    EX_FMAC_DPP(B, As[lj][1], "row_shr:2 row_mask:0xf");
    EX_FMAC_DPP(B, As[lj][2], "row_shr:4 row_mask:0xf");
    EX_FMAC_DPP(B, As[lj][3], "row_shr:8 row_mask:0xf");
    EX_FMAC_DPP(B, As[lj][4], "row_bcast:15 row_mask:0xa");  // into rows 1 and 3
    EX_FMAC_DPP(B, As[lj][5], "row_bcast:31 row_mask:0xc");  // into rows 2 and 3
    if (lk < TK - 1) xfin[lk][lj & 1][lane] = B;
    if (fillBits & (1u << lj)) Av[((long)k * d1 + (j0 + 1 + lj)) * d0 + i] = B;
    xj = B;
    __syncthreads();
  }
  for (int q = lk; q < TK - 1; q++) __syncthreads();
}
#undef EX_DPP
#undef EX_FMAC_DPP

// one launch per tile diagonal of one context
// the b-th tile of a diagonal (scalar arithmetic: no table load ahead of
// the tile's own loads)
__device__ __forceinline__ int4 ex_tile_of(const ExDiag &D, int b) {
  int comp = 0;
  while (comp < 2 && b >= D.cnt[comp]) b -= D.cnt[comp++];
  const int ti = D.ti[comp], tj = D.tj[comp], tk = D.tk[comp];
  for (int K = 0; K < tk; K++) {
    const int jlo = max(0, D.d - K - (ti - 1)), jhi = min(tj - 1, D.d - K);
    const int n = jhi >= jlo ? jhi - jlo + 1 : 0;
    if (b < n) return make_int4(comp, D.d - K - (jlo + b), jlo + b, K);
    b -= n;
  }
  return make_int4(comp, 0, 0, 0);  // not reached: b < the diagonal's count
}

__global__ __launch_bounds__(EX_THREADS) void k_extrap_diag(Geo g, ExDiag D, float *__restrict__ vel,
                                                     uint8_t *__restrict__ valid) {
  extrap_tile(g, ex_tile_of(D, __builtin_amdgcn_readfirstlane(blockIdx.x)), vel, valid);
}

// one launch per tile diagonal of many planets: block b takes tile b of the
// merged schedule, on planet planetOf[b] (planets are independent, and a
// planet's tiles of one diagonal depend only on its earlier diagonals)
__global__ __launch_bounds__(EX_THREADS) void k_extrap_diag_batch(const ExPlanet *__restrict__ planets,
                                                           const int4 *__restrict__ tiles,
                                                           const int *__restrict__ planetOf, int off) {
  const int b = off + blockIdx.x;
  const ExPlanet P = planets[planetOf[b]];
  extrap_tile(P.g, tiles[b], P.vel, P.valid);
}

}  // namespace ws
