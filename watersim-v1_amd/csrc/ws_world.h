// Terrain queries of the planet world on the device (js/solver.js:512-587),
// shared by the marking, G2P and advection kernels.
#pragma once
#include "ws_df.h"
#include "ws_internal.h"

namespace ws {

struct World {
  const uint8_t *st;   // static raster (SOLID/AIR), nx*ny*nz
  const float *R;      // terrain radius lattice, tn^3
  const float *thr;    // exact floor(x/dx) thresholds
  int tn;
  float dv, inv_dv;
  df dv_df;
  df cxd, cyd, czd;    // domain centre in float-float
  float oceanR, skyAlt, atmosphereH, domainR, coreR;
  df lim_df, skyAlt_df, p3_df;  // (oceanR+0.3)^2, max(2dx,0.3), 0.3
};

// terrainRadiusAt, solver.js:512-535 (float)
__device__ __forceinline__ float terrain_radius(const World &w, float x, float y, float z) {
  if (!w.R) return w.coreR;
  int n = w.tn;
  float gx = x * w.inv_dv - 0.5f, gy = y * w.inv_dv - 0.5f, gz = z * w.inv_dv - 0.5f;
  int i = clampi((int)gx, 0, n - 2), j = clampi((int)gy, 0, n - 2), k = clampi((int)gz, 0, n - 2);
  float fx = fminf(fmaxf(gx - (float)i, 0.f), 1.f);
  float fy = fminf(fmaxf(gy - (float)j, 0.f), 1.f);
  float fz = fminf(fmaxf(gz - (float)k, 0.f), 1.f);
  long b = ((long)k * n + j) * n + i, s7 = (long)n * n;
  const float *R = w.R;
  float v000 = R[b], v100 = R[b + 1], v010 = R[b + n], v110 = R[b + n + 1];
  float v001 = R[b + s7], v101 = R[b + s7 + 1], v011 = R[b + s7 + n], v111 = R[b + s7 + n + 1];
  float a0 = v000 + (v100 - v000) * fx, a1 = v010 + (v110 - v010) * fx;
  float a2 = v001 + (v101 - v001) * fx, a3 = v011 + (v111 - v011) * fx;
  float b0 = a0 + (a1 - a0) * fy, b1 = a2 + (a3 - a2) * fy;
  return b0 + (b1 - b0) * fz;
}

// the same lookup in float-float for a float point
__device__ __forceinline__ df terrain_radius_df(const World &w, float x, float y, float z) {
  if (!w.R) return df_make(w.coreR);
  int n = w.tn;
  df h = df_make(0.5f);
  df gx = df_sub(df_div(df_make(x), w.dv_df), h);
  df gy = df_sub(df_div(df_make(y), w.dv_df), h);
  df gz = df_sub(df_div(df_make(z), w.dv_df), h);
  int i = clampi(df_floor(gx), 0, n - 2), j = clampi(df_floor(gy), 0, n - 2), k = clampi(df_floor(gz), 0, n - 2);
  df fx = df_sub(gx, df_make((float)i)), fy = df_sub(gy, df_make((float)j)), fz = df_sub(gz, df_make((float)k));
  if (fx.hi < 0.f) fx = df_make(0.f); else if (fx.hi > 1.f) fx = df_make(1.f);
  if (fy.hi < 0.f) fy = df_make(0.f); else if (fy.hi > 1.f) fy = df_make(1.f);
  if (fz.hi < 0.f) fz = df_make(0.f); else if (fz.hi > 1.f) fz = df_make(1.f);
  long b = ((long)k * n + j) * n + i, s7 = (long)n * n;
  const float *R = w.R;
  df v000 = df_make(R[b]), v100 = df_make(R[b + 1]), v010 = df_make(R[b + n]), v110 = df_make(R[b + n + 1]);
  df v001 = df_make(R[b + s7]), v101 = df_make(R[b + s7 + 1]), v011 = df_make(R[b + s7 + n]);
  df v111 = df_make(R[b + s7 + n + 1]);
  df a0 = df_add(v000, df_mul(df_sub(v100, v000), fx)), a1 = df_add(v010, df_mul(df_sub(v110, v010), fx));
  df a2 = df_add(v001, df_mul(df_sub(v101, v001), fx)), a3 = df_add(v011, df_mul(df_sub(v111, v011), fx));
  df b0 = df_add(a0, df_mul(df_sub(a1, a0), fy)), b1 = df_add(a2, df_mul(df_sub(a3, a2), fy));
  return df_add(b0, df_mul(df_sub(b1, b0), fz));
}

// _rockCellAt, solver.js:542-559, through the static raster
__device__ __forceinline__ bool rock_cell(const World &w, const Geo &g, float x, float y, float z) {
  if (!w.R) return false;
  int i = clampi(cell_floor(x, g.inv_dx, w.thr, g.nx), 1, g.nx - 2);
  int j = clampi(cell_floor(y, g.inv_dx, w.thr, g.ny), 1, g.ny - 2);
  int k = clampi(cell_floor(z, g.inv_dx, w.thr, g.nz), 1, g.nz - 2);
  return w.st[(k * g.ny + j) * g.nx + i] == SOLID;
}

}  // namespace ws
