// Surface kernels: density splat (js/solver.js:3785-3820), marching
// tetrahedra (js/surface.js:17-177) with FP64 volume, the iso bisection of
// autotuneIso (solver.js:3838-3849) and the kinetic energy (:3914-3921).
#include "ws_kernels.h"

namespace ws {

// Corner-gather density splat over particles binned by their base cell: a
// corner receives from the particles of the 8 cells around it; the kernel
// weights are summed in 64-bit fixed point, so the sum is exact and
// independent of the order the bins hold.
__global__ void k_splat_density(Geo g, World w, int nP, const unsigned *__restrict__ start,
                                const unsigned *__restrict__ sorted, const float *__restrict__ px,
                                const float *__restrict__ py, const float *__restrict__ pz, float *__restrict__ dens,
                                float r2) {
  long nD = (long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  long cI = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cI >= nD) return;
  int ci = (int)(cI % (g.nx + 1)), cj = (int)((cI / (g.nx + 1)) % (g.ny + 1)), ck = (int)(cI / ((long)(g.nx + 1) * (g.ny + 1)));
  long long acc = 0;
  for (int dk = -1; dk <= 0; dk++) {
    int k0 = ck + dk;
    if (k0 < 0 || k0 >= g.nz) continue;
    for (int dj = -1; dj <= 0; dj++) {
      int j0 = cj + dj;
      if (j0 < 0 || j0 >= g.ny) continue;
      for (int di = -1; di <= 0; di++) {
        int i0 = ci + di;
        if (i0 < 0 || i0 >= g.nx) continue;
        int c = (k0 * g.ny + j0) * g.nx + i0;
        unsigned a = start[c], e = start[c + 1];
        for (unsigned s = a; s < e; s++) {
          unsigned p = sorted[s];
          float fx = px[p] - (float)i0 * g.dx, fy = py[p] - (float)j0 * g.dx, fz = pz[p] - (float)k0 * g.dx;
          float qx = di == 0 ? fx : fx - g.dx, qy = dj == 0 ? fy : fy - g.dx, qz = dk == 0 ? fz : fz - g.dx;
          float d2 = qx * qx + qy * qy + qz * qz;
          if (d2 >= r2) continue;
          float ww = 1.f - d2 / r2;
          ww = ww * ww;
          acc += (long long)__builtin_rintf(ww * kFixScaleF);
        }
      }
    }
  }
  dens[cI] = (float)acc * kFixInvF;
}

// Marching tetrahedra tables (surface.js:17-46): cube corners, the 6 tets
// around the 0-6 diagonal, tet edges and the 16 sign cases.
__constant__ int MT_CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int MT_TET[6][4] = {{0, 5, 1, 6}, {0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}};
__constant__ int MT_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
__constant__ int MT_NTRI[16] = {0, 1, 1, 2, 1, 2, 2, 1, 1, 2, 2, 1, 2, 1, 1, 0};
__constant__ int MT_TRI[16][2][3] = {
    {{0, 0, 0}, {0, 0, 0}}, {{0, 2, 3}, {0, 0, 0}}, {{0, 1, 4}, {0, 0, 0}}, {{2, 3, 4}, {2, 4, 1}},
    {{2, 1, 5}, {0, 0, 0}}, {{0, 3, 5}, {0, 5, 1}}, {{0, 4, 5}, {0, 5, 2}}, {{3, 5, 4}, {0, 0, 0}},
    {{3, 4, 5}, {0, 0, 0}}, {{0, 2, 5}, {0, 5, 4}}, {{0, 1, 5}, {0, 5, 3}}, {{2, 1, 5}, {0, 0, 0}},
    {{2, 1, 4}, {2, 4, 3}}, {{0, 1, 4}, {0, 0, 0}}, {{0, 2, 3}, {0, 0, 0}}, {{0, 0, 0}, {0, 0, 0}}};

// The corner values of a cell and its triangle count. The reference
// compares the float field values with a double contour level, and does
// all vertex, normal, orientation and volume arithmetic in double; so does
// this code (the file is compiled without FMA contraction), storing only
// the final positions and normals as float.
__device__ __forceinline__ int mt_load(int nx, int ny, const float *__restrict__ dens, int i, int j, int k,
                                       double iso, float cv[8]) {
  long sj = nx + 1, sk = (long)(ny + 1) * sj;
  float mn = INFINITY, mx = -INFINITY;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    float v = dens[(k + MT_CORNER[c][2]) * sk + (j + MT_CORNER[c][1]) * sj + i + MT_CORNER[c][0]];
    cv[c] = v;
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  if ((double)mx < iso || (double)mn > iso) return 0;
  int tris = 0;
#pragma unroll
  for (int t = 0; t < 6; t++) {
    int code = 0;
    for (int c = 0; c < 4; c++)
      if ((double)cv[MT_TET[t][c]] > iso) code |= 1 << c;
    tris += MT_NTRI[code];
  }
  return tris;
}

__global__ void k_mt_count(int nx, int ny, int nz, const float *__restrict__ dens, double iso, unsigned *cnt) {
  long nC = (long)nx * ny * nz;
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
  float cv[8];
  cnt[c] = 3u * (unsigned)mt_load(nx, ny, dens, i, j, k, iso, cv);
}

// Triangles of one cell in the reference's order (tets 0..5, case rows),
// surface.js:89-170. Emits when pos != NULL; returns the signed volume.
__device__ double mt_cell(int nx, int ny, int nz, double dx, const float *__restrict__ dens, int i, int j, int k,
                          double iso, const float cv[8], float *pos, float *nrm, long o, long cap) {
  long sj = nx + 1, sk = (long)(ny + 1) * sj;
  double gxv[8], gyv[8], gzv[8];
  double dx2 = 2 * dx;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    int ci = i + MT_CORNER[c][0], cj = j + MT_CORNER[c][1], ck = k + MT_CORNER[c][2];
    int ip = min(ci + 1, nx), im = max(ci - 1, 0), jp = min(cj + 1, ny), jm = max(cj - 1, 0);
    int kp = min(ck + 1, nz), km = max(ck - 1, 0);
    gxv[c] = ((double)dens[ck * sk + cj * sj + ip] - (double)dens[ck * sk + cj * sj + im]) / dx2;
    gyv[c] = ((double)dens[ck * sk + jp * sj + ci] - (double)dens[ck * sk + jm * sj + ci]) / dx2;
    gzv[c] = ((double)dens[kp * sk + cj * sj + ci] - (double)dens[km * sk + cj * sj + ci]) / dx2;
  }
  double vol = 0;
  for (int t = 0; t < 6; t++) {
    int code = 0;
    for (int c = 0; c < 4; c++)
      if ((double)cv[MT_TET[t][c]] > iso) code |= 1 << c;
    for (int e = 0; e < MT_NTRI[code]; e++) {
      double P[3][3], N[3][3];
      for (int vi = 0; vi < 3; vi++) {
        int ed = MT_TRI[code][e][vi];
        int a = MT_TET[t][MT_EDGE[ed][0]], b = MT_TET[t][MT_EDGE[ed][1]];
        double va = cv[a], vb = cv[b], den = vb - va;
        double tl = den != 0 ? (iso - va) / den : 0.5;
        if (tl < 0) tl = 0; else if (tl > 1) tl = 1;
        double wxx = (double)(i + MT_CORNER[a][0]) + (double)(MT_CORNER[b][0] - MT_CORNER[a][0]) * tl;
        double wyy = (double)(j + MT_CORNER[a][1]) + (double)(MT_CORNER[b][1] - MT_CORNER[a][1]) * tl;
        double wzz = (double)(k + MT_CORNER[a][2]) + (double)(MT_CORNER[b][2] - MT_CORNER[a][2]) * tl;
        P[vi][0] = wxx * dx; P[vi][1] = wyy * dx; P[vi][2] = wzz * dx;
        double nxx = -(gxv[a] + (gxv[b] - gxv[a]) * tl);
        double nyy = -(gyv[a] + (gyv[b] - gyv[a]) * tl);
        double nzz = -(gzv[a] + (gzv[b] - gzv[a]) * tl);
        double nl = sqrt(nxx * nxx + nyy * nyy + nzz * nzz);
        if (!(nl != 0)) nl = 1;
        N[vi][0] = nxx / nl; N[vi][1] = nyy / nl; N[vi][2] = nzz / nl;
      }
      double ux = P[1][0] - P[0][0], uy = P[1][1] - P[0][1], uz = P[1][2] - P[0][2];
      double wx = P[2][0] - P[0][0], wy = P[2][1] - P[0][1], wz = P[2][2] - P[0][2];
      double fnx = uy * wz - uz * wy, fny = uz * wx - ux * wz, fnz = ux * wy - uy * wx;
      double cgx = 0, cgy = 0, cgz = 0;
      for (int vi = 0; vi < 3; vi++) { cgx += N[vi][0]; cgy += N[vi][1]; cgz += N[vi][2]; }
      if (fnx * cgx + fny * cgy + fnz * cgz < 0) {
        for (int d = 0; d < 3; d++) {
          double tp = P[2][d]; P[2][d] = P[1][d]; P[1][d] = tp;
          tp = N[2][d]; N[2][d] = N[1][d]; N[1][d] = tp;
        }
      }
      vol += (P[0][0] * (P[1][1] * P[2][2] - P[2][1] * P[1][2]) - P[0][1] * (P[1][0] * P[2][2] - P[2][0] * P[1][2]) +
              P[0][2] * (P[1][0] * P[2][1] - P[2][0] * P[1][1])) / 6;
      if (pos) {
        for (int vi = 0; vi < 3; vi++)
          if (o + vi < cap)
            for (int d = 0; d < 3; d++) {
              pos[(o + vi) * 3 + d] = (float)P[vi][d];
              nrm[(o + vi) * 3 + d] = (float)N[vi][d];
            }
      }
      o += 3;
    }
  }
  return vol;
}

__device__ __forceinline__ void block_sum_f64(double v, double *out) {
  __shared__ double sh[256];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

__global__ __launch_bounds__(256) void k_mt_emit(int nx, int ny, int nz, double dx, const float *__restrict__ dens,
                                                 double iso, const unsigned *__restrict__ off, float *pos, float *nrm,
                                                 long cap, double *vpart) {
  long nC = (long)nx * ny * nz;
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double vol = 0;
  if (c < nC) {
    int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
    float cv[8];
    if (mt_load(nx, ny, dens, i, j, k, iso, cv)) vol = mt_cell(nx, ny, nz, dx, dens, i, j, k, iso, cv, pos, nrm, off[c], cap);
  }
  block_sum_f64(vol, &vpart[blockIdx.x]);
}

// volume and vertex count only (the bisection of autotuneIso)
__global__ __launch_bounds__(256) void k_mt_volume(int nx, int ny, int nz, double dx, const float *__restrict__ dens,
                                                   double iso, double *vpart, unsigned *cnt) {
  long nC = (long)nx * ny * nz;
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double vol = 0;
  unsigned n = 0;
  if (c < nC) {
    int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
    float cv[8];
    int tris = mt_load(nx, ny, dens, i, j, k, iso, cv);
    if (tris) {
      n = 3u * tris;
      vol = mt_cell(nx, ny, nz, dx, dens, i, j, k, iso, cv, nullptr, nullptr, 0, 0);
    }
  }
  for (int o = 32; o > 0; o >>= 1) n += (unsigned)__shfl_xor((int)n, o);
  if ((threadIdx.x & 63) == 0 && n) atomicAdd(cnt, n);
  block_sum_f64(vol, &vpart[blockIdx.x]);
}

__global__ void k_sum_f64(const double *__restrict__ part, int n, double *out) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  block_sum_f64(s, out);
}

__global__ __launch_bounds__(256) void k_kinetic(int nP, const float *__restrict__ pvx, const float *__restrict__ pvy,
                                                 const float *__restrict__ pvz, double *part) {
  double e = 0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nP; p += gridDim.x * blockDim.x) {
    double v2 = (double)pvx[p] * pvx[p] + (double)pvy[p] * pvy[p] + (double)pvz[p] * pvz[p];
    e += v2;
  }
  block_sum_f64(e, &part[blockIdx.x]);
}

}  // namespace ws
