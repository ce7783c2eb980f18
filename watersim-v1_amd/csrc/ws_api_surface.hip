// C ABI of the surface path: splatDensity (js/solver.js:3785-3820),
// MarchingTetrahedra.build (js/surface.js:66-177), autotuneIso
// (solver.js:3838-3849) and kineticEnergy (solver.js:3914-3921).
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

#include "ws_frame.h"

using namespace ws;

namespace ws {

// counting-sort bins of the grid-coupled particles by their base cell
// (floor(x/dx), clamped) for the corner gather of the density splat
__global__ void k_bin_particles(Geo g, const float *__restrict__ thr, int nP, const float *__restrict__ px,
                                const float *__restrict__ py, const float *__restrict__ pz,
                                const uint8_t *__restrict__ pflag, unsigned *__restrict__ cnt, int *__restrict__ pcell,
                                unsigned *__restrict__ prank) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  if (pflag[p] != 0) { pcell[p] = -1; return; }
  int i = clampi(cell_floor(px[p], g.inv_dx, thr, g.nx), 0, g.nx - 1);
  int j = clampi(cell_floor(py[p], g.inv_dx, thr, g.ny), 0, g.ny - 1);
  int k = clampi(cell_floor(pz[p], g.inv_dx, thr, g.nz), 0, g.nz - 1);
  int c = (k * g.ny + j) * g.nx + i;
  pcell[p] = c;
  prank[p] = atomicAdd(&cnt[c], 1u);
}

__global__ void k_zero_counts(unsigned *p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_bin_scatter(int nP, const int *__restrict__ pcell, const unsigned *__restrict__ prank,
                              const unsigned *__restrict__ start, unsigned *__restrict__ sorted) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  int c = pcell[p];
  if (c >= 0) sorted[start[c] + prank[p]] = (unsigned)p;
}

}  // namespace ws

namespace {

template <class T>
int grow(ws_ctx *c, T **p, long &have, long need) {
  if (have >= need) return WS_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (hipMalloc((void **)p, sizeof(T) * (need > 0 ? need : 1)) != hipSuccess) {
    have = 0;
    c->err = "hipMalloc failed in surface scratch";
    return WS_EHIP;
  }
  have = need;
  return WS_OK;
}

int ensure_scan(ws_ctx *c, long n) {
  size_t b = 0;
  WS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (unsigned *)nullptr, (unsigned *)nullptr, (int)n, c->stream));
  if (b <= c->scanTmpBytes) return WS_OK;
  if (c->scanTmp) (void)hipFree(c->scanTmp);
  c->scanTmp = nullptr;
  if (hipMalloc(&c->scanTmp, b) != hipSuccess) { c->err = "hipMalloc scan scratch"; return WS_EHIP; }
  c->scanTmpBytes = b;
  return WS_OK;
}

}  // namespace

namespace ws {

// marching tetrahedra over a device corner field of (nx+1)(ny+1)(nz+1)
// values: count, scan, emit with the FP64 volume
int mt_build_dev(ws_ctx *c, const float *dd, int nx, int ny, int nz, double dx, double iso, float *pos, float *nrm,
                 long cap, long *count, double *volume) {
  long nC = (long)nx * ny * nz;
  int r;
  long have = c->mtCells;
  if ((r = grow(c, &c->mtOff, have, nC + 1))) return r;
  have = c->mtCells;
  if ((r = grow(c, &c->mtCount, have, nC + 1))) return r;
  c->mtCells = have;
  if ((r = ensure_scan(c, nC + 1))) return r;
  hipLaunchKernelGGL(k_zero_counts, dim3(blocks_for(nC + 1, 256)), dim3(256), 0, c->stream, c->mtCount, nC + 1);
  hipLaunchKernelGGL(k_mt_count, dim3(blocks_for(nC, 256)), dim3(256), 0, c->stream, nx, ny, nz, dd, iso,
                     c->mtCount);
  WS_LAUNCHED();
  WS_HIP(hipcub::DeviceScan::ExclusiveSum(c->scanTmp, c->scanTmpBytes, c->mtCount, c->mtOff, (int)(nC + 1),
                                          c->stream));
  unsigned total = 0;
  WS_HIP(hipMemcpyAsync(&total, c->mtOff + nC, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  long need = pos ? (long)total : 0;
  if (need > c->mtCap) {  // vertex capacity of the device mesh buffers
    if (c->mtPos) (void)hipFree(c->mtPos);
    if (c->mtNrm) (void)hipFree(c->mtNrm);
    c->mtPos = c->mtNrm = nullptr;
    c->mtCap = 0;
    WS_HIP(hipMalloc((void **)&c->mtPos, sizeof(float) * 3 * need));
    WS_HIP(hipMalloc((void **)&c->mtNrm, sizeof(float) * 3 * need));
    c->mtCap = need;
  }
  int nb = blocks_for(nC, 256);
  double *vpart = nullptr;
  WS_HIP(hipMallocAsync((void **)&vpart, sizeof(double) * (nb + 1), c->stream));
  hipLaunchKernelGGL(k_mt_emit, dim3(nb), dim3(256), 0, c->stream, nx, ny, nz, dx, dd, iso, c->mtOff,
                     need ? c->mtPos : nullptr, need ? c->mtNrm : nullptr, need, vpart);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, c->stream, vpart, nb, vpart + nb);
  WS_LAUNCHED();
  double vol = 0;
  WS_HIP(hipMemcpyAsync(&vol, vpart + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  long ncopy = need < cap ? need : cap;
  if (pos && ncopy > 0)
    WS_HIP(hipMemcpyAsync(pos, c->mtPos, sizeof(float) * 3 * ncopy, hipMemcpyDeviceToHost, c->stream));
  if (pos && nrm && ncopy > 0)
    WS_HIP(hipMemcpyAsync(nrm, c->mtNrm, sizeof(float) * 3 * ncopy, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipFreeAsync(vpart, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  if (count) *count = (long)total;
  if (volume) *volume = std::fabs(vol);
  return WS_OK;
}


}  // namespace ws

extern "C" {

int ws_splat_density(ws_ctx *c) {
  if (!c) return WS_EINVAL;
  int nC = (int)c->nC;
  hipLaunchKernelGGL(k_zero_counts, dim3(blocks_for(nC + 1, 256)), dim3(256), 0, c->stream, c->cellCount,
                     (long)nC + 1);
  if (c->nP > 0)
    hipLaunchKernelGGL(k_bin_particles, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->geo, c->thr, c->nP,
                       c->px, c->py, c->pz, c->pflag, c->cellCount, c->pcell, c->prank);
  WS_LAUNCHED();
  if (int r = scan_excl(c, c->cellCount, c->cellStart, (long)nC + 1)) return r;
  if (c->nP > 0)
    hipLaunchKernelGGL(k_bin_scatter, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->nP, c->pcell,
                       c->prank, c->cellStart, c->sortedIdx);
  double r = 1.6 * c->spacing;
  World w{};
  hipLaunchKernelGGL(k_splat_density, dim3(blocks_for(c->nD, 256)), dim3(256), 0, c->stream, c->geo, w, c->nP,
                     c->cellStart, c->sortedIdx, c->px, c->py, c->pz, c->dens, (float)(r * r));
  WS_LAUNCHED();
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_mt_build(ws_ctx *c, const float *dens, int nx, int ny, int nz, double dx, double iso, float *pos, float *nrm,
                long cap, long *count, double *volume) {
  if (!c || nx < 1 || ny < 1 || nz < 1 || !(dx > 0)) return WS_EINVAL;
  long nD = (long)(nx + 1) * (ny + 1) * (nz + 1);
  const float *dd = c->dens;
  int r;
  if (dens) {
    if ((r = grow(c, &c->mtDens, c->mtDensN, nD))) return r;
    WS_HIP(hipMemcpyAsync(c->mtDens, dens, sizeof(float) * nD, hipMemcpyHostToDevice, c->stream));
    dd = c->mtDens;
  } else if (nx != c->nx || ny != c->ny || nz != c->nz) {
    c->err = "ws_mt_build on the context field needs the context dimensions";
    return WS_EINVAL;
  }
  return mt_build_dev(c, dd, nx, ny, nz, dx, iso, pos, nrm, cap, count, volume);
}

int ws_autotune_iso(ws_ctx *c, double *iso) {
  if (!c || !iso) return WS_EINVAL;
  // autotuneIso, solver.js:3838-3849: 8 bisections of [0.3, 3.0] on the
  // meshed volume against nP * particleVolume
  double Vt = c->nP * c->particleVolume, lo = 0.3, hi = 3.0;
  long nC = c->nC;
  int nb = blocks_for(nC, 256);
  double *vpart = nullptr;
  unsigned *cnt = nullptr;
  WS_HIP(hipMallocAsync((void **)&vpart, sizeof(double) * (nb + 1), c->stream));
  WS_HIP(hipMallocAsync((void **)&cnt, sizeof(unsigned), c->stream));
  for (int it = 0; it < 8; it++) {
    double mid = 0.5 * (lo + hi);
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, c->stream, cnt, 1L);
    hipLaunchKernelGGL(k_mt_volume, dim3(nb), dim3(256), 0, c->stream, c->nx, c->ny, c->nz, c->dx, c->dens,
                       mid, vpart, cnt);
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, c->stream, vpart, nb, vpart + nb);
    WS_LAUNCHED();
    double vol = 0;
    unsigned n = 0;
    WS_HIP(hipMemcpyAsync(&vol, vpart + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    WS_HIP(hipMemcpyAsync(&n, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    WS_HIP(hipStreamSynchronize(c->stream));
    if (n > 900000u || std::fabs(vol) > Vt) lo = mid; else hi = mid;
  }
  WS_HIP(hipFreeAsync(vpart, c->stream));
  WS_HIP(hipFreeAsync(cnt, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  *iso = 0.5 * (lo + hi);
  c->iso = *iso;
  return WS_OK;
}

int ws_kinetic_energy(ws_ctx *c, double *ke) {
  if (!c || !ke) return WS_EINVAL;
  const int nb = 256;
  double *part = c->partials;  // 2*512+2048 doubles
  if (c->nP > 0) {
    hipLaunchKernelGGL(k_kinetic, dim3(nb), dim3(256), 0, c->stream, c->nP, c->pvx, c->pvy, c->pvz, part);
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, c->stream, part, nb, part + nb);
    WS_LAUNCHED();
  }
  double s = 0;
  if (c->nP > 0) {
    WS_HIP(hipMemcpyAsync(&s, part + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    WS_HIP(hipStreamSynchronize(c->stream));
  }
  *ke = 0.5 * (1000 * c->particleVolume) * s;
  return WS_OK;
}

}  // extern "C"
