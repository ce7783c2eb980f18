// Metaballs.build (js/surface.js:178-284): a Blinn-style skin of the liquid
// particles. Every grid-coupled particle (flag 0) adds (1 - d^2/R^2)^2 to
// the corners of the metaball lattice within R of it (and within `rad`
// corners of its nearest corner); the field is contoured at the tension's
// level by the marching-tetrahedra kernels of the density surface.
//
// The membership tests are the reference's double expressions (this file is
// compiled without FMA contraction), so exactly the reference's corners
// receive each particle. The sums are 64-bit fixed point (40 fractional
// bits) added with integer atomics: exact and independent of the order, one
// rounding to float at the end. The reference adds in particle order with a
// float rounding after every addition, so a corner's value differs from the
// reference's by those roundings (a few float ulps).
#include <cmath>

#include "ws_frame.h"

namespace ws {

namespace {

constexpr double kMbFix = 1099511627776.0;  // 2^40
constexpr double kMbRMin = 1.6, kMbRSpan = 2.2, kMbIso0 = 1.5, kMbAlpha = 1.5, kMbDMin = 0.5, kMbDMax = 1.6;

// Math.round: nearest integer, halves toward +infinity
__host__ __device__ inline double js_round(double x) {
  double r = floor(x);
  return x - r >= 0.5 ? r + 1 : r;
}

__global__ void k_mb_zero(long n, unsigned long long *__restrict__ acc) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) acc[i] = 0ull;
}

// one particle per thread: the corners of the reference's stencil
// (surface.js:262-280) that pass its three pruning tests
__global__ __launch_bounds__(256) void k_mb_splat(int nP, const float *__restrict__ px, const float *__restrict__ py,
                                                  const float *__restrict__ pz, const uint8_t *__restrict__ pflag,
                                                  int m, double mdx, double R2, int rad,
                                                  unsigned long long *__restrict__ acc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP || pflag[p] != 0) return;
  double gx = (double)px[p] / mdx, gy = (double)py[p] / mdx, gz = (double)pz[p] / mdx;
  int i0 = (int)js_round(gx), j0 = (int)js_round(gy), k0 = (int)js_round(gz);
  int ia = max(i0 - rad, 0), ib = min(i0 + rad, m);
  int ja = max(j0 - rad, 0), jb = min(j0 + rad, m);
  int ka = max(k0 - rad, 0), kb = min(k0 + rad, m);
  long sj = m + 1, sk = (long)(m + 1) * (m + 1);
  for (int k = ka; k <= kb; k++) {
    double ez = k - gz, dz2 = ez * ez;
    if (dz2 >= R2) continue;
    for (int j = ja; j <= jb; j++) {
      double ey = j - gy, dy2 = dz2 + ey * ey;
      if (dy2 >= R2) continue;
      long base = k * sk + j * sj;
      for (int i = ia; i <= ib; i++) {
        double ex = i - gx;
        double d2 = dy2 + ex * ex;
        if (d2 >= R2) continue;
        double u = 1 - d2 / R2;
        atomicAdd(&acc[base + i], (unsigned long long)llrint(u * u * kMbFix));
      }
    }
  }
}

__global__ void k_mb_finish(long n, const unsigned long long *__restrict__ acc, float *__restrict__ field) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) field[i] = (float)((double)acc[i] / kMbFix);
}

}  // namespace

void metaballs_release(ws_ctx *c) {
  (void)hipFree(c->mbField);
  (void)hipFree(c->mbAcc);
  c->mbField = nullptr;
  c->mbAcc = nullptr;
  c->mbCap = 0;
}

}  // namespace ws

using namespace ws;

extern "C" {

int ws_metaballs_build(ws_ctx *c, double tension, double detail, float *pos, float *nrm, long cap, long *count,
                       double *volume, double *iso_out, int *lattice) {
  if (!c) return WS_EINVAL;
  // lattice, kernel radius and contour level (surface.js:198-252)
  double f = 1;
  if (std::isfinite(detail) && detail > 0) f = detail < kMbDMin ? kMbDMin : (detail > kMbDMax ? kMbDMax : detail);
  int m = (int)js_round(c->nx * f);
  if (m < 6) m = 6;
  double W = c->nx * c->dx;
  double mdx = m == c->nx ? c->dx : W / m;
  double t = std::isfinite(tension) ? tension : 0.5;
  t = t < 0 ? 0 : (t > 1 ? 1 : t);
  double R = (kMbRMin + kMbRSpan * t) * c->spacing / mdx;
  double R2 = R * R;
  double rc = std::ceil(R);
  int rad = (int)(rc < 3 ? rc : 3);
  double iso = kMbIso0 * std::pow((kMbRMin + kMbRSpan * t) / kMbRMin, kMbAlpha);
  long n = (long)(m + 1) * (m + 1) * (m + 1);
  if (n > c->mbCap) {
    metaballs_release(c);
    WS_HIP(hipMalloc((void **)&c->mbField, sizeof(float) * n));
    WS_HIP(hipMalloc((void **)&c->mbAcc, sizeof(unsigned long long) * n));
    c->mbCap = n;
  }
  c->mbM = m;
  hipLaunchKernelGGL(k_mb_zero, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->mbAcc);
  if (c->nP > 0 && c->spacing > 0)
    hipLaunchKernelGGL(k_mb_splat, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->nP, c->px, c->py, c->pz,
                       c->pflag, m, mdx, R2, rad, c->mbAcc);
  hipLaunchKernelGGL(k_mb_finish, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->mbAcc, c->mbField);
  WS_LAUNCHED();
  if (iso_out) *iso_out = iso;
  if (lattice) *lattice = m;
  return mt_build_dev(c, c->mbField, m, m, m, mdx, iso, pos, nrm, cap, count, volume);
}

}  // extern "C"
