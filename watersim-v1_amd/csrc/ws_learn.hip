// Kernels of the learned pressure warm start: a 3-level 3-D U-Net on
// channels-last bf16 grids, trained and run on the device. Convolutions are
// implicit GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation): one wavefront computes 16 voxels x all output
// channels, the K dimension runs over (tap, input channel) in chunks of 32.
// Weight gradients are the same product with voxels as the K dimension,
// split over voxel chunks and summed in a fixed order (deterministic).
#include "ws_learn.h"
#include "../../include/watersim_learn.h"

namespace ws {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ void tap_offset(int tap, int flip, int &dz, int &dy, int &dx) {
  dz = tap / 9 - 1;
  dy = (tap / 3) % 3 - 1;
  dx = tap % 3 - 1;
  if (flip) { dz = -dz; dy = -dy; dx = -dx; }
}

// out[v][n] = sum_k A[v][k] W[n][k] (+ bias), A[v][tap*Cin + ci] =
// in[v + off(tap)][ci] (zero outside the grid; off negated when flip).
// OUT 0: bf16 with ReLU, 1: bf16, 2: f32. Cin is a multiple of 8 so the
// eight K values of a lane lie in one tap; K is a multiple of 32.
// (block blk of the product: 64 voxels, 16 per wave)
template <int NT, int OUT>
__device__ __forceinline__ void conv3_blk(NetDims d, int Cin, int K, int flip, const bf16_t *__restrict__ in,
                                          const bf16_t *__restrict__ W, const float *__restrict__ bias,
                                          void *__restrict__ out, long blk) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long V = (long)d.Z * d.Y * d.X;
  const long v0 = (blk * 4 + wv) * 16;
  const long va = v0 + (lane & 15);
  const bool okA = va < V;
  const int x = (int)(va % d.X), y = (int)((va / d.X) % d.Y), z = (int)(va / ((long)d.X * d.Y));
  const int kq = 8 * (lane >> 4);
  f4 acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
  const bf16_t *wrow = W + (long)(lane & 15) * K;
  for (int kc = 0; kc < K; kc += 32) {
    const int kk = kc + kq;
    const int tap = kk / Cin, ci = kk - tap * Cin;
    bf8 a = {};
    if (okA && tap < 27) {
      int dz, dy, dx;
      tap_offset(tap, flip, dz, dy, dx);
      const int zz = z + dz, yy = y + dy, xx = x + dx;
      if (zz >= 0 && zz < d.Z && yy >= 0 && yy < d.Y && xx >= 0 && xx < d.X)
        a = *(const bf8 *)(in + (((long)zz * d.Y + yy) * d.X + xx) * Cin + ci);
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const bf8 b = *(const bf8 *)(wrow + (long)nt * 16 * K + kk);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
    }
  }
  constexpr int Cout = NT * 16;
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int n = nt * 16 + (lane & 15);
    const float bn = bias ? bias[n] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const long v = v0 + 4 * (lane >> 4) + i;
      if (v >= V) continue;
      float val = acc[nt][i] + bn;
      if (OUT == 0) val = fmaxf(val, 0.f);
      if (OUT == 2) ((float *)out)[v * Cout + n] = val;
      else ((bf16_t *)out)[v * Cout + n] = f2bf(val);
    }
  }
}

template <int NT, int OUT>
__global__ __launch_bounds__(256) void k_conv3(NetDims d, int Cin, int K, int flip, const bf16_t *__restrict__ in,
                                               const bf16_t *__restrict__ W, const float *__restrict__ bias,
                                               void *__restrict__ out) {
  conv3_blk<NT, OUT>(d, Cin, K, flip, in, W, bias, out, blockIdx.x);
}

// The same product for small grids, where one wave per 16 voxels leaves
// most of the GPU idle: the S waves of a block share 16 voxels, wave w takes
// the steps of 32 K values w, w + S, ..., and the S partial sums are added
// in wave order through LDS (a fixed order: the result does not depend on
// timing).
// (block blk: voxels 16 blk to 16 blk + 15)
template <int NT, int OUT, int S>
__device__ __forceinline__ void conv3s_blk(NetDims d, int Cin, int K, int flip, const bf16_t *__restrict__ in,
                                           const bf16_t *__restrict__ W, const float *__restrict__ bias,
                                           void *__restrict__ out, long blk) {
  __shared__ f4 part[S - 1][NT][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long V = (long)d.Z * d.Y * d.X;
  const long v0 = blk * 16;
  const long va = v0 + (lane & 15);
  const bool okA = va < V;
  const int x = (int)(va % d.X), y = (int)((va / d.X) % d.Y), z = (int)(va / ((long)d.X * d.Y));
  const int kq = 8 * (lane >> 4);
  f4 acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
  const bf16_t *wrow = W + (long)(lane & 15) * K;
#pragma unroll 2
  for (int kc = 32 * wv; kc < K; kc += 32 * S) {
    const int kk = kc + kq;
    const int tap = kk / Cin, ci = kk - tap * Cin;
    bf8 a = {};
    if (okA && tap < 27) {
      int dz, dy, dx;
      tap_offset(tap, flip, dz, dy, dx);
      const int zz = z + dz, yy = y + dy, xx = x + dx;
      if (zz >= 0 && zz < d.Z && yy >= 0 && yy < d.Y && xx >= 0 && xx < d.X)
        a = *(const bf8 *)(in + (((long)zz * d.Y + yy) * d.X + xx) * Cin + ci);
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const bf8 b = *(const bf8 *)(wrow + (long)nt * 16 * K + kk);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
    }
  }
  if (wv > 0) {
#pragma unroll
    for (int nt = 0; nt < NT; nt++) part[wv - 1][nt][lane] = acc[nt];
  }
  __syncthreads();
  if (wv != 0) return;
  for (int w = 1; w < S; w++) {
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[nt] += part[w - 1][nt][lane];
  }
  constexpr int Cout = NT * 16;
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int n = nt * 16 + (lane & 15);
    const float bn = bias ? bias[n] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const long v = v0 + 4 * (lane >> 4) + i;
      if (v >= V) continue;
      float val = acc[nt][i] + bn;
      if (OUT == 0) val = fmaxf(val, 0.f);
      if (OUT == 2) ((float *)out)[v * Cout + n] = val;
      else ((bf16_t *)out)[v * Cout + n] = f2bf(val);
    }
  }
}

template <int NT, int OUT, int S>
__global__ __launch_bounds__(64 * S) void k_conv3s(NetDims d, int Cin, int K, int flip,
                                                   const bf16_t *__restrict__ in, const bf16_t *__restrict__ W,
                                                   const float *__restrict__ bias, void *__restrict__ out) {
  conv3s_blk<NT, OUT, S>(d, Cin, K, flip, in, W, bias, out, blockIdx.x);
}

// The same convolution with the input tile staged in LDS. A block computes
// a 16 x 8 x (MT/2) voxel tile: each of its four waves owns MT rows of 16
// voxels along x and all output channels. The input region (tile plus a
// one-voxel halo) is loaded into LDS one chunk of CH channels at a time
// (CH = 16, or 8 for an 8-channel input), so every input element comes from
// L2 once per block and the 27 taps read it from LDS. One MFMA step covers
// 32 K values: two taps of 16 channels (four taps of 8 when CH = 8).
// (tile blk of the grid, x fastest)
template <int NT, int OUT, int MT>
__device__ __forceinline__ void conv3l_blk(NetDims d, int Cin, int K, int flip, const bf16_t *__restrict__ in,
                                           const bf16_t *__restrict__ W, const float *__restrict__ bias,
                                           void *__restrict__ out, long blk) {
  constexpr int TX = 16, TY = 8, TZ = MT / 2, RX = TX + 2, RY = TY + 2, RZ = TZ + 2, RN = RX * RY * RZ;
  __shared__ uint4 reg[RN * 2];  // RN voxels x 16 channels of bf16 (32 bytes each)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tilesX = (d.X + TX - 1) / TX, tilesY = (d.Y + TY - 1) / TY;
  const int b = (int)blk;
  const int x0 = (b % tilesX) * TX, y0 = ((b / tilesX) % tilesY) * TY, z0 = (b / (tilesX * tilesY)) * TZ;
  const int CH = Cin < 16 ? 8 : 16;  // channels per staged chunk
  const int tpl = CH == 16 ? 2 : 4;  // taps per MFMA step
  const int steps = (27 + tpl - 1) / tpl;
  // this lane's tap within a step and its 8-channel group
  const int tsub = CH == 16 ? (lane >> 5) : (lane >> 4);
  const int cgrp = CH == 16 ? ((lane >> 4) & 1) : 0;
  f4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[m][nt] = f4{0.f, 0.f, 0.f, 0.f};
  const int nchunk = Cin / CH;
  for (int ch = 0; ch < nchunk; ch++) {
    __syncthreads();
    // stage the region: one uint4 (8 channels) per element
    const int per = CH / 8;
    for (int e = threadIdx.x; e < RN * per; e += 256) {
      int r = e / per, hlf = e - r * per;
      int rx = r % RX, ry = (r / RX) % RY, rz = r / (RX * RY);
      int gx = x0 - 1 + rx, gy = y0 - 1 + ry, gz = z0 - 1 + rz;
      uint4 v = {0u, 0u, 0u, 0u};
      if (gx >= 0 && gx < d.X && gy >= 0 && gy < d.Y && gz >= 0 && gz < d.Z)
        v = *(const uint4 *)(in + (((long)gz * d.Y + gy) * d.X + gx) * Cin + ch * CH + hlf * 8);
      reg[r * 2 + hlf] = v;
    }
    __syncthreads();
    for (int st = 0; st < steps; st++) {
      const int tap = st * tpl + tsub;
      const bool tapOk = tap < 27;
      int dz = 0, dy = 0, dx = 0;
      if (tapOk) tap_offset(tap, flip, dz, dy, dx);
      bf8 bw[NT];
      const long kk = (long)tap * Cin + ch * CH + cgrp * 8;
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        bw[nt] = tapOk ? *(const bf8 *)(W + (long)(nt * 16 + (lane & 15)) * K + kk) : bf8{};
      }
#pragma unroll
      for (int m = 0; m < MT; m++) {
        const int row = wv * MT + m;          // row of 16 voxels within the tile
        const int ly = row % TY, lz = row / TY;
        const int rx = (lane & 15) + 1 + dx, ry = ly + 1 + dy, rz = lz + 1 + dz;
        const uint4 av = reg[((rz * RY + ry) * RX + rx) * 2 + cgrp];
        const bf8 a = __builtin_bit_cast(bf8, av);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[nt], acc[m][nt], 0, 0, 0);
      }
    }
  }
  constexpr int Cout = NT * 16;
#pragma unroll
  for (int m = 0; m < MT; m++) {
    const int row = wv * MT + m;
    const int gy = y0 + row % TY, gz = z0 + row / TY;
    if (gy >= d.Y || gz >= d.Z) continue;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const int n = nt * 16 + (lane & 15);
      const float bn = bias ? bias[n] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int gx = x0 + 4 * (lane >> 4) + i;
        if (gx >= d.X) continue;
        const long v = ((long)gz * d.Y + gy) * d.X + gx;
        float val = acc[m][nt][i] + bn;
        if (OUT == 0) val = fmaxf(val, 0.f);
        if (OUT == 2) ((float *)out)[v * Cout + n] = val;
        else ((bf16_t *)out)[v * Cout + n] = f2bf(val);
      }
    }
  }
}

template <int NT, int OUT, int MT>
__global__ __launch_bounds__(256) void k_conv3l(NetDims d, int Cin, int K, int flip, const bf16_t *__restrict__ in,
                                                const bf16_t *__restrict__ W, const float *__restrict__ bias,
                                                void *__restrict__ out) {
  conv3l_blk<NT, OUT, MT>(d, Cin, K, flip, in, W, bias, out, blockIdx.x);
}

// Weight gradient partials: part[chunk][co][k] = sum over the chunk's
// voxels v of gz[v][co] * x[v + off(tap)][ci], k = tap*Cin + ci. The
// product has M = output channels, N = k, and voxels as its K dimension;
// the four waves of a block take interleaved 32-voxel slices of the chunk
// and are summed through LDS in a fixed order.
template <int MT>
__global__ __launch_bounds__(256) void k_wgrad(NetDims d, int Cin, int K, int vchunk, const bf16_t *__restrict__ x,
                                               const bf16_t *__restrict__ gz, float *__restrict__ part) {
  constexpr int Cout = MT * 16;
  __shared__ float red[4][MT][4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = blockIdx.x * 16 + (lane & 15);
  const int tap = col / Cin, ci = col - tap * Cin;
  int dz = 0, dy = 0, dx = 0;
  const bool colOk = tap < 27;
  if (colOk) tap_offset(tap, 0, dz, dy, dx);
  const long V = (long)d.Z * d.Y * d.X;
  const long vbeg = (long)blockIdx.y * vchunk;
  const long vend = vbeg + vchunk < V ? vbeg + vchunk : V;
  f4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; mt++) acc[mt] = f4{0.f, 0.f, 0.f, 0.f};
  for (long v0 = vbeg + 32 * wv; v0 < vend; v0 += 128) {
    bf8 b;
    bf8 a[MT];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const long v = v0 + 8 * (lane >> 4) + t;
      bf16_t xv = 0;
      if (v < vend && colOk) {
        const int xx = (int)(v % d.X) + dx, yy = (int)((v / d.X) % d.Y) + dy, zz = (int)(v / ((long)d.X * d.Y)) + dz;
        if (zz >= 0 && zz < d.Z && yy >= 0 && yy < d.Y && xx >= 0 && xx < d.X)
          xv = x[(((long)zz * d.Y + yy) * d.X + xx) * Cin + ci];
      }
      b[t] = __builtin_bit_cast(__bf16, xv);
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        const bf16_t gv = v < vend ? gz[v * Cout + mt * 16 + (lane & 15)] : (bf16_t)0;
        a[mt][t] = __builtin_bit_cast(__bf16, gv);
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; mt++) acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b, acc[mt], 0, 0, 0);
  }
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int i = 0; i < 4; i++) red[wv][mt][i][lane] = acc[mt][i];
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float s = red[0][mt][i][lane] + red[1][mt][i][lane] + red[2][mt][i][lane] + red[3][mt][i][lane];
        const int co = mt * 16 + 4 * (lane >> 4) + i;
        part[((long)blockIdx.y * Cout + co) * K + col] = s;
      }
  }
}

// grad[W(co, tap, ci)] (+)= sum over chunks of part[chunk][co][tap*Cin_p + ci]
// for ci < Cin (the real input channels), tap < taps; fixed chunk order.
__global__ void k_wgrad_reduce(int nchunk, int Cout, int K, int Cin_p, int Cin, int taps, const float *__restrict__ part,
                               float *__restrict__ grad, int accumulate) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long)Cout * taps * Cin;
  if (e >= n) return;
  int ci = (int)(e % Cin), tap = (int)((e / Cin) % taps), co = (int)(e / ((long)Cin * taps));
  int k = tap * Cin_p + ci;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += part[((long)c * Cout + co) * K + k];
  grad[e] = accumulate ? grad[e] + s : s;
}

// bias gradient: gb[c] (+)= sum_v g[v][c], per-block partials then a fixed-order sum
__global__ void k_colsum_part(long V, int C, const bf16_t *__restrict__ g, float *__restrict__ part) {
  __shared__ float s[256];
  int c = blockIdx.y;
  float acc = 0.f;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x)
    acc += bf2f(g[v * C + c]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)c * gridDim.x + blockIdx.x] = s[0];
}
__global__ void k_colsum_final(int C, int nb, const float *__restrict__ part, float *__restrict__ gb, int accumulate) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int b = 0; b < nb; b++) s += part[(long)c * nb + b];
  gb[c] = accumulate ? gb[c] + s : s;
}

// weights: master FP32 W[co][tap][ci] -> forward bf16 Wf[co][K] (k = tap*Cin_p
// + ci) and input-gradient bf16 Wd[ci][Kd] (k = tap*Cout + co), zero padding
__global__ void k_pack_weights(int Cout, int Cin, int Cin_p, int K, int Kd, const float *__restrict__ w,
                               bf16_t *__restrict__ wf, bf16_t *__restrict__ wd) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nf = (long)Cout * K;
  if (e < nf) {
    int co = (int)(e / K), k = (int)(e % K);
    int tap = k / Cin_p, ci = k - tap * Cin_p;
    float v = (tap < 27 && ci < Cin) ? w[((long)co * 27 + tap) * Cin + ci] : 0.f;
    wf[e] = f2bf(v);
  }
  if (wd) {
    long nd = (long)Cin * Kd;
    if (e < nd) {
      int ci = (int)(e / Kd), k = (int)(e % Kd);
      int tap = k / Cout, co = k - tap * Cout;
      float v = tap < 27 ? w[((long)co * 27 + tap) * Cin + ci] : 0.f;
      wd[e] = f2bf(v);
    }
  }
}

// 2x2x2 average pooling (average in f32, stored bf16)
// (element e: a coarse voxel and group of 8 channels; C is a multiple of 8)
__device__ __forceinline__ void pool_at(NetDims dc, int C, const bf16_t *__restrict__ fine, bf16_t *__restrict__ coarse,
                                        long e) {
  long Vc = (long)dc.Z * dc.Y * dc.X;
  const int G = C >> 3;
  if (e >= Vc * G) return;
  const int grp = (int)(e % G);
  long v = e / G;
  int x = (int)(v % dc.X), y = (int)((v / dc.X) % dc.Y), z = (int)(v / ((long)dc.X * dc.Y));
  int FX = 2 * dc.X, FY = 2 * dc.Y;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int q = 0; q < 8; q++) {
    int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
    const uint4 w = *(const uint4 *)(fine + (((long)zz * FY + yy) * FX + xx) * C + 8 * grp);
    const unsigned u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      s[2 * i] += bf2f((bf16_t)(u[i] & 0xffffu));
      s[2 * i + 1] += bf2f((bf16_t)(u[i] >> 16));
    }
  }
  unsigned o[4];
#pragma unroll
  for (int i = 0; i < 4; i++) o[i] = (unsigned)f2bf(s[2 * i] * 0.125f) | ((unsigned)f2bf(s[2 * i + 1] * 0.125f) << 16);
  *(uint4 *)(coarse + v * C + 8 * grp) = uint4{o[0], o[1], o[2], o[3]};
}
__global__ void k_pool(NetDims dc, int C, const bf16_t *__restrict__ fine, bf16_t *__restrict__ coarse) {
  pool_at(dc, C, fine, coarse, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// nearest upsampling x2 of the coarse tensor concatenated with the skip:
// out[v][0..Cu) = coarse[parent(v)], out[v][Cu..Cu+Cs) = skip[v]; one
// element per fine voxel and group of 8 channels (Cu, Cs multiples of 8)
__device__ __forceinline__ void up_concat_at(NetDims df, int Cu, int Cs, const bf16_t *__restrict__ coarse,
                                             const bf16_t *__restrict__ skip, bf16_t *__restrict__ out, long e) {
  long Vf = (long)df.Z * df.Y * df.X;
  const int C = Cu + Cs, G = C >> 3;
  if (e >= Vf * G) return;
  const int grp = (int)(e % G), c = 8 * grp;
  long v = e / G;
  uint4 w;
  if (c >= Cu) {
    w = *(const uint4 *)(skip + v * Cs + (c - Cu));
  } else {
    int x = (int)(v % df.X), y = (int)((v / df.X) % df.Y), z = (int)(v / ((long)df.X * df.Y));
    long p = ((long)(z >> 1) * (df.Y >> 1) + (y >> 1)) * (df.X >> 1) + (x >> 1);
    w = *(const uint4 *)(coarse + p * Cu + c);
  }
  *(uint4 *)(out + v * C + c) = w;
}
__global__ void k_up_concat(NetDims df, int Cu, int Cs, const bf16_t *__restrict__ coarse,
                            const bf16_t *__restrict__ skip, bf16_t *__restrict__ out) {
  up_concat_at(df, Cu, Cs, coarse, skip, out, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// gz = bf16(g * [y > 0]), g a channel slice [off, off + C) of a row of Cs floats
__global__ void k_relu_back(long V, int C, int Cs, int off, const float *__restrict__ g, const bf16_t *__restrict__ y,
                            bf16_t *__restrict__ gz) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= V * C) return;
  long v = e / C;
  int c = (int)(e % C);
  gz[e] = bf2f(y[e]) > 0.f ? f2bf(g[v * Cs + off + c]) : (bf16_t)0;
}

// gradient through nearest upsampling: the sum over the eight children of
// channels [0, C) of a fine gradient row of Cs floats, then ReLU of yc
__global__ void k_upback_relu(NetDims dc, int C, int Cs, const float *__restrict__ gf, const bf16_t *__restrict__ yc,
                              bf16_t *__restrict__ gz) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long Vc = (long)dc.Z * dc.Y * dc.X;
  if (e >= Vc * C) return;
  int c = (int)(e % C);
  long v = e / C;
  int x = (int)(v % dc.X), y = (int)((v / dc.X) % dc.Y), z = (int)(v / ((long)dc.X * dc.Y));
  int FX = 2 * dc.X, FY = 2 * dc.Y;
  float s = 0.f;
  for (int q = 0; q < 8; q++) {
    int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
    s += gf[(((long)zz * FY + yy) * FX + xx) * Cs + c];
  }
  gz[e] = bf2f(yc[e]) > 0.f ? f2bf(s) : (bf16_t)0;
}

// gradient into a pooled layer's output: its pooled child's gradient / 8
// plus the skip connection's gradient (a channel slice of a row of Cs),
// then ReLU of y
__global__ void k_poolback_add_relu(NetDims df, int C, const float *__restrict__ gc, int Cs, int off,
                                    const float *__restrict__ gskip, const bf16_t *__restrict__ y,
                                    bf16_t *__restrict__ gz) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long Vf = (long)df.Z * df.Y * df.X;
  if (e >= Vf * C) return;
  int c = (int)(e % C);
  long v = e / C;
  int x = (int)(v % df.X), yy = (int)((v / df.X) % df.Y), z = (int)(v / ((long)df.X * df.Y));
  long p = ((long)(z >> 1) * (df.Y >> 1) + (yy >> 1)) * (df.X >> 1) + (x >> 1);
  float g = gc[p * C + c] * 0.125f + gskip[v * Cs + off + c];
  gz[e] = bf2f(y[e]) > 0.f ? f2bf(g) : (bf16_t)0;
}

// 1x1x1 head: out[v] = sum_c w[c] a[v][c] + b
__device__ __forceinline__ void head_at(long V, const bf16_t *__restrict__ a, const float *__restrict__ w,
                                        float *__restrict__ out, long v) {
  if (v >= V) return;
  float s = 0.f;
  for (int c = 0; c < 16; c++) s += bf2f(f2bf(w[c])) * bf2f(a[v * 16 + c]);
  out[v] = s + w[16];
}
__global__ void k_head_fwd(long V, const bf16_t *__restrict__ a, const float *__restrict__ w, float *__restrict__ out) {
  head_at(V, a, w, out, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// head backward: gz9[v][c] = bf16(gout[v] w[c] [a9 > 0]); per-block partial
// sums of gout * a9 (weights) and gout (bias)
__global__ void k_head_back(long V, const bf16_t *__restrict__ a, const float *__restrict__ w,
                            const float *__restrict__ gout, bf16_t *__restrict__ gz, float *__restrict__ part) {
  __shared__ float s[17][256];
  float acc[17];
  for (int c = 0; c < 17; c++) acc[c] = 0.f;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
    float go = gout[v];
    for (int c = 0; c < 16; c++) {
      float av = bf2f(a[v * 16 + c]);
      acc[c] += go * av;
      gz[v * 16 + c] = av > 0.f ? f2bf(go * bf2f(f2bf(w[c]))) : (bf16_t)0;
    }
    acc[16] += go;
  }
  for (int c = 0; c < 17; c++) s[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int c = 0; c < 17; c++) s[c][threadIdx.x] += s[c][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 17) part[(long)blockIdx.x * 17 + threadIdx.x] = s[threadIdx.x][0];
}
__global__ void k_head_back_final(int nb, const float *__restrict__ part, float *__restrict__ g, int accumulate) {
  int c = threadIdx.x;
  if (c >= 17) return;
  float s = 0.f;
  for (int b = 0; b < nb; b++) s += part[(long)b * 17 + c];
  g[c] = accumulate ? g[c] + s : s;
}

// masked squared error normalised by the target energy: block partials of
// sum m (y - t)^2 and sum m t^2
__global__ void k_loss_part(long V, const float *__restrict__ y, const float *__restrict__ t,
                            const uint8_t *__restrict__ m, double *__restrict__ part) {
  __shared__ double se[256], st[256];
  double a = 0, b = 0;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
    if (!m[v]) continue;
    double d = (double)y[v] - t[v];
    a += d * d;
    b += (double)t[v] * t[v];
  }
  se[threadIdx.x] = a;
  st[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) { se[threadIdx.x] += se[threadIdx.x + o]; st[threadIdx.x] += st[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = se[0]; part[2 * blockIdx.x + 1] = st[0]; }
}
__global__ void k_loss_final(int nb, const double *__restrict__ part, double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0, b = 0;
  for (int i = 0; i < nb; i++) { a += part[2 * i]; b += part[2 * i + 1]; }
  out[0] = a;
  out[1] = b;
}
// d loss / d y = 2 m (y - t) / sum m t^2
__global__ void k_loss_grad(long V, const float *__restrict__ y, const float *__restrict__ t,
                            const uint8_t *__restrict__ m, const double *__restrict__ sums, float *__restrict__ g) {
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  double den = sums[1] > 1e-30 ? sums[1] : 1e-30;
  g[v] = m[v] ? (float)(2.0 * ((double)y[v] - t[v]) / den) : 0.f;
}

// coarse-scale term of the loss: the masked error and target averaged over
// 4x4x4 blocks (the smooth part SOR removes slowest), block partials of
// sum e4^2 and sum t4^2
__global__ void k_loss4_part(NetDims d, const float *__restrict__ y, const float *__restrict__ t,
                             const uint8_t *__restrict__ m, float *__restrict__ e4, double *__restrict__ part) {
  __shared__ double se[256], st[256];
  NetDims c4{d.Z / 4, d.Y / 4, d.X / 4};
  long V4 = (long)c4.Z * c4.Y * c4.X;
  double a = 0, b = 0;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V4; v += (long)gridDim.x * blockDim.x) {
    int x = (int)(v % c4.X), yy = (int)((v / c4.X) % c4.Y), z = (int)(v / ((long)c4.X * c4.Y));
    float se4 = 0.f, st4 = 0.f;
    for (int q = 0; q < 64; q++) {
      long f = ((long)(4 * z + (q >> 4)) * d.Y + (4 * yy + ((q >> 2) & 3))) * d.X + (4 * x + (q & 3));
      if (m[f]) { se4 += y[f] - t[f]; st4 += t[f]; }
    }
    se4 *= (1.f / 64.f);
    st4 *= (1.f / 64.f);
    e4[v] = se4;
    a += (double)se4 * se4;
    b += (double)st4 * st4;
  }
  se[threadIdx.x] = a;
  st[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) { se[threadIdx.x] += se[threadIdx.x + o]; st[threadIdx.x] += st[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = se[0]; part[2 * blockIdx.x + 1] = st[0]; }
}
// gout += w * d(sum e4^2 / sum t4^2) / dy = w * m * 2 e4(parent) / 64 / sum t4^2
__global__ void k_loss4_grad(NetDims d, const uint8_t *__restrict__ m, const float *__restrict__ e4,
                             const double *__restrict__ sums, float w, float *__restrict__ g) {
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long V = (long)d.Z * d.Y * d.X;
  if (v >= V || !m[v]) return;
  int x = (int)(v % d.X), y = (int)((v / d.X) % d.Y), z = (int)(v / ((long)d.X * d.Y));
  long p = ((long)(z >> 2) * (d.Y >> 2) + (y >> 2)) * (d.X >> 2) + (x >> 2);
  double den = sums[1] > 1e-30 ? sums[1] : 1e-30;
  g[v] += (float)(w * 2.0 * e4[p] / 64.0 / den);
}

// Adam on the flat parameter vector (FP32 master weights)
__global__ void k_adam(long n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                       float gscale) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float gv = g[e] * gscale;
  float a = b1 * m1[e] + (1.f - b1) * gv;
  float b = b2 * m2[e] + (1.f - b2) * gv * gv;
  m1[e] = a;
  m2[e] = b;
  w[e] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_fill_f32(long n, float *__restrict__ p, float v) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

// host-supplied input channels (f32 [V][5]) -> bf16 [V][8]
__global__ void k_pack_input(long V, const float *__restrict__ in5, bf16_t *__restrict__ x) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= V * 8) return;
  long v = e >> 3;
  int c = (int)(e & 7);
  x[e] = c < 5 ? f2bf(in5[v * 5 + c]) : (bf16_t)0;
}

// a crop of a bank sample: x [S^3][8], target, mask (fluid channel above
// mth: a fluid cell, or on the coarsened grid a block with a fluid cell)
__global__ void k_crop(NetDims full, NetDims cd, int oz, int oy, int ox, const bf16_t *__restrict__ sx,
                       const float *__restrict__ st, bf16_t *__restrict__ x, float *__restrict__ t,
                       uint8_t *__restrict__ m, float mth) {
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long Vc = (long)cd.Z * cd.Y * cd.X;
  if (v >= Vc) return;
  int xx = (int)(v % cd.X) + ox, yy = (int)((v / cd.X) % cd.Y) + oy, zz = (int)(v / ((long)cd.X * cd.Y)) + oz;
  long f = ((long)zz * full.Y + yy) * full.X + xx;
  const uint4 *src = (const uint4 *)(sx + f * 8);
  *(uint4 *)(x + v * 8) = *src;
  t[v] = st[f];
  m[v] = bf2f(sx[f * 8 + 2]) > mth ? 1 : 0;
}

// ---- the solver side: inputs from the pressure system, the correction
// initial residual r0 = rhs - A q over fluid cells and FP64 sums for the
// two scales (sum r0^2, sum q^2, fluid count)
__device__ __forceinline__ void k_ws_residual_body(long bx, long nbx, NetDims d, const uint8_t *__restrict__ ct,
                                                   const uint8_t *__restrict__ nbm, const uint8_t *__restrict__ dg,
                                                   const float *__restrict__ rhs, const float *__restrict__ q,
                                                   float *__restrict__ r0, double *__restrict__ part) {
  __shared__ double sa[256], sb[256], sc[256];
  double a = 0, b = 0, n = 0;
  long C = (long)d.Z * d.Y * d.X;
  int nx = d.X, nxy = d.X * d.Y;
  for (long c = (long)bx * blockDim.x + threadIdx.x; c < C; c += (long)nbx * blockDim.x) {
    unsigned mk = nbm[c];
    if (!(mk & 0x40)) { r0[c] = 0.f; continue; }
    const float sum = nb_sum(q, c, mk, nx, nxy);
    float r = rhs[c] - ((float)dg[c] * q[c] - sum);
    r0[c] = r;
    a += (double)r * r;
    b += (double)q[c] * q[c];
    n += 1;
  }
  sa[threadIdx.x] = a; sb[threadIdx.x] = b; sc[threadIdx.x] = n;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sa[threadIdx.x] += sa[threadIdx.x + o]; sb[threadIdx.x] += sb[threadIdx.x + o]; sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[3 * bx] = sa[0]; part[3 * bx + 1] = sb[0]; part[3 * bx + 2] = sc[0]; }
}
__global__ void k_ws_residual(NetDims d, const uint8_t *__restrict__ ct, const uint8_t *__restrict__ nbm,
                              const uint8_t *__restrict__ dg, const float *__restrict__ rhs,
                              const float *__restrict__ q, float *__restrict__ r0, double *__restrict__ part) {
  k_ws_residual_body(blockIdx.x, gridDim.x, d, ct, nbm, dg, rhs, q, r0, part);
}
WS_PLANETS_OPS(k_ws_residual_ops, k_ws_residual, k_ws_residual_body)
// the block partials summed in a fixed order by 256 threads (thread t takes
// partials t, t + 256, ... in turn, then a pairwise tree); nb <= 1024
__device__ __forceinline__ void k_ws_scales_body(long bx, long nbx, int nb, const double *__restrict__ part,
                                                 float *__restrict__ scales) {
  __shared__ double sa[256], sb[256], sn[256];
  double a = 0, b = 0, n = 0;
  for (int i = threadIdx.x; i < nb; i += 256) { a += part[3 * i]; b += part[3 * i + 1]; n += part[3 * i + 2]; }
  sa[threadIdx.x] = a; sb[threadIdx.x] = b; sn[threadIdx.x] = n;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sa[threadIdx.x] += sa[threadIdx.x + o]; sb[threadIdx.x] += sb[threadIdx.x + o]; sn[threadIdx.x] += sn[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  a = sa[0]; b = sb[0]; n = sn[0];
  double s = n > 0 ? sqrt(a / n) : 1.0, sq = n > 0 ? sqrt(b / n) : 1.0;
  scales[0] = s > 0 ? (float)s : 1.f;
  scales[1] = sq > 0 ? (float)sq : 1.f;
}
__global__ __launch_bounds__(256) void k_ws_scales(int nb, const double *__restrict__ part,
                                                   float *__restrict__ scales) {
  k_ws_scales_body(blockIdx.x, gridDim.x, nb, part, scales);
}
WS_PLANETS_OPS(k_ws_scales_ops, k_ws_scales, k_ws_scales_body)
// the five input channels [r0/s, air, fluid, solid, q/sq] as bf16 [V][8]
__device__ __forceinline__ void k_ws_inputs_body(long bx, long nbx, long C, const uint8_t *__restrict__ ct,
                                                 const float *__restrict__ r0, const float *__restrict__ q,
                                                 const float *__restrict__ scales, bf16_t *__restrict__ x) {
  long c = (long)bx * blockDim.x + threadIdx.x;
  if (c >= C) return;
  uint8_t t = ct[c];
  float s = scales[0], sq = scales[1];
  bool fl = t == FLUID;
  bf16_t v[8] = {fl ? f2bf(r0[c] / s) : (bf16_t)0, f2bf(t == AIR ? 1.f : 0.f), f2bf(fl ? 1.f : 0.f),
                 f2bf(t == SOLID ? 1.f : 0.f), fl ? f2bf(q[c] / sq) : (bf16_t)0, 0, 0, 0};
  for (int i = 0; i < 8; i++) x[c * 8 + i] = v[i];
}
__global__ void k_ws_inputs(long C, const uint8_t *__restrict__ ct, const float *__restrict__ r0,
                            const float *__restrict__ q, const float *__restrict__ scales, bf16_t *__restrict__ x) {
  k_ws_inputs_body(blockIdx.x, gridDim.x, C, ct, r0, q, scales, x);
}
WS_PLANETS_OPS(k_ws_inputs_ops, k_ws_inputs, k_ws_inputs_body)
// the five input channels of the grid coarsened 2x per axis: each coarse
// cell averages its eight cells' full-resolution channels (r0/s and q/sq on
// fluid cells, 0 elsewhere; the type one-hots, so types become fractions)
__device__ __forceinline__ void k_ws_inputs_coarse_body(long bx, long nbx, NetDims cd, const uint8_t *__restrict__ ct,
                                                        const float *__restrict__ r0, const float *__restrict__ q,
                                                        const float *__restrict__ scales, bf16_t *__restrict__ x) {
  long v = (long)bx * blockDim.x + threadIdx.x;
  long V = (long)cd.Z * cd.Y * cd.X;
  if (v >= V) return;
  const int X = 2 * cd.X, XY = 4 * cd.X * cd.Y;
  const int cx = (int)(v % cd.X), cy = (int)((v / cd.X) % cd.Y), cz = (int)(v / ((long)cd.X * cd.Y));
  const float s = scales[0], sq = scales[1];
  float a[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  for (int o = 0; o < 8; o++) {
    long c = (long)(2 * cz + (o >> 2)) * XY + (long)(2 * cy + ((o >> 1) & 1)) * X + (2 * cx + (o & 1));
    const uint8_t t = ct[c];
    const bool fl = t == FLUID;
    a[0] += fl ? r0[c] / s : 0.f;
    a[1] += t == AIR ? 1.f : 0.f;
    a[2] += fl ? 1.f : 0.f;
    a[3] += t == SOLID ? 1.f : 0.f;
    a[4] += fl ? q[c] / sq : 0.f;
  }
  for (int i = 0; i < 8; i++) x[v * 8 + i] = i < 5 ? f2bf(a[i] * 0.125f) : (bf16_t)0;
}
__global__ void k_ws_inputs_coarse(NetDims cd, const uint8_t *__restrict__ ct, const float *__restrict__ r0,
                                   const float *__restrict__ q, const float *__restrict__ scales,
                                   bf16_t *__restrict__ x) {
  k_ws_inputs_coarse_body(blockIdx.x, gridDim.x, cd, ct, r0, q, scales, x);
}
WS_PLANETS_OPS(k_ws_inputs_coarse_ops, k_ws_inputs_coarse, k_ws_inputs_coarse_body)
// q += s y(parent) on fluid cells: the coarse prediction, constant over
// each 2x2x2 block
__device__ __forceinline__ void apply_coarse_at(NetDims cd, const uint8_t *__restrict__ ct, const float *__restrict__ y,
                                                const float *__restrict__ scales, float *__restrict__ q, long c) {
  const long C = 8L * cd.Z * cd.Y * cd.X;
  if (c >= C || ct[c] != FLUID) return;
  const int X = 2 * cd.X, Y = 2 * cd.Y;
  const int x = (int)(c % X), yy = (int)((c / X) % Y), z = (int)(c / ((long)X * Y));
  q[c] = q[c] + scales[0] * y[((long)(z >> 1) * cd.Y + (yy >> 1)) * cd.X + (x >> 1)];
}
__global__ void k_ws_apply_coarse(NetDims cd, const uint8_t *__restrict__ ct, const float *__restrict__ y,
                                  const float *__restrict__ scales, float *__restrict__ q) {
  apply_coarse_at(cd, ct, y, scales, q, (long)blockIdx.x * blockDim.x + threadIdx.x);
}
// training target on the coarsened grid: the mean of (q* - q_prev) / s over
// each block's fluid cells (0 for a block without one), the value that a
// block-constant correction fits best in least squares
__global__ void k_ws_target_coarse(NetDims cd, const uint8_t *__restrict__ ct, const float *__restrict__ q,
                                   const float *__restrict__ qprev, const float *__restrict__ scales,
                                   float *__restrict__ t) {
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long V = (long)cd.Z * cd.Y * cd.X;
  if (v >= V) return;
  const int X = 2 * cd.X, XY = 4 * cd.X * cd.Y;
  const int cx = (int)(v % cd.X), cy = (int)((v / cd.X) % cd.Y), cz = (int)(v / ((long)cd.X * cd.Y));
  float sum = 0.f;
  int n = 0;
  for (int o = 0; o < 8; o++) {
    long c = (long)(2 * cz + (o >> 2)) * XY + (long)(2 * cy + ((o >> 1) & 1)) * X + (2 * cx + (o & 1));
    if (ct[c] == FLUID) { sum += q[c] - qprev[c]; n++; }
  }
  t[v] = n ? sum / (float)n / scales[0] : 0.f;
}
// q += s y on fluid cells (the prediction of the correction)
__device__ __forceinline__ void apply_at(long C, const uint8_t *__restrict__ ct, const float *__restrict__ y,
                                         const float *__restrict__ scales, float *__restrict__ q, long c) {
  if (c >= C) return;
  if (ct[c] == FLUID) q[c] = q[c] + scales[0] * y[c];
}
__global__ void k_ws_apply(long C, const uint8_t *__restrict__ ct, const float *__restrict__ y,
                           const float *__restrict__ scales, float *__restrict__ q) {
  apply_at(C, ct, y, scales, q, (long)blockIdx.x * blockDim.x + threadIdx.x);
}
// training target of a recorded sample: (q* - q_prev) / s on fluid cells
__global__ void k_ws_target(long C, const uint8_t *__restrict__ ct, const float *__restrict__ q,
                            const float *__restrict__ qprev, const float *__restrict__ scales,
                            float *__restrict__ t) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  t[c] = ct[c] == FLUID ? (q[c] - qprev[c]) / scales[0] : 0.f;
}

// ---- the forward pass of many planets at once
// Each step of the forward pass (a convolution, a pooling, an upsampling,
// the head, the correction) runs as one launch over every planet: planet j
// owns the launch's blocks [starts[j], starts[j + 1]) and its block k there
// does what block k of its own launch would do, so the results equal the
// planets' separate passes bit for bit.
struct NetJob {
  NetDims d;  // the grid of the step (for the correction: the network's grid)
  const void *in, *in2;  // input, and the skip tensor of an upsampling
  const bf16_t *W;
  const float *b;  // bias; the head's weights
  void *out;       // output; the pressure of a correction
  const uint8_t *ct;
  const float *scales;
};

// the planet that owns block b: the last j with starts[j] <= b (starts
// ascending), as a wave-uniform value
__device__ __forceinline__ int net_job_of(const long *__restrict__ starts, int nj, long b) {
  const int lane = threadIdx.x & 63;
  int k = 0;
  for (int base = 0; base < nj; base += 64) {
    const int j = base + lane;
    k += __popcll(__ballot(j < nj && starts[j] <= b));
  }
  return __builtin_amdgcn_readfirstlane(k - 1);
}

// a convolution layer over many planets (forward, bf16 + ReLU); V selects the
// form each planet's own pass uses: 0 untiled, 1 small-grid, 2 LDS-tiled
template <int NT, int V>
__global__ __launch_bounds__(256) void k_conv_batch(const NetJob *__restrict__ J, const long *__restrict__ starts,
                                                    int nj, int Cin, int K) {
  const int j = net_job_of(starts, nj, blockIdx.x);
  const NetJob q = J[j];
  const long blk = (long)blockIdx.x - starts[j];
  const bf16_t *in = (const bf16_t *)q.in;
  if constexpr (V == 0) conv3_blk<NT, 0>(q.d, Cin, K, 0, in, q.W, q.b, q.out, blk);
  else if constexpr (V == 1) conv3s_blk<NT, 0, 4>(q.d, Cin, K, 0, in, q.W, q.b, q.out, blk);
  else conv3l_blk<NT, 0, NT >= 3 ? 4 : 8>(q.d, Cin, K, 0, in, q.W, q.b, q.out, blk);
}

// the element-wise steps over many planets, 256 elements per block:
// 0 pooling (a channels), 1 upsampling with concatenation (a coarse, a2 skip
// channels), 2 the head, 3 the correction from a coarsened grid, 4 the
// correction
__global__ __launch_bounds__(256) void k_net_elem_batch(const NetJob *__restrict__ J, const long *__restrict__ starts,
                                                        int nj, int kind, int a, int a2) {
  const int j = net_job_of(starts, nj, blockIdx.x);
  const NetJob q = J[j];
  const long e = ((long)blockIdx.x - starts[j]) * 256 + threadIdx.x;
  const long V = (long)q.d.Z * q.d.Y * q.d.X;
  switch (kind) {
    case 0: pool_at(q.d, a, (const bf16_t *)q.in, (bf16_t *)q.out, e); break;
    case 1: up_concat_at(q.d, a, a2, (const bf16_t *)q.in, (const bf16_t *)q.in2, (bf16_t *)q.out, e); break;
    case 2: head_at(V, (const bf16_t *)q.in, q.b, (float *)q.out, e); break;
    case 3: apply_coarse_at(q.d, q.ct, (const float *)q.in, q.scales, (float *)q.out, e); break;
    default: apply_at(V, q.ct, (const float *)q.in, q.scales, (float *)q.out, e); break;
  }
}

// explicit instantiations
#define CONV(NT, OUT)                                                                                     \
  template __global__ void k_conv3<NT, OUT>(NetDims, int, int, int, const bf16_t *, const bf16_t *, const float *, \
                                            void *);
CONV(1, 0) CONV(2, 0) CONV(4, 0) CONV(1, 2) CONV(2, 2) CONV(3, 2) CONV(4, 2) CONV(6, 2)
#undef CONV
#define CONVS(NT, OUT)                                                                                     \
  template __global__ void k_conv3s<NT, OUT, 4>(NetDims, int, int, int, const bf16_t *, const bf16_t *,     \
                                                const float *, void *);
CONVS(1, 0) CONVS(2, 0) CONVS(4, 0) CONVS(1, 2) CONVS(2, 2) CONVS(3, 2) CONVS(4, 2) CONVS(6, 2)
#undef CONVS
#define CONVL(NT, OUT, MT)                                                                                    \
  template __global__ void k_conv3l<NT, OUT, MT>(NetDims, int, int, int, const bf16_t *, const bf16_t *,     \
                                                 const float *, void *);
CONVL(1, 0, 8) CONVL(2, 0, 8) CONVL(4, 0, 4) CONVL(1, 2, 8) CONVL(2, 2, 8) CONVL(3, 2, 4) CONVL(4, 2, 4)
CONVL(6, 2, 4)
#undef CONVL
template __global__ void k_wgrad<1>(NetDims, int, int, int, const bf16_t *, const bf16_t *, float *);
template __global__ void k_wgrad<2>(NetDims, int, int, int, const bf16_t *, const bf16_t *, float *);
template __global__ void k_wgrad<4>(NetDims, int, int, int, const bf16_t *, const bf16_t *, float *);

// ====================================================================== host
namespace {

const int kLayerDef[kNetLayers][4] = {  // cin, cinp, cout, taps (oracle/unet_ref.py LAYERS)
    {5, 8, 16, 27},   {16, 16, 16, 27}, {16, 16, 32, 27}, {32, 32, 32, 27}, {32, 32, 64, 27}, {64, 64, 64, 27},
    {96, 96, 32, 27}, {32, 32, 32, 27}, {48, 48, 16, 27}, {16, 16, 16, 27}, {16, 16, 1, 1}};

int round32(int k) { return (k + 31) / 32 * 32; }

template <typename T>
int nalloc(ws_ctx *c, Net *n, T **p, size_t count) {
  hipError_t e = hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
  if (e != hipSuccess) {
    c->err = std::string("hipMalloc (warm-start network): ") + hipGetErrorString(e);
    return WS_EHIP;
  }
  n->owned.push_back(*p);
  return WS_OK;
}

void release(Net *n, void *p) {
  if (!p) return;
  (void)hipFree(p);
  for (auto &q : n->owned)
    if (q == p) q = nullptr;
}

long vox(NetDims d) { return (long)d.Z * d.Y * d.X; }
NetDims half(NetDims d) { return NetDims{d.Z / 2, d.Y / 2, d.X / 2}; }

// the form of a product with NT x 16 output channels on grid d: 0 untiled,
// 1 four waves per 16 voxels with K split (netTiled 1, on grids with too few
// LDS tiles to fill the GPU alone), 2 LDS-tiled; and its blocks
int conv_form(const ws_ctx *c, NetDims d, int NT) {
  if (!c->netTiled) return 0;
  if (c->netTiled == 2) return 2;
  const int MT = NT >= 3 ? 4 : 8;
  const long tiles = (long)((d.X + 15) / 16) * ((d.Y + 7) / 8) * ((d.Z + MT / 2 - 1) / (MT / 2));
  return tiles < 256 ? 1 : 2;
}
long conv_blocks(int form, NetDims d, int NT) {
  const long V = vox(d);
  if (form == 0) return blocks_for(V, 64);
  if (form == 1) return blocks_for(V, 16);
  const int MT = NT >= 3 ? 4 : 8;
  return (long)((d.X + 15) / 16) * ((d.Y + 7) / 8) * ((d.Z + MT / 2 - 1) / (MT / 2));
}

template <int NT, int OUT>
void conv(ws_ctx *c, NetDims d, int cin, int K, int flip, const bf16_t *in, const bf16_t *W, const float *b, void *out) {
  const int form = conv_form(c, d, NT);
  const dim3 grid(conv_blocks(form, d, NT));
  constexpr int MT = NT >= 3 ? 4 : 8;
  if (form == 0)
    hipLaunchKernelGGL((k_conv3<NT, OUT>), grid, dim3(256), 0, c->stream, d, cin, K, flip, in, W, b, out);
  else if (form == 1)
    hipLaunchKernelGGL((k_conv3s<NT, OUT, 4>), grid, dim3(256), 0, c->stream, d, cin, K, flip, in, W, b, out);
  else
    hipLaunchKernelGGL((k_conv3l<NT, OUT, MT>), grid, dim3(256), 0, c->stream, d, cin, K, flip, in, W, b, out);
}

// forward conv of layer l (bf16 + ReLU)
void conv_fwd(ws_ctx *c, int l, NetDims d, const bf16_t *in, bf16_t *out) {
  Net *n = c->net;
  const NetLayer &L = n->L[l];
  const float *b = n->w + L.boff;
  switch (L.cout / 16) {
    case 1: conv<1, 0>(c, d, L.cinp, L.K, 0, in, L.wf, b, out); break;
    case 2: conv<2, 0>(c, d, L.cinp, L.K, 0, in, L.wf, b, out); break;
    default: conv<4, 0>(c, d, L.cinp, L.K, 0, in, L.wf, b, out); break;
  }
}

// input gradient of layer l: f32 [V][cinp] from gz [V][cout]
void conv_dgrad(ws_ctx *c, int l, NetDims d, const bf16_t *gz, float *gx) {
  Net *n = c->net;
  const NetLayer &L = n->L[l];
  switch (L.cinp / 16) {
    case 1: conv<1, 2>(c, d, L.cout, L.Kd, 1, gz, L.wd, nullptr, gx); break;
    case 2: conv<2, 2>(c, d, L.cout, L.Kd, 1, gz, L.wd, nullptr, gx); break;
    case 3: conv<3, 2>(c, d, L.cout, L.Kd, 1, gz, L.wd, nullptr, gx); break;
    case 4: conv<4, 2>(c, d, L.cout, L.Kd, 1, gz, L.wd, nullptr, gx); break;
    default: conv<6, 2>(c, d, L.cout, L.Kd, 1, gz, L.wd, nullptr, gx); break;
  }
}

// weight and bias gradients of layer l from its input x and gz
int conv_wgrad(ws_ctx *c, int l, NetDims d, const bf16_t *x, const bf16_t *gz, int accumulate) {
  Net *n = c->net;
  const NetLayer &L = n->L[l];
  long V = vox(d);
  int vchunk = 4096;
  while ((V + vchunk - 1) / vchunk > 64) vchunk *= 2;
  int nchunk = (int)((V + vchunk - 1) / vchunk);
  if ((long)nchunk * L.cout * L.K > n->partN) {
    c->err = "warm-start network: gradient scratch too small";
    return WS_EINVAL;
  }
  dim3 grid(L.K / 16, nchunk);
  switch (L.cout / 16) {
    case 1: hipLaunchKernelGGL(k_wgrad<1>, grid, dim3(256), 0, c->stream, d, L.cinp, L.K, vchunk, x, gz, n->part); break;
    case 2: hipLaunchKernelGGL(k_wgrad<2>, grid, dim3(256), 0, c->stream, d, L.cinp, L.K, vchunk, x, gz, n->part); break;
    default: hipLaunchKernelGGL(k_wgrad<4>, grid, dim3(256), 0, c->stream, d, L.cinp, L.K, vchunk, x, gz, n->part); break;
  }
  long nw = (long)L.cout * L.taps * L.cin;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(nw, 256)), dim3(256), 0, c->stream, nchunk, L.cout, L.K, L.cinp,
                     L.cin, L.taps, n->part, n->g + L.woff, accumulate);
  // bias: column sums of gz (the scratch after the weight partials)
  float *bpart = n->part + (long)nchunk * L.cout * L.K;
  hipLaunchKernelGGL(k_colsum_part, dim3(64, L.cout), dim3(256), 0, c->stream, V, L.cout, gz, bpart);
  hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(64), 0, c->stream, L.cout, 64, bpart, n->g + L.boff, accumulate);
  WS_LAUNCHED();
  return WS_OK;
}

}  // namespace

long net_param_count() {
  long p = 0;
  for (int l = 0; l < kNetLayers; l++) p += (long)kLayerDef[l][2] * kLayerDef[l][3] * kLayerDef[l][0] + kLayerDef[l][2];
  return p;
}

int net_ensure(ws_ctx *c) {
  if (c->net) return WS_OK;
  Net *n = new Net();
  c->net = n;
  long off = 0;
  for (int l = 0; l < kNetLayers; l++) {
    NetLayer &L = n->L[l];
    L.cin = kLayerDef[l][0]; L.cinp = kLayerDef[l][1]; L.cout = kLayerDef[l][2]; L.taps = kLayerDef[l][3];
    L.woff = off;
    off += (long)L.cout * L.taps * L.cin;
    L.boff = off;
    off += L.cout;
    L.K = round32(27 * L.cinp);
    L.Kd = round32(27 * L.cout);
  }
  n->nparam = off;
  int r;
  if ((r = nalloc(c, n, &n->w, off)) || (r = nalloc(c, n, &n->g, off)) || (r = nalloc(c, n, &n->m1, off)) ||
      (r = nalloc(c, n, &n->m2, off)) || (r = nalloc(c, n, &n->sums, 4096)) || (r = nalloc(c, n, &n->scales, 4)) ||
      (r = nalloc(c, n, &n->rpart, 3 * 1024)))
    return r;
  // zeroed on the context's stream (a null-stream memset would not be
  // ordered before the stream's later kernels)
  for (float *z : {n->w, n->m1, n->m2})
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(off, 256)), dim3(256), 0, c->stream, off, z, 0.f);
  WS_LAUNCHED();
  for (int l = 0; l < kNetLayers - 1; l++) {
    NetLayer &L = n->L[l];
    if ((r = nalloc(c, n, &L.wf, (size_t)L.cout * L.K))) return r;
    if (l > 0 && (r = nalloc(c, n, &L.wd, (size_t)L.cinp * L.Kd))) return r;
  }
  return WS_OK;
}

int net_pack(ws_ctx *c) {
  Net *n = c->net;
  for (int l = 0; l < kNetLayers - 1; l++) {
    NetLayer &L = n->L[l];
    long cnt = (long)L.cout * L.K;
    long cd = (long)L.cinp * L.Kd;
    long m = cnt > cd ? cnt : cd;
    hipLaunchKernelGGL(k_pack_weights, dim3(blocks_for(m, 256)), dim3(256), 0, c->stream, L.cout, L.cin, L.cinp, L.K,
                       L.Kd, n->w + L.woff, L.wf, l > 0 ? L.wd : nullptr);
  }
  WS_LAUNCHED();
  n->packed = true;
  return WS_OK;
}

// activation buffers for grids of up to V voxels (three levels)
static int net_buffers(ws_ctx *c, long V) {
  Net *n = c->net;
  if (V <= n->capV) return WS_OK;
  bf16_t **bb[] = {&n->x0, &n->a0, &n->s1, &n->p1, &n->a2, &n->s2, &n->p2, &n->a4, &n->a5, &n->u1,
                   &n->a6, &n->a7, &n->u2, &n->a8, &n->a9, &n->gz};
  for (auto p : bb) { release(n, *p); *p = nullptr; }
  float **ff[] = {&n->out, &n->gout, &n->tgt, &n->gA, &n->gB, &n->gC, &n->gD, &n->gE, &n->gF, &n->gG, &n->part};
  for (auto p : ff) { release(n, *p); *p = nullptr; }
  release(n, n->mask);
  n->mask = nullptr;
  long h = V / 8, q = V / 64;
  int r;
  if ((r = nalloc(c, n, &n->x0, V * 8)) || (r = nalloc(c, n, &n->a0, V * 16)) || (r = nalloc(c, n, &n->s1, V * 16)) ||
      (r = nalloc(c, n, &n->p1, h * 16)) || (r = nalloc(c, n, &n->a2, h * 32)) || (r = nalloc(c, n, &n->s2, h * 32)) ||
      (r = nalloc(c, n, &n->p2, q * 32)) || (r = nalloc(c, n, &n->a4, q * 64)) || (r = nalloc(c, n, &n->a5, q * 64)) ||
      (r = nalloc(c, n, &n->u1, h * 96)) || (r = nalloc(c, n, &n->a6, h * 32)) || (r = nalloc(c, n, &n->a7, h * 32)) ||
      (r = nalloc(c, n, &n->u2, V * 48)) || (r = nalloc(c, n, &n->a8, V * 16)) || (r = nalloc(c, n, &n->a9, V * 16)) ||
      (r = nalloc(c, n, &n->gz, V * 16)) || (r = nalloc(c, n, &n->out, V)) || (r = nalloc(c, n, &n->gout, V)) ||
      (r = nalloc(c, n, &n->tgt, V)) || (r = nalloc(c, n, &n->mask, V)) || (r = nalloc(c, n, &n->gA, V * 48)) ||
      (r = nalloc(c, n, &n->gB, V * 16)) || (r = nalloc(c, n, &n->gC, h * 32)) || (r = nalloc(c, n, &n->gD, h * 96)) ||
      (r = nalloc(c, n, &n->gE, q * 64)) || (r = nalloc(c, n, &n->gF, q * 32)) || (r = nalloc(c, n, &n->gG, h * 16)))
    return r;
  // weight-gradient partials: at most 64 chunks of the widest product, plus
  // the bias partials
  long pw = 0;
  for (int l = 0; l < kNetLayers - 1; l++) {
    long x = 64L * n->L[l].cout * n->L[l].K + 64L * n->L[l].cout;
    pw = x > pw ? x : pw;
  }
  if ((r = nalloc(c, n, &n->part, pw))) return r;
  n->partN = pw;
  n->capV = V;
  return WS_OK;
}

int net_forward(ws_ctx *c, NetDims d) {
  Net *n = c->net;
  if (!n->packed) { int r = net_pack(c); if (r) return r; }
  NetDims h = half(d), q = half(h);
  long V = vox(d), Vh = vox(h), Vq = vox(q);
  conv_fwd(c, 0, d, n->x0, n->a0);
  conv_fwd(c, 1, d, n->a0, n->s1);
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(Vh * 2, 256)), dim3(256), 0, c->stream, h, 16, n->s1, n->p1);
  conv_fwd(c, 2, h, n->p1, n->a2);
  conv_fwd(c, 3, h, n->a2, n->s2);
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(Vq * 4, 256)), dim3(256), 0, c->stream, q, 32, n->s2, n->p2);
  conv_fwd(c, 4, q, n->p2, n->a4);
  conv_fwd(c, 5, q, n->a4, n->a5);
  hipLaunchKernelGGL(k_up_concat, dim3(blocks_for(Vh * 12, 256)), dim3(256), 0, c->stream, h, 64, 32, n->a5, n->s2,
                     n->u1);
  conv_fwd(c, 6, h, n->u1, n->a6);
  conv_fwd(c, 7, h, n->a6, n->a7);
  hipLaunchKernelGGL(k_up_concat, dim3(blocks_for(V * 6, 256)), dim3(256), 0, c->stream, d, 32, 16, n->a7, n->s1,
                     n->u2);
  conv_fwd(c, 8, d, n->u2, n->a8);
  conv_fwd(c, 9, d, n->a8, n->a9);
  hipLaunchKernelGGL(k_head_fwd, dim3(blocks_for(V, 256)), dim3(256), 0, c->stream, V, n->a9, n->w + n->L[10].woff,
                     n->out);
  WS_LAUNCHED();
  return WS_OK;
}

// gradients of the loss whose d/d out is in gout (accumulated into g when
// accumulate); the layer inputs and outputs of the last forward are used
int net_backward(ws_ctx *c, NetDims d, int acc) {
  Net *n = c->net;
  NetDims h = half(d), q = half(h);
  long V = vox(d), Vh = vox(h), Vq = vox(q);
  int r;
  // head, and c9's output gradient through its ReLU
  hipLaunchKernelGGL(k_head_back, dim3(64), dim3(256), 0, c->stream, V, n->a9, n->w + n->L[10].woff, n->gout, n->gz,
                     n->part);
  hipLaunchKernelGGL(k_head_back_final, dim3(1), dim3(64), 0, c->stream, 64, n->part, n->g + n->L[10].woff, acc);
  WS_LAUNCHED();
  // c9 (input a8)
  if ((r = conv_wgrad(c, 9, d, n->a8, n->gz, acc))) return r;
  conv_dgrad(c, 9, d, n->gz, n->gB);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(V * 16, 256)), dim3(256), 0, c->stream, V, 16, 16, 0, n->gB, n->a8,
                     n->gz);
  // c8 (input u2 = [up(a7) 32, s1 16])
  if ((r = conv_wgrad(c, 8, d, n->u2, n->gz, acc))) return r;
  conv_dgrad(c, 8, d, n->gz, n->gA);
  hipLaunchKernelGGL(k_upback_relu, dim3(blocks_for(Vh * 32, 256)), dim3(256), 0, c->stream, h, 32, 48, n->gA, n->a7,
                     n->gz);
  // c7 (input a6)
  if ((r = conv_wgrad(c, 7, h, n->a6, n->gz, acc))) return r;
  conv_dgrad(c, 7, h, n->gz, n->gC);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(Vh * 32, 256)), dim3(256), 0, c->stream, Vh, 32, 32, 0, n->gC, n->a6,
                     n->gz);
  // c6 (input u1 = [up(a5) 64, s2 32])
  if ((r = conv_wgrad(c, 6, h, n->u1, n->gz, acc))) return r;
  conv_dgrad(c, 6, h, n->gz, n->gD);
  hipLaunchKernelGGL(k_upback_relu, dim3(blocks_for(Vq * 64, 256)), dim3(256), 0, c->stream, q, 64, 96, n->gD, n->a5,
                     n->gz);
  // c5 (input a4)
  if ((r = conv_wgrad(c, 5, q, n->a4, n->gz, acc))) return r;
  conv_dgrad(c, 5, q, n->gz, n->gE);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(Vq * 64, 256)), dim3(256), 0, c->stream, Vq, 64, 64, 0, n->gE, n->a4,
                     n->gz);
  // c4 (input p2)
  if ((r = conv_wgrad(c, 4, q, n->p2, n->gz, acc))) return r;
  conv_dgrad(c, 4, q, n->gz, n->gF);
  // c3: pooled gradient / 8 plus the skip slice [64, 96) of gD, ReLU of s2
  hipLaunchKernelGGL(k_poolback_add_relu, dim3(blocks_for(Vh * 32, 256)), dim3(256), 0, c->stream, h, 32, n->gF, 96,
                     64, n->gD, n->s2, n->gz);
  if ((r = conv_wgrad(c, 3, h, n->a2, n->gz, acc))) return r;
  conv_dgrad(c, 3, h, n->gz, n->gC);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(Vh * 32, 256)), dim3(256), 0, c->stream, Vh, 32, 32, 0, n->gC, n->a2,
                     n->gz);
  // c2 (input p1)
  if ((r = conv_wgrad(c, 2, h, n->p1, n->gz, acc))) return r;
  conv_dgrad(c, 2, h, n->gz, n->gG);
  // c1: pooled gradient / 8 plus the skip slice [32, 48) of gA, ReLU of s1
  hipLaunchKernelGGL(k_poolback_add_relu, dim3(blocks_for(V * 16, 256)), dim3(256), 0, c->stream, d, 16, n->gG, 48, 32,
                     n->gA, n->s1, n->gz);
  if ((r = conv_wgrad(c, 1, d, n->a0, n->gz, acc))) return r;
  conv_dgrad(c, 1, d, n->gz, n->gB);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(V * 16, 256)), dim3(256), 0, c->stream, V, 16, 16, 0, n->gB, n->a0,
                     n->gz);
  // c0 (input x0): weights only
  if ((r = conv_wgrad(c, 0, d, n->x0, n->gz, acc))) return r;
  WS_LAUNCHED();
  return WS_OK;
}

static int dims_ok(ws_ctx *c, NetDims d) {
  if (d.Z < 4 || d.Y < 4 || d.X < 4 || (d.Z & 3) || (d.Y & 3) || (d.X & 3)) {
    c->err = "warm-start network: grid sides must be multiples of 4";
    return WS_EINVAL;
  }
  return WS_OK;
}

// loss of the last forward against tgt/mask; writes gout, returns the loss
static int net_loss(ws_ctx *c, NetDims d, double *loss) {
  Net *n = c->net;
  long V = vox(d);
  hipLaunchKernelGGL(k_loss_part, dim3(256), dim3(256), 0, c->stream, V, n->out, n->tgt, n->mask, n->sums);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, c->stream, 256, n->sums, n->sums + 1024);
  hipLaunchKernelGGL(k_loss_grad, dim3(blocks_for(V, 256)), dim3(256), 0, c->stream, V, n->out, n->tgt, n->mask,
                     n->sums + 1024, n->gout);
  WS_LAUNCHED();
  float wc = (float)c->netLossCoarse;
  if (wc > 0) {
    // the coarse error goes into the f32 gradient scratch gB (free before the backward pass)
    hipLaunchKernelGGL(k_loss4_part, dim3(256), dim3(256), 0, c->stream, d, n->out, n->tgt, n->mask, n->gB, n->sums);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, c->stream, 256, n->sums, n->sums + 1026);
    hipLaunchKernelGGL(k_loss4_grad, dim3(blocks_for(V, 256)), dim3(256), 0, c->stream, d, n->mask, n->gB,
                       n->sums + 1026, wc, n->gout);
    WS_LAUNCHED();
  }
  if (loss) {
    double h[4];
    WS_HIP(hipMemcpyAsync(h, n->sums + 1024, sizeof h, hipMemcpyDeviceToHost, c->stream));
    WS_HIP(hipStreamSynchronize(c->stream));
    *loss = h[1] > 1e-30 ? h[0] / h[1] : h[0] / 1e-30;
    if (wc > 0) *loss += wc * (h[3] > 1e-30 ? h[2] / h[3] : 0.0);
  }
  return WS_OK;
}

// ---- the solver side
static NetDims grid_dims(ws_ctx *c) { return NetDims{c->nz, c->ny, c->nx}; }
// the grid the network sees: the solver's, or with netCoarse the solver's
// coarsened 2x per axis
static NetDims net_grid(ws_ctx *c, int coarse) {
  return coarse ? NetDims{c->nz / 2, c->ny / 2, c->nx / 2} : grid_dims(c);
}

static int solver_inputs(ws_ctx *c) {
  Net *n = c->net;
  NetDims d = grid_dims(c);
  long C = c->nC;
  stage_launch(c, k_ws_residual_ops, k_ws_residual, 1024, 256, d, c->cellType, c->nbm, c->dg, c->rhs, c->q, n->r0,
               n->rpart);
  stage_launch(c, k_ws_scales_ops, k_ws_scales, 1, 256, 1024, n->rpart, n->scales);
  if (c->netCoarse) {
    NetDims cd = net_grid(c, 1);
    stage_launch(c, k_ws_inputs_coarse_ops, k_ws_inputs_coarse, blocks_for(vox(cd), 256), 256, cd, c->cellType, n->r0,
                 c->q, n->scales, n->x0);
  } else {
    stage_launch(c, k_ws_inputs_ops, k_ws_inputs, blocks_for(C, 256), 256, C, c->cellType, n->r0, c->q, n->scales,
                 n->x0);
  }
  WS_LAUNCHED();
  return WS_OK;
}

static int solver_side(ws_ctx *c) {
  int r;
  if (c->netCoarse && ((c->nx | c->ny | c->nz) & 1)) {
    c->err = "netCoarse: the grid sides must be even";
    return WS_EINVAL;
  }
  NetDims d = net_grid(c, c->netCoarse);
  if ((r = net_ensure(c)) || (r = dims_ok(c, d)) || (r = net_buffers(c, vox(d)))) return r;
  Net *n = c->net;
  if (!n->r0 && ((r = nalloc(c, n, &n->r0, c->nC)) || (r = nalloc(c, n, &n->qprev, c->nC)))) return r;
  return WS_OK;
}

int net_warm_start(ws_ctx *c) {
  int r;
  if ((r = solver_side(c)) || (r = solver_inputs(c)) || (r = net_forward(c, net_grid(c, c->netCoarse)))) return r;
  Net *n = c->net;
  if (c->netCoarse)
    hipLaunchKernelGGL(k_ws_apply_coarse, dim3(blocks_for(c->nC, 256)), dim3(256), 0, c->stream, net_grid(c, 1),
                       c->cellType, n->out, n->scales, c->q);
  else
    hipLaunchKernelGGL(k_ws_apply, dim3(blocks_for(c->nC, 256)), dim3(256), 0, c->stream, c->nC, c->cellType, n->out,
                       n->scales, c->q);
  WS_LAUNCHED();
  return WS_OK;
}

// the warm start of a solve whose forward pass and correction run for many
// planets at once (net_batch_forward): the network's inputs only
int net_warm_inputs(ws_ctx *c) {
  int r;
  if ((r = solver_side(c))) return r;
  if (!c->net->packed && (r = net_pack(c))) return r;
  return solver_inputs(c);
}

struct NetBatch {
  struct Step {
    int kind;  // -1 a convolution (form, nt: its product), else k_net_elem_batch's kind
    int form, nt, a, a2;
    int nj, off;  // planets, and their first entry in jobs / starts
    long blocks;
  };
  std::vector<Step> steps;
  NetJob *jobs = nullptr;
  long *starts = nullptr;
};

void net_batch_free(NetBatch *nb) {
  if (!nb) return;
  if (nb->jobs) (void)hipFree(nb->jobs);
  if (nb->starts) (void)hipFree(nb->starts);
  delete nb;
}

// the steps of the forward pass (the order of net_forward) and the
// correction, each over the planets with the warm start on
int net_batch_plan(ws_ctx *const *cs, int n, NetBatch **out, int *bad) {
  *out = nullptr;
  std::vector<ws_ctx *> P;
  for (int i = 0; i < n; i++) {
    if (!cs[i]->warmStart) continue;
    *bad = i;
    int r;
    if ((r = solver_side(cs[i]))) return r;
    if (!cs[i]->net->packed && (r = net_pack(cs[i]))) return r;
    P.push_back(cs[i]);
  }
  *bad = -1;
  if (P.empty()) return WS_OK;
  NetBatch *nb = new NetBatch();
  std::vector<NetJob> hj;
  std::vector<long> hs;
  auto add = [&](NetBatch::Step st, const std::vector<NetJob> &jobs, const std::vector<long> &blocks) {
    if (jobs.empty()) return;
    st.nj = (int)jobs.size();
    st.off = (int)hj.size();
    st.blocks = 0;
    for (size_t j = 0; j < jobs.size(); j++) {
      hj.push_back(jobs[j]);
      hs.push_back(st.blocks);
      st.blocks += blocks[j];
    }
    nb->steps.push_back(st);
  };
  // level 0, 1, 2: the network's grid, halved, quartered
  auto dims = [&](ws_ctx *c, int lev) {
    NetDims d = net_grid(c, c->netCoarse);
    for (int k = 0; k < lev; k++) d = half(d);
    return d;
  };
  auto layer = [&](int l, int lev, bf16_t *Net::*in, bf16_t *Net::*o) {
    const int nt = P[0]->net->L[l].cout / 16;
    for (int form = 0; form < 3; form++) {
      std::vector<NetJob> jobs;
      std::vector<long> blocks;
      for (ws_ctx *c : P) {
        const NetDims d = dims(c, lev);
        if (conv_form(c, d, nt) != form) continue;
        Net *m = c->net;
        const NetLayer &L = m->L[l];
        jobs.push_back(NetJob{d, m->*in, nullptr, L.wf, m->w + L.boff, m->*o, nullptr, nullptr});
        blocks.push_back(conv_blocks(form, d, nt));
      }
      add(NetBatch::Step{-1, form, nt, P[0]->net->L[l].cinp, P[0]->net->L[l].K, 0, 0, 0}, jobs, blocks);
    }
  };
  auto elem = [&](int kind, int lev, int a, int a2, auto job, auto count) {
    std::vector<NetJob> jobs;
    std::vector<long> blocks;
    for (ws_ctx *c : P) {
      jobs.push_back(job(c, dims(c, lev)));
      blocks.push_back(blocks_for(count(c, dims(c, lev)), 256));
    }
    add(NetBatch::Step{kind, 0, 0, a, a2, 0, 0, 0}, jobs, blocks);
  };
  auto pool = [&](int lev, int C, bf16_t *Net::*in, bf16_t *Net::*o) {
    elem(0, lev, C, 0, [&](ws_ctx *c, NetDims d) { return NetJob{d, c->net->*in, nullptr, nullptr, nullptr, c->net->*o, nullptr, nullptr}; },
         [&](ws_ctx *, NetDims d) { return vox(d) * C / 8; });
  };
  auto up = [&](int lev, int Cu, int Cs, bf16_t *Net::*in, bf16_t *Net::*skip, bf16_t *Net::*o) {
    elem(1, lev, Cu, Cs, [&](ws_ctx *c, NetDims d) { return NetJob{d, c->net->*in, c->net->*skip, nullptr, nullptr, c->net->*o, nullptr, nullptr}; },
         [&](ws_ctx *, NetDims d) { return vox(d) * (Cu + Cs) / 8; });
  };
  layer(0, 0, &Net::x0, &Net::a0);
  layer(1, 0, &Net::a0, &Net::s1);
  pool(1, 16, &Net::s1, &Net::p1);
  layer(2, 1, &Net::p1, &Net::a2);
  layer(3, 1, &Net::a2, &Net::s2);
  pool(2, 32, &Net::s2, &Net::p2);
  layer(4, 2, &Net::p2, &Net::a4);
  layer(5, 2, &Net::a4, &Net::a5);
  up(1, 64, 32, &Net::a5, &Net::s2, &Net::u1);
  layer(6, 1, &Net::u1, &Net::a6);
  layer(7, 1, &Net::a6, &Net::a7);
  up(0, 32, 16, &Net::a7, &Net::s1, &Net::u2);
  layer(8, 0, &Net::u2, &Net::a8);
  layer(9, 0, &Net::a8, &Net::a9);
  elem(2, 0, 0, 0,
       [&](ws_ctx *c, NetDims d) {
         Net *m = c->net;
         return NetJob{d, m->a9, nullptr, nullptr, m->w + m->L[10].woff, m->out, nullptr, nullptr};
       },
       [&](ws_ctx *, NetDims d) { return vox(d); });
  // the correction: from the coarsened grid or the solver's own
  for (int coarse = 1; coarse >= 0; coarse--) {
    std::vector<NetJob> jobs;
    std::vector<long> blocks;
    for (ws_ctx *c : P) {
      if ((c->netCoarse != 0) != (coarse != 0)) continue;
      Net *m = c->net;
      jobs.push_back(NetJob{net_grid(c, coarse), m->out, nullptr, nullptr, nullptr, c->q, c->cellType, m->scales});
      blocks.push_back(blocks_for(c->nC, 256));
    }
    add(NetBatch::Step{coarse ? 3 : 4, 0, 0, 0, 0, 0, 0, 0}, jobs, blocks);
  }
  hipError_t e = hipMalloc((void **)&nb->jobs, sizeof(NetJob) * hj.size());
  if (e == hipSuccess) e = hipMalloc((void **)&nb->starts, sizeof(long) * hs.size());
  if (e == hipSuccess) e = hipMemcpy(nb->jobs, hj.data(), sizeof(NetJob) * hj.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(nb->starts, hs.data(), sizeof(long) * hs.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    P[0]->err = std::string("warm-start network, batched pass: ") + hipGetErrorString(e);
    *bad = 0;
    for (int i = 0; i < n; i++)
      if (cs[i] == P[0]) *bad = i;
    net_batch_free(nb);
    return WS_EHIP;
  }
  *out = nb;
  return WS_OK;
}

void net_batch_forward(const NetBatch *nb, hipStream_t s) {
  if (!nb) return;
  for (const NetBatch::Step &st : nb->steps) {
    const NetJob *J = nb->jobs + st.off;
    const long *S = nb->starts + st.off;
    const dim3 grid(st.blocks), blk(256);
    if (st.kind >= 0) {
      hipLaunchKernelGGL(k_net_elem_batch, grid, blk, 0, s, J, S, st.nj, st.kind, st.a, st.a2);
      continue;
    }
#define NB_CONV(NT, F) \
  if (st.nt == NT && st.form == F) hipLaunchKernelGGL((k_conv_batch<NT, F>), grid, blk, 0, s, J, S, st.nj, st.a, st.a2);
    NB_CONV(1, 0) NB_CONV(1, 1) NB_CONV(1, 2) NB_CONV(2, 0) NB_CONV(2, 1) NB_CONV(2, 2) NB_CONV(4, 0) NB_CONV(4, 1)
    NB_CONV(4, 2)
#undef NB_CONV
  }
}

int net_record_inputs(ws_ctx *c) {
  int r;
  if ((r = solver_side(c))) return r;
  Net *n = c->net;
  if (n->bankCap < 1) { c->err = "recordSamples: no sample bank (ws_net_bank)"; return WS_ESTATE; }
  if (n->bankCoarse != c->netCoarse) {
    c->err = "recordSamples: the sample bank was made for the other netCoarse";
    return WS_ESTATE;
  }
  if ((r = solver_inputs(c))) return r;
  const long V = vox(net_grid(c, c->netCoarse));
  WS_HIP(hipMemcpyAsync(n->bankX + (long)n->bankNext * V * 8, n->x0, sizeof(bf16_t) * V * 8, hipMemcpyDeviceToDevice,
                        c->stream));
  WS_HIP(hipMemcpyAsync(n->qprev, c->q, sizeof(float) * c->nC, hipMemcpyDeviceToDevice, c->stream));
  return WS_OK;
}

int net_record_target(ws_ctx *c) {
  Net *n = c->net;
  long C = c->nC;
  if (n->bankCoarse) {
    NetDims cd = net_grid(c, 1);
    hipLaunchKernelGGL(k_ws_target_coarse, dim3(blocks_for(vox(cd), 256)), dim3(256), 0, c->stream, cd, c->cellType,
                       c->q, n->qprev, n->scales, n->bankT + (long)n->bankNext * vox(cd));
  } else {
    hipLaunchKernelGGL(k_ws_target, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, C, c->cellType, c->q, n->qprev,
                       n->scales, n->bankT + (long)n->bankNext * C);
  }
  WS_LAUNCHED();
  n->bankNext = (n->bankNext + 1) % n->bankCap;
  if (n->bankCount < n->bankCap) n->bankCount++;
  return WS_OK;
}

void net_free(ws_ctx *c) {
  Net *n = c->net;
  if (!n) return;
  for (void *p : n->owned)
    if (p) (void)hipFree(p);
  delete n;
  c->net = nullptr;
}

}  // namespace ws

// ======================================================================= ABI
using namespace ws;

extern "C" {

long ws_net_param_count(void) { return net_param_count(); }

int ws_net_set_params(ws_ctx *c, const float *w, long n) {
  if (!c || !w) return WS_EINVAL;
  int r = net_ensure(c);
  if (r) return r;
  if (n != c->net->nparam) { c->err = "ws_net_set_params: wrong parameter count"; return WS_EINVAL; }
  // on the context's stream, after the zeroing of a new network and any
  // queued training
  WS_HIP(hipMemcpyAsync(c->net->w, w, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  c->net->packed = false;
  return WS_OK;
}

int ws_net_get_params(ws_ctx *c, float *w, long n) {
  if (!c || !w) return WS_EINVAL;
  int r = net_ensure(c);
  if (r) return r;
  if (n != c->net->nparam) { c->err = "ws_net_get_params: wrong parameter count"; return WS_EINVAL; }
  WS_HIP(hipMemcpyAsync(w, c->net->w, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_net_forward(ws_ctx *c, int nz, int ny, int nx, const float *in5, float *out) {
  if (!c || !in5 || !out) return WS_EINVAL;
  NetDims d{nz, ny, nx};
  int r;
  if ((r = net_ensure(c)) || (r = dims_ok(c, d)) || (r = net_buffers(c, vox(d)))) return r;
  Net *n = c->net;
  long V = vox(d);
  float *tmp = nullptr;
  WS_HIP(hipMallocAsync((void **)&tmp, sizeof(float) * V * 5, c->stream));
  WS_HIP(hipMemcpyAsync(tmp, in5, sizeof(float) * V * 5, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(V * 8, 256)), dim3(256), 0, c->stream, V, tmp, n->x0);
  WS_LAUNCHED();
  if ((r = net_forward(c, d))) return r;
  WS_HIP(hipMemcpyAsync(out, n->out, sizeof(float) * V, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipFreeAsync(tmp, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_net_grad(ws_ctx *c, int nz, int ny, int nx, const float *in5, const float *target, const uint8_t *mask,
                float *grad, double *loss) {
  if (!c || !in5 || !target || !mask || !grad) return WS_EINVAL;
  NetDims d{nz, ny, nx};
  int r;
  if ((r = net_ensure(c)) || (r = dims_ok(c, d)) || (r = net_buffers(c, vox(d)))) return r;
  Net *n = c->net;
  long V = vox(d);
  float *tmp = nullptr;
  WS_HIP(hipMallocAsync((void **)&tmp, sizeof(float) * V * 5, c->stream));
  WS_HIP(hipMemcpyAsync(tmp, in5, sizeof(float) * V * 5, hipMemcpyHostToDevice, c->stream));
  WS_HIP(hipMemcpyAsync(n->tgt, target, sizeof(float) * V, hipMemcpyHostToDevice, c->stream));
  WS_HIP(hipMemcpyAsync(n->mask, mask, V, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(V * 8, 256)), dim3(256), 0, c->stream, V, tmp, n->x0);
  WS_LAUNCHED();
  if ((r = net_forward(c, d)) || (r = net_loss(c, d, loss)) || (r = net_backward(c, d, 0))) return r;
  WS_HIP(hipMemcpyAsync(grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipFreeAsync(tmp, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_net_bank(ws_ctx *c, int capacity) {
  if (!c || capacity < 1) return WS_EINVAL;
  int r;
  if ((r = solver_side(c))) return r;
  Net *n = c->net;
  release(n, n->bankX);
  release(n, n->bankT);
  n->bankX = nullptr;
  n->bankT = nullptr;
  n->bankCap = n->bankCount = n->bankNext = 0;
  const long V = vox(net_grid(c, c->netCoarse));
  if ((r = nalloc(c, n, &n->bankX, (size_t)capacity * V * 8)) || (r = nalloc(c, n, &n->bankT, (size_t)capacity * V)))
    return r;
  n->bankCap = capacity;
  n->bankCoarse = c->netCoarse;
  return WS_OK;
}

int ws_net_bank_count(ws_ctx *c, int *count) {
  if (!c || !count) return WS_EINVAL;
  *count = c->net ? c->net->bankCount : 0;
  return WS_OK;
}

int ws_net_train(ws_ctx *c, int steps, int batch, int crop, double lr, unsigned seed, double *loss) {
  if (!c || steps < 0 || batch < 1 || crop < 4 || (crop & 3)) return WS_EINVAL;
  Net *n = c->net;
  if (!n || n->bankCount < 1) { c->err = "ws_net_train: no recorded samples"; return WS_ESTATE; }
  NetDims cd{crop, crop, crop}, full = net_grid(c, n->bankCoarse);
  if (crop > full.X || crop > full.Y || crop > full.Z) {
    c->err = "ws_net_train: crop larger than the network's grid";
    return WS_EINVAL;
  }
  const long Vs = vox(full);  // one bank sample
  int r;
  if ((r = net_buffers(c, Vs > vox(cd) ? Vs : vox(cd)))) return r;
  long Vc = vox(cd);
  uint32_t s = seed * 2654435761u + 12345u;
  auto rnd = [&](int m) { s = s * 1664525u + 1013904223u; return (int)((s >> 8) % (uint32_t)m); };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double last = 0;
  for (int it = 0; it < steps; it++) {
    double lsum = 0;
    for (int bi = 0; bi < batch; bi++) {
      int smp = rnd(n->bankCount);
      int oz = 4 * rnd((full.Z - crop) / 4 + 1), oy = 4 * rnd((full.Y - crop) / 4 + 1);
      int ox = 4 * rnd((full.X - crop) / 4 + 1);
      hipLaunchKernelGGL(k_crop, dim3(blocks_for(Vc, 256)), dim3(256), 0, c->stream, full, cd, oz, oy, ox,
                         n->bankX + (long)smp * Vs * 8, n->bankT + (long)smp * Vs, n->x0, n->tgt, n->mask,
                         n->bankCoarse ? 0.0625f : 0.5f);
      WS_LAUNCHED();
      double l = 0;
      if ((r = net_forward(c, cd)) || (r = net_loss(c, cd, &l)) || (r = net_backward(c, cd, bi > 0))) return r;
      lsum += l;
    }
    n->adamStep++;
    float c1 = 1.f - std::pow(b1, (float)n->adamStep), c2 = 1.f - std::pow(b2, (float)n->adamStep);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n->nparam, 256)), dim3(256), 0, c->stream, n->nparam, n->w, n->g, n->m1,
                       n->m2, (float)lr, b1, b2, eps, c1, c2, 1.f / batch);
    WS_LAUNCHED();
    if ((r = net_pack(c))) return r;
    last = lsum / batch;
  }
  WS_HIP(hipStreamSynchronize(c->stream));
  if (loss) *loss = last;
  return WS_OK;
}

}  // extern "C"
