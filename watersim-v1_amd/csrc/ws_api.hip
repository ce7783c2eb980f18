// C ABI of the watersim HIP library (include/watersim.h): context setup,
// named field transfer, and the host-side orchestration of the substep
// stages on one HIP stream.
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <chrono>
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ws_frame.h"
#include "ws_learn.h"

using namespace ws;

namespace {

df to_df(double d) {
  float hi = (float)d;
  return df{hi, (float)(d - (double)hi)};
}

// thr[i] = smallest float x with floor(double(x)/dx) >= i (thr[0] = -inf):
// the exact cell index of a float coordinate as JavaScript computes it.
std::vector<float> floor_thresholds(double dx, int n) {
  std::vector<float> t(n + 2);
  t[0] = -FLT_MAX;
  for (int i = 1; i < n + 2; i++) {
    float hiF = (float)((i + 1) * dx);
    uint32_t lo = 0, hi;
    std::memcpy(&hi, &hiF, 4);
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      float x;
      std::memcpy(&x, &mid, 4);
      if (std::floor((double)x / dx) >= i) hi = mid; else lo = mid + 1;
    }
    std::memcpy(&t[i], &lo, 4);
  }
  return t;
}

__global__ void k_zero_words(unsigned *p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Device memory of n elements, zeroed on the context's stream, so that the
// zeros land before any later work of that stream. (A hipMemset would run on
// the null stream, which the context's non-blocking stream does not wait
// for: a kernel or copy queued right after could be overwritten by it.)
template <class T>
int dalloc(ws_ctx *c, T **p, size_t n) {
  const size_t bytes = sizeof(T) * (n ? n : 1), words = (bytes + 3) / 4;
  hipError_t e = hipMalloc((void **)p, words * 4);
  if (e != hipSuccess) {
    c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return WS_EHIP;
  }
  hipLaunchKernelGGL(k_zero_words, dim3(blocks_for((long)words, 256)), dim3(256), 0, c->stream, (unsigned *)*p,
                     (long)words);
  e = hipGetLastError();
  if (e != hipSuccess) {
    c->err = std::string("zeroing a new buffer: ") + hipGetErrorString(e);
    return WS_EHIP;
  }
  return WS_OK;
}

World world_of(ws_ctx *c) {
  World w;
  w.st = c->staticTypes;
  w.R = c->haveTerrain ? c->terrainR : nullptr;
  w.thr = c->thr;
  w.tn = c->tn;
  w.dv = (float)c->dv;
  w.inv_dv = c->dv > 0 ? (float)(1.0 / c->dv) : 0.f;
  w.dv_df = to_df(c->dv);
  w.cxd = to_df(c->cx); w.cyd = to_df(c->cy); w.czd = to_df(c->cz);
  w.oceanR = (float)c->oceanR;
  double skyAlt = 2 * c->dx > 0.3 ? 2 * c->dx : 0.3;
  w.skyAlt = (float)skyAlt;
  w.skyAlt_df = to_df(skyAlt);
  w.lim_df = to_df((c->oceanR + 0.3) * (c->oceanR + 0.3));
  w.p3_df = to_df(0.3);
  w.atmosphereH = (float)c->atmosphereH;
  w.domainR = (float)c->domainR;
  w.coreR = (float)c->coreR;
  return w;
}

Par make_par(ws_ctx *c) {
  Par P;
  P.pic = (float)c->pic; P.maxSpeed = (float)c->maxSpeed; P.gravity = (float)c->gravity;
  P.oceanR = (float)c->oceanR; P.atmosphereH = (float)c->atmosphereH; P.domainR = (float)c->domainR;
  P.coreR = (float)c->coreR; P.W = (float)c->W; P.H = (float)c->H; P.D = (float)c->D;
  double sx = 1, sy = 0, sz = 0;
  if (c->hasSun) {
    sx = c->sunX - c->cx; sy = c->sunY - c->cy; sz = c->sunZ - c->cz;
    double l = std::sqrt(sx * sx + sy * sy + sz * sz);
    if (!(l != 0)) l = 1;
    sx /= l; sy /= l; sz /= l;
  }
  P.sunX = (float)sx; P.sunY = (float)sy; P.sunZ = (float)sz;
  double rte = c->rainT < c->cloudT - 0.02 ? c->rainT : c->cloudT - 0.02;
  P.rainTe = (float)(rte < 0 ? 0 : rte);
  P.rainLiftAge = (float)c->rainLiftAge;
  P.ceilReflect = c->ceilReflect;
  return P;
}

int need_world(ws_ctx *c) {
  if (!c->haveStatic) {
    c->err = "no static world uploaded (staticTypes, gu, gv, gw)";
    return WS_ESTATE;
  }
  return WS_OK;
}

__global__ void k_set_dt(DevScalars *s, double dt) {
  if (threadIdx.x == 0) { s->dt = (float)dt; s->dtD = dt; }
}

__device__ __forceinline__ void k_reset_counters_body(long bx, long nbx, DevScalars *s) {
  if (threadIdx.x == 0) { s->umax_bits = 0; s->airborne = 0; }
}
__global__ void k_reset_counters(DevScalars *s) {
  k_reset_counters_body(blockIdx.x, gridDim.x, s);
}
WS_PLANETS_OPS(k_reset_counters_ops, k_reset_counters, k_reset_counters_body)

__global__ void k_zero_u32(unsigned *p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// The device scalars published into the pinned host mirror by a kernel: it
// copies them, makes its writes visible system-wide, then stores the call's
// sequence number into a pinned word that the host polls. This replaces the
// runtime's copy and blocking wait (a copy launch, then a wake-up of tens of
// microseconds) on the round trips of the to-residual solve and of every
// substep's end. A wait longer than 2 ms falls back to blocking on the
// stream, which also reports a failed launch.
__global__ void k_publish_scalars(const DevScalars *__restrict__ ds, DevScalars *hs, unsigned *seqWord,
                                  unsigned seq) {
  static_assert(sizeof(DevScalars) % 4 == 0, "whole words");
  constexpr int W = (int)(sizeof(DevScalars) / 4);
  const unsigned *src = (const unsigned *)ds;
  unsigned *dst = (unsigned *)hs;
  for (int i = threadIdx.x; i < W; i += blockDim.x) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(seqWord, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

unsigned next_publication(ws_ctx *c) {
  unsigned seq = ++c->hsPub;
  return seq ? seq : ++c->hsPub;
}

int wait_scalars(ws_ctx *c, unsigned seq);

int read_scalars(ws_ctx *c) {
  const unsigned seq = next_publication(c);
  hipLaunchKernelGGL(k_publish_scalars, dim3(1), dim3(256), 0, c->stream, c->ds, c->hs, c->hsSeq, seq);
  WS_HIP(hipGetLastError());
  return wait_scalars(c, seq);
}

// the wait for publication seq (queued on the context's stream)
int wait_scalars(ws_ctx *c, unsigned seq) {
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned spins = 0;; spins++) {
    if (__atomic_load_n(c->hsSeq, __ATOMIC_ACQUIRE) == seq) return WS_OK;
    if ((spins & 255u) == 255u && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
    __builtin_ia32_pause();
  }
  WS_HIP(hipStreamSynchronize(c->stream));
  if (__atomic_load_n(c->hsSeq, __ATOMIC_ACQUIRE) != seq) {
    c->err = "the device scalars were not published";
    return WS_EHIP;
  }
  return WS_OK;
}

// ------------------------------------------------------------------ stages
int st_rasterize(ws_ctx *c) {
  int nC = (int)c->nC;
  stage_launch(c, k_init_cells_ops, k_init_cells, blocks_for(nC + 1, 256), 256, c->staticTypes, c->cellType,
               c->cellCount, nC);
  int r = balls_raster(c);
  if (r) return r;
  if (c->nP > 0)
    stage_launch(c, k_mark_ops, k_mark, blocks_for(c->nP, 256), 256, c->geo, world_of(c), c->nP, c->px, c->py, c->pz,
                 c->pflag, c->cellType, c->cellCount, c->pcell, c->prank, c->ds);
  WS_LAUNCHED();
  if (int r = scan_excl(c, c->cellCount, c->cellStart, nC + 1)) return r;
  if (c->nP > 0)
    stage_launch(c, k_scatter_ops, k_scatter, blocks_for(c->nP, 256), 256, c->nP, c->pcell, c->prank, c->cellStart,
                 c->sortedIdx);
  WS_LAUNCHED();
  return WS_OK;
}

template <int B>
static void p2g_launch(ws_ctx *c, const StageOps &ops) {
  constexpr int TX = P2G_BINS[B][0], TY = P2G_BINS[B][1], TZ = P2G_BINS[B][2];
  int tx = (c->nx + TX - 1) / TX, ty = (c->ny + TY - 1) / TY, tz = (c->nz + TZ - 1) / TZ;
  stage_launch(c, ops, k_p2g_tile<TX, TY, TZ>, (long)tx * ty * tz, 256, c->geo, tx, ty, c->cellStart, c->sortedIdx,
               c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz, c->vel, c->wgt);
}

int st_p2g(ws_ctx *c) {
  static_assert(P2G_NBINS == 4, "one launch per bin shape");
  switch (c->p2gBin) {
    case 1: p2g_launch<1>(c, k_p2g_tile_ops1); break;
    case 2: p2g_launch<2>(c, k_p2g_tile_ops2); break;
    case 3: p2g_launch<3>(c, k_p2g_tile_ops3); break;
    default: p2g_launch<0>(c, k_p2g_tile_ops0); break;
  }
  WS_LAUNCHED();
  return WS_OK;
}

long nfaces(ws_ctx *c) { return c->uN + c->vN + c->wN; }

int st_bc(ws_ctx *c, float *snapshot, int fuse_gravity) {
  stage_launch(c, k_bc_snapshot_ops, k_bc_snapshot, blocks_for(nfaces(c), 256), 256, c->geo, c->cellType, c->vel,
               snapshot, c->wgt, c->valid, fuse_gravity, c->gu, c->ds, 1, c->nBalls ? c->sv : nullptr);
  WS_LAUNCHED();
  return WS_OK;
}

int visc_iterations(double coef) {
  // error after m iterations ~ 12 coef (3 coef)^m relative to |u|
  int m = 1;
  double e = 12 * coef * 3 * coef;
  while (e > 1e-8 && m < 15) { e *= 3 * coef; m++; }
  if (!(m & 1)) m++;
  return m;
}

// dt_bound: the substep length, or an upper bound of it when dt is decided on
// the device (the kernels read the actual dt; the bound only sets the
// iteration count, and more iterations stay exact)
int st_viscosity(ws_ctx *c, double dt_bound, const float *old) {
  double nu = c->viscosity;
  if (nu <= 1e-6) return WS_OK;
  double coef = nu * dt_bound / (c->dx * c->dx);
  if (coef > 0.24) coef = 0.24;
  int m = visc_iterations(coef);
  long F = nfaces(c);
  // x1 -> vel, x2 -> scr2, x3 -> vel, ... (m odd: the result lands in vel)
  const float *cur = old;
  for (int it = 0; it < m; it++) {
    float *out = (it & 1) ? c->scr2 : c->vel;
    stage_launch(c, k_visc_iter_ops, k_visc_iter, blocks_for(F, 256), 256, c->geo, c->cellType, old, cur, out,
                 (float)(nu / (c->dx * c->dx)), c->ds, it == 0 ? 1 : 0);
    cur = out;
  }
  WS_LAUNCHED();
  return WS_OK;
}

int st_gravity(ws_ctx *c) {
  stage_launch(c, k_gravity_bc_ops, k_gravity_bc, blocks_for(nfaces(c), 256), 256, c->geo, c->cellType, c->vel, c->gu,
               c->ds, c->nBalls ? c->sv : nullptr);
  WS_LAUNCHED();
  return WS_OK;
}

enum { SPAN_SOR = 0, SPAN_PRESSURE = 1, SPAN_NET = 2 };
// spans queued before they are read (a read synchronizes with the stream);
// their events are created when timing is switched on
constexpr size_t kSpanMax = 512;

int fold_spans(ws_ctx *c);

// the start event of a timed span, recorded on the context stream
int span_begin(ws_ctx *c, hipEvent_t *a) {
  if (c->evFree.empty()) {
    hipEvent_t e;
    WS_HIP(hipEventCreate(&e));
    c->evFree.push_back(e);
  }
  *a = c->evFree.back();
  c->evFree.pop_back();
  WS_HIP(hipEventRecord(*a, c->stream));
  return WS_OK;
}

// the end event; the pair is read later by fold_spans
int span_end(ws_ctx *c, hipEvent_t a, int kind, long launches) {
  if (c->evFree.empty()) {
    hipEvent_t e;
    WS_HIP(hipEventCreate(&e));
    c->evFree.push_back(e);
  }
  hipEvent_t b = c->evFree.back();
  c->evFree.pop_back();
  WS_HIP(hipEventRecord(b, c->stream));
  c->spans.push_back(ws_ctx::Span{a, b, kind, launches});
  return c->spans.size() >= kSpanMax ? fold_spans(c) : WS_OK;
}

// reads every queued span into the totals and recycles its events
int fold_spans(ws_ctx *c) {
  for (auto &sp : c->spans) {
    WS_HIP(hipEventSynchronize(sp.b));
    float ms = 0;
    WS_HIP(hipEventElapsedTime(&ms, sp.a, sp.b));
    if (sp.kind == SPAN_SOR) {
      c->sorMs += ms;
      c->sorLaunches += sp.launches;
    } else if (sp.kind == SPAN_PRESSURE) {
      c->pressureMs += ms;
      c->pressureSolves++;
    } else {
      c->netMs += ms;
      c->netCalls++;
    }
    c->evFree.push_back(sp.a);
    c->evFree.push_back(sp.b);
  }
  c->spans.clear();
  return WS_OK;
}

// sorMode 4 and 5 run the whole solve, stopping rule included, on the device
static inline bool dev_solve(const ws_ctx *c) { return c->sorMode == 4 || c->sorMode == 5; }

// the XCD-resident launch of n sweeps (sorMode 6); false when the lists
// do not fit any workgroup size it is built for (the caller then runs the
// list kernels)
bool sor_xcd_launch(ws_ctx *c, int n, int *err) {
  *err = WS_OK;
  int G = c->xcdGroups;
  if (G <= 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      *err = WS_EHIP;
      return false;
    }
    G = std::min(kXcdMaxG, std::max(1, cus / 8));
  }
  // a workgroup holds U x 1024 cells of each colour; ranges are cut from
  // the red list, and the black cells of a range differ from its red ones
  // by a few percent, so the choice keeps a margin (a range that does not
  // fit makes the launch decline)
  const long per = (std::max(c->nRed, c->nBlack) + G - 1) / G;
  const long want = per + per / 32 + 32;
  int U = c->xcdUnits ? c->xcdUnits : (want <= 1024 ? 1 : want <= 2048 ? 2 : want <= 3072 ? 3 : 0);
  if (!c->xcdUnits && U) U = std::max(U, c->xcdMinUnits);
  if (!U || U > 3) return false;
  if (!c->xcdX) {
    // exchange words are written by their launch before they are read (no
    // clearing); the arrival counters start at zero, below every epoch
    if (hipMalloc((void **)&c->xcdX, sizeof(unsigned long long) * (size_t)c->nC) != hipSuccess ||
        hipMalloc((void **)&c->xcdFlow, sizeof(unsigned) * kFlowStride * kXcdMaxG) != hipSuccess) {
      *err = WS_EHIP;
      return false;
    }
    c->xcdEpoch = 0;
  }
  if (c->xcdEpoch == 0 || c->xcdEpoch > 0xf0000000u - 2u * (unsigned)n) {
    hipLaunchKernelGGL(k_xcd_arm, dim3(1), dim3(256), 0, c->stream, kXcdMaxG, c->xcdFlow);
    c->xcdEpoch = 1;
  }
  XcdSweep J;
  J.g = c->geo;
  J.ds = c->ds;
  J.code = c->sorCode;
  J.lrhs = c->sorRhs;
  J.lpw = c->sorPw;
  J.off = c->sorOff;
  J.q = c->q;
  J.half = c->sorHalf;
  J.om1 = (float)(1 - c->omega);
  J.G = G;
  J.sweeps = n;
  J.epoch = c->xcdEpoch;
  J.flow = c->xcdFlow;
  J.xq = c->xcdX;
  c->xcdEpoch += 2u * (unsigned)n + 2u;
  const dim3 grid(8 * (G - 1) + 1);
  if (U == 1) hipLaunchKernelGGL((k_sor_xcd<1024, 1, 1024>), grid, dim3(1024), 0, c->stream, J);
  else if (U == 2) hipLaunchKernelGGL((k_sor_xcd<1024, 2, 2048>), grid, dim3(1024), 0, c->stream, J);
  else hipLaunchKernelGGL((k_sor_xcd<1024, 3, 3072>), grid, dim3(1024), 0, c->stream, J);
  if (hipGetLastError() != hipSuccess) {
    *err = WS_EHIP;
    return false;
  }
  c->xcdPending = n;
  c->xcdPendingUnits = U;
  return true;
}

int sor_sweeps(ws_ctx *c, int n) {
  c->tbLaunchesPending = 0;
  if (c->sorMode == 6 && n > 0) {
    hipEvent_t span = nullptr;
    int r;
    if (c->timing && (r = span_begin(c, &span))) return r;
    if (sor_xcd_launch(c, n, &r)) {
      c->sweepsTotal += n;
      c->sorCellSweeps += (c->nRed + c->nBlack) * (long)n;
      if (span && (r = span_end(c, span, SPAN_SOR, 1))) return r;
      return WS_OK;
    }
    if (r) return r;
    // lists too long for the XCD kernel: this context sweeps with the list
    // kernel from now on
    c->sorMode = 2;
    c->xcdFallbacks++;
    if (span && (r = span_end(c, span, SPAN_SOR, 0))) return r;
  }
  int half = (c->nx + 1) / 2;
  long nthr = (long)half * c->ny * c->nz;
  float om1 = (float)(1 - c->omega);
  hipEvent_t span = nullptr;
  int r;
  if (c->timing && n > 0 && (r = span_begin(c, &span))) return r;
  int nT = c->sorTiles[0] * c->sorTiles[1] * c->sorTiles[2];
  for (int it = 0; it < n; it++) {
    if (c->sorMode == 3) {
      int nTB = c->tbTiles[0] * c->tbTiles[1] * c->tbTiles[2];
      if (it + 1 < n) {
        hipLaunchKernelGGL(k_sor_tb<2>, dim3(nTB), dim3(SOR_TB_THREADS), 0, c->stream, c->geo, c->tbTiles[0],
                           c->tbTiles[1], c->tbActive, c->q, c->q2, c->tbPk, c->tbRhs, c->tbCnt, (float)c->omega,
                           om1);
        it++;
      } else {
        hipLaunchKernelGGL(k_sor_tb<1>, dim3(nTB), dim3(SOR_TB_THREADS), 0, c->stream, c->geo, c->tbTiles[0],
                           c->tbTiles[1], c->tbActive, c->q, c->q2, c->tbPk, c->tbRhs, c->tbCnt, (float)c->omega,
                           om1);
      }
      float *t = c->q; c->q = c->q2; c->q2 = t;
      c->tbLaunchesPending++;
    } else if (c->sorMode == 2) {
      if (c->nRed > 0)
        hipLaunchKernelGGL(k_sor_list, dim3(blocks_for(c->nRed, 256)), dim3(256), 0, c->stream, c->geo, (int)c->nRed,
                           c->sorCode, c->sorRhs, c->sorPw, c->q, om1);
      if (c->nBlack > 0)
        hipLaunchKernelGGL(k_sor_list, dim3(blocks_for(c->nBlack, 256)), dim3(256), 0, c->stream, c->geo,
                           (int)c->nBlack, c->sorCode + c->sorHalf, c->sorRhs + c->sorHalf, c->sorPw + c->sorHalf,
                           c->q, om1);
    } else if (c->sorMode == 1) {
      hipLaunchKernelGGL(k_sor_tile, dim3(nT), dim3(256), 0, c->stream, c->geo, c->sorTiles[0], c->sorTiles[1],
                         c->sorActive, c->q, c->q2, c->rhs, c->pw, c->nbm, om1);
      float *t = c->q; c->q = c->q2; c->q2 = t;  // the new sweep is current
    } else {
      for (int color = 0; color < 2; color++)
        hipLaunchKernelGGL(k_sor, dim3(blocks_for(nthr, 256)), dim3(256), 0, c->stream, c->geo, c->q, c->rhs, c->pw,
                           c->nbm, color, om1);
    }
  }
  WS_LAUNCHED();
  c->sweepsTotal += n;
  c->sorCellSweeps += (c->sorMode >= 2 ? c->nRed + c->nBlack : c->nC) * (long)n;
  if (span) {
    // the span holds only back-to-back SOR launches: its length over the
    // launch count is the kernel's average duration (gaps included)
    long launches;
    if (c->sorMode == 3) {
      launches = c->tbLaunchesPending;
    } else {
      long per = c->sorMode == 1 ? 1L : (c->sorMode == 2 ? (long)(c->nRed > 0) + (c->nBlack > 0) : 2L);
      launches = per * n;
    }
    if ((r = span_end(c, span, SPAN_SOR, launches))) return r;
  }
  return WS_OK;
}

int fold_dev_stats(ws_ctx *c);
void take_counts(ws_ctx *c);
int xcd_confirm(ws_ctx *c, int *redo);
int device_failures(ws_ctx *c);
int xcd_settle(ws_ctx *c);
int sor_sweeps_list(ws_ctx *c, int n);

// the on-device solve of one context (sorMode 4)
SorJob sor_job(ws_ctx *c) {
  SorJob J;
  J.g = c->geo;
  J.ds = c->ds;
  J.code = c->sorCode;
  J.lrhs = c->sorRhs;
  J.lpw = c->sorPw;
  J.ldg = c->sorDg;
  J.q = c->q;
  J.half = c->sorHalf;
  J.om1 = (float)(1 - c->omega);
  J.every = c->residualEvery > 0 ? c->residualEvery : 1;
  J.maxIters = c->maxPressureIters;
  J.adaptive = c->adaptiveChecks;
  J.fixedIters = c->pressureIters;
  J.measure = c->measureResidual;
  J.tol = c->pressureTol;
  J.chunk = c->sorChunk;
  J.arrive = nullptr;
  J.G = 1;
  return J;
}

int sor_solve_dev(ws_ctx *c) {
  SorJob J = sor_job(c);
  if (c->timing) WS_HIP(hipEventRecord(c->ev[0], c->stream));
  if (c->sorMode == 5) {
    // neighbour-synchronised solve: flowGroups workgroups on the first
    // flowXcds XCDs, all resident at once (at most one per compute unit)
    const int G = c->flowGroups, X = c->flowXcds;
    J.G = G;
    J.arrive = c->sorArrive;
    J.off = c->sorOff;
    J.flow = c->sorFlow;
    hipLaunchKernelGGL(k_flow_arm, dim3(1), dim3(256), 0, c->stream, G, c->sorFlow, c->sorArrive);
    const int grid = (G / X) * 8 + G % X;
    if (c->sorThreads == 256)
      hipLaunchKernelGGL(k_sor_flow<256>, dim3(grid), dim3(256), 0, c->stream, J, X);
    else if (c->sorThreads == 512)
      hipLaunchKernelGGL(k_sor_flow<512>, dim3(grid), dim3(512), 0, c->stream, J, X);
    else
      hipLaunchKernelGGL(k_sor_flow<1024>, dim3(grid), dim3(1024), 0, c->stream, J, X);
  } else if (c->sorGroups > 1) {
    // several workgroups meeting at a barrier: all must be resident, so at
    // most one per compute unit (a barrier that times out fails the solve)
    J.G = c->sorGroups;
    J.arrive = c->sorArrive;
    hipLaunchKernelGGL(k_sor_arm, dim3(1), dim3(64), 0, c->stream, 1, c->sorArrive);
    if (c->sorThreads == 256)
      hipLaunchKernelGGL(k_sor_solve_multi<256>, dim3(J.G), dim3(256), 0, c->stream, J);
    else if (c->sorThreads == 512)
      hipLaunchKernelGGL(k_sor_solve_multi<512>, dim3(J.G), dim3(512), 0, c->stream, J);
    else
      hipLaunchKernelGGL(k_sor_solve_multi<1024>, dim3(J.G), dim3(1024), 0, c->stream, J);
  } else if (c->sorThreads == 256)
    hipLaunchKernelGGL(k_sor_solve<256>, dim3(1), dim3(256), 0, c->stream, J);
  else if (c->sorThreads == 512)
    hipLaunchKernelGGL(k_sor_solve<512>, dim3(1), dim3(512), 0, c->stream, J);
  else
    hipLaunchKernelGGL(k_sor_solve<1024>, dim3(1), dim3(1024), 0, c->stream, J);
  WS_LAUNCHED();
  c->devStatsPending = true;
  if (c->timing) {
    WS_HIP(hipEventRecord(c->ev[1], c->stream));
    WS_HIP(hipEventSynchronize(c->ev[1]));
    float ms = 0;
    WS_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->sorMs += ms;
    c->sorLaunches += 1;
    int r = fold_dev_stats(c);
    if (r) return r;
    c->sorCellSweeps += (c->nRed + c->nBlack) * (long)c->itersDone;
  }
  return WS_OK;
}

// the sweep and residual counts of on-device solves, folded into the host
// counters (and the last solve's sweeps and residual)
int fold_dev_stats(ws_ctx *c) {
  if (!c->devStatsPending) return WS_OK;
  if (int r = read_scalars(c)) return r;
  c->devStatsPending = false;
  if (c->hs->sorRes < 0) {
    c->err = "on-device pressure solve: its workgroups' barrier timed out (not all resident)";
    return WS_ESTATE;
  }
  c->itersDone = (int)c->hs->sorIters;
  c->residual = c->hs->sorRes;
  c->sweepsTotal += (long)(c->hs->sorSweeps - c->sorSweepsSeen);
  c->residualsTotal += (long)(c->hs->sorChecks - c->sorChecksSeen);
  c->sorSweepsSeen = c->hs->sorSweeps;
  c->sorChecksSeen = c->hs->sorChecks;
  take_counts(c);
  return WS_OK;
}

// the list lengths, as the last device-scalar copy saw them
void take_counts(ws_ctx *c) {
  c->nRed = c->hs->nRed;
  c->nBlack = c->hs->nBlack;
}

int residual(ws_ctx *c, double *res) {
  c->residualsTotal++;
  if (c->sorMode >= 2)
    hipLaunchKernelGGL(k_residual_list, dim3(c->nPartials), dim3(256), 0, c->stream, c->geo, c->ds, c->sorHalf,
                       c->sorCode, c->sorRhs, c->sorDg, c->q, c->partials);
  else
    hipLaunchKernelGGL(k_residual, dim3(c->nPartials), dim3(256), 0, c->stream, c->geo, c->q, c->rhs, c->nbm, c->dg,
                       c->partials);
  // the final sum publishes the scalars to the host mirror itself
  const unsigned seq = next_publication(c);
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(256), 0, c->stream, c->partials, c->nPartials, c->ds, c->hs,
                     c->hsSeq, seq);
  WS_LAUNCHED();
  if (int r = wait_scalars(c, seq)) return r;
  take_counts(c);
  if (int e = device_failures(c)) return e;
  int redo = 0, r = xcd_confirm(c, &redo);
  if (r) return r;
  if (redo) {
    // the XCD launch declined and q is as it was: the list kernel runs
    // the same sweeps, then the residual is taken again
    c->residualsTotal--;
    if ((r = sor_sweeps_list(c, redo))) return r;
    return residual(c, res);
  }
  *res = c->hs->bb > 0 ? std::sqrt(c->hs->rr) / std::sqrt(c->hs->bb) : 0.0;
  return WS_OK;
}

// After a read of the device scalars: did the last XCD-resident launch run?
// A declined launch (q untouched) switches the context to the list kernel
// and returns its sweeps in *redo; a wait that gave up fails the solve.
int xcd_confirm(ws_ctx *c, int *redo) {
  *redo = 0;
  if (!c->xcdPending) return WS_OK;
  const int n = c->xcdPending;
  c->xcdPending = 0;
  if (c->hs->xcdTimedOut) {
    c->err = "XCD-resident pressure sweeps: a workgroup's wait for its neighbours timed out";
    return WS_ESTATE;
  }
  if (c->hs->xcdDeclined != c->xcdDeclinedSeen) {
    const bool big = c->hs->xcdTooBig != c->xcdTooBigSeen;
    c->xcdDeclinedSeen = c->hs->xcdDeclined;
    c->xcdTooBigSeen = c->hs->xcdTooBig;
    if (big && !c->xcdUnits && c->xcdPendingUnits < 3)
      c->xcdMinUnits = c->xcdPendingUnits + 1;  // later launches hold more cells per thread
    else
      c->sorMode = 2;  // placement, or no larger size: the list kernel from now on
    c->xcdFallbacks++;
    c->sweepsTotal -= n;
    c->sorCellSweeps -= (c->nRed + c->nBlack) * (long)n;
    *redo = n;
  }
  return WS_OK;
}

// the device scalars read now, to confirm a pending XCD-resident launch
// (paths that sweep without a residual check afterwards)
int xcd_settle(ws_ctx *c) {
  if (!c->xcdPending) return WS_OK;
  if (int r = read_scalars(c)) return r;
  int redo = 0, r = xcd_confirm(c, &redo);
  if (r) return r;
  return redo ? sor_sweeps_list(c, redo) : WS_OK;
}

// n sweeps with the list kernel whatever the mode (a declined XCD launch)
int sor_sweeps_list(ws_ctx *c, int n) {
  const int keep = c->sorMode;
  c->sorMode = 2;
  const int r = sor_sweeps(c, n);
  c->sorMode = keep;
  return r;
}

int st_pressure_solve(ws_ctx *c);

// Measurement of the learned warm start: both solves on the same assembled
// system, plain SOR from the previous pressure and then, from the same
// previous pressure, the network's prediction followed by SOR to the same
// residual. Times (HIP events on the solver stream) and sweep counts are
// accumulated per method; the warm-started result is kept.
int st_pressure_compare(ws_ctx *c) {
  int r, keep = c->warmStart;
  WS_HIP(hipMemcpyAsync(c->scr2, c->q, sizeof(float) * c->nC, hipMemcpyDeviceToDevice, c->stream));
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) WS_HIP(hipMemcpyAsync(c->q, c->scr2, sizeof(float) * c->nC, hipMemcpyDeviceToDevice, c->stream));
    c->warmStart = pass;
    WS_HIP(hipEventRecord(c->ev[2], c->stream));
    r = st_pressure_solve(c);
    if (r) { c->warmStart = keep; return r; }
    WS_HIP(hipEventRecord(c->ev[3], c->stream));
    WS_HIP(hipEventSynchronize(c->ev[3]));
    float ms = 0;
    WS_HIP(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    if (pass == 0) { c->cmpMsPlain += ms; c->cmpSweepsPlain += c->itersDone; }
    else { c->cmpMsLearned += ms; c->cmpSweepsLearned += c->itersDone; }
  }
  c->cmpSolves++;
  c->warmStart = keep;
  return WS_OK;
}

int st_pressure(ws_ctx *c) {
  if (c->pressureCompare && c->timing) return st_pressure_compare(c);
  hipEvent_t span = nullptr;
  int r;
  if (c->timing && (r = span_begin(c, &span))) return r;
  if ((r = st_pressure_solve(c))) return r;
  if (span && (r = span_end(c, span, SPAN_PRESSURE, 0))) return r;
  return WS_OK;
}

int pressure_setup(ws_ctx *c);
int pressure_iterate(ws_ctx *c);

int st_pressure_solve(ws_ctx *c) {
  int r = pressure_setup(c);
  return r ? r : pressure_iterate(c);
}

// the system, its fluid lists and the starting guess (warm start)
int pressure_setup(ws_ctx *c) {
  stage_launch(c, k_assemble_ops, k_assemble, blocks_for(c->nC, 256), 256, c->geo, c->cellType, c->vel, c->q, c->q2,
               c->rhs, c->pw, c->nbm, c->dg, (float)c->omega);
  if (c->sorMode == 1) {
    int nT = c->sorTiles[0] * c->sorTiles[1] * c->sorTiles[2];
    hipLaunchKernelGGL(k_sor_tiles_mark, dim3(nT), dim3(256), 0, c->stream, c->geo, c->sorTiles[0], c->sorTiles[1],
                       nT, c->nbm, c->sorActive);
  }
  if (c->sorMode == 3) {
    int nTB = c->tbTiles[0] * c->tbTiles[1] * c->tbTiles[2];
    if (!c->tbPk) {
      int r;
      if ((r = dalloc(c, &c->tbPk, (size_t)nTB * TB_CAP)) || (r = dalloc(c, &c->tbRhs, (size_t)nTB * TB_CAP)) ||
          (r = dalloc(c, &c->tbCnt, (size_t)nTB * 4)))
        return r;
    }
    hipLaunchKernelGGL(k_sor_tb_mark, dim3(nTB), dim3(256), 0, c->stream, c->geo, c->tbTiles[0], c->tbTiles[1], nTB,
                       c->nbm, c->tbActive);
    hipLaunchKernelGGL(k_sor_tb_build, dim3(nTB), dim3(SOR_TB_THREADS), 0, c->stream, c->geo, c->tbTiles[0],
                       c->tbTiles[1], c->tbActive, c->nbm, c->dg, c->rhs, c->tbPk, c->tbRhs, c->tbCnt);
  }
  WS_LAUNCHED();
  int r;
  if (c->sorMode >= 2) {
    int nb = blocks_for(c->nC + 1, 256);
    stage_launch(c, k_sor_count_ops, k_sor_count, nb, 256, c->geo, c->nbm, c->sorOff);
    WS_LAUNCHED();
    if (int r = scan_excl(c, c->sorOff, c->sorOff, (long)c->nC + 1)) return r;
    stage_launch(c, k_sor_compact_ops, k_sor_compact, nb, 256, c->geo, c->nbm, c->dg, c->rhs, c->pw, c->sorOff,
                 c->sorHalf, c->sorCode, c->sorRhs, c->sorPw, c->sorDg, c->ds);
    WS_LAUNCHED();
    if (c->pressureTol <= 0 && !dev_solve(c)) {  // the to-residual path learns the counts with its first check
      if ((r = read_scalars(c))) return r;
      take_counts(c);
    }
  }
  // the learned warm start: record this solve's inputs for training and/or
  // start the sweeps from the network's prediction
  if (c->recordSamples && (r = net_record_inputs(c))) return r;
  if (c->warmStart) {
    hipEvent_t span = nullptr;
    if (c->timing && (r = span_begin(c, &span))) return r;
    if ((r = c->netDeferred ? net_warm_inputs(c) : net_warm_start(c))) return r;
    if (span && (r = span_end(c, span, SPAN_NET, 0))) return r;
  }
  return WS_OK;
}

// the sweeps, to the tolerance or the fixed count
int pressure_iterate(ws_ctx *c) {
  int r;
  if (dev_solve(c)) {
    if ((r = sor_solve_dev(c))) return r;
    if (c->recordSamples && (r = net_record_target(c))) return r;
    return WS_OK;
  }
  if (c->pressureTol <= 0) {
    if ((r = sor_sweeps(c, c->pressureIters))) return r;
    c->itersDone = c->pressureIters;
    if (c->measureResidual) {
      if ((r = residual(c, &c->residual))) return r;
    } else if ((r = xcd_settle(c))) {
      return r;
    }
    if (c->recordSamples && (r = net_record_target(c))) return r;
    return WS_OK;
  }
  // to a relative residual: sweep in batches, checking in between; stop
  // early if the residual no longer falls (float resolution reached)
  // The batch length follows the observed convergence rate: after each check
  // the sweeps still needed at that rate are predicted, so only a few
  // residual evaluations run per solve. With adaptive checks the first batch
  // is three quarters of the previous solve's sweeps (consecutive solves of
  // a context need similar counts).
  int done = 0, flat = 0, k = c->residualEvery;
  if (c->adaptiveChecks && c->itersDone > 0) k = std::max(k, (int)(0.75 * c->itersDone));
  double res = 0, best, prev;
  if ((r = residual(c, &res))) return r;
  best = prev = res;
  while (res > c->pressureTol && done < c->maxPressureIters && flat < 12) {
    if (done + k > c->maxPressureIters) k = c->maxPressureIters - done;
    if ((r = sor_sweeps(c, k))) return r;
    done += k;
    if ((r = residual(c, &res))) return r;
    if (res < 0.995 * best) { best = res; flat = 0; } else flat++;
    int kn = c->residualEvery;
    if (c->adaptiveChecks && res < prev && res > c->pressureTol) {
      double rate = std::pow(res / prev, 1.0 / k);  // per sweep
      double need = std::log(c->pressureTol / res) / std::log(rate);
      kn = need < 4 ? 4 : (need > 64 ? 64 : (int)std::ceil(need));
    }
    prev = res;
    k = kn;
  }
  c->itersDone = done;
  c->residual = res;
  if (c->recordSamples && (r = net_record_target(c))) return r;
  return WS_OK;
}

// face-sweep grid with the umax reduction: a bounded grid-stride launch
int reduce_grid(long F) {
  int b = blocks_for(F, 256);
  return b < 2048 ? b : 2048;
}

int st_project(ws_ctx *c, int do_clamp) {
  stage_launch(c, k_project_ops, k_project, reduce_grid(nfaces(c)), 256, c->geo, c->cellType, c->q, c->vel, do_clamp,
               (float)(c->maxSpeed * 1.1), c->ds);
  WS_LAUNCHED();
  return WS_OK;
}

int st_vorticity(ws_ctx *c) {
  if (!(c->vorticity > 0)) return WS_OK;
  int nC = (int)c->nC;
  float *vc = c->vort, *om = c->vort + 3 * c->nC, *fc = c->vort + 7 * c->nC;
  stage_launch(c, k_vort_center_ops, k_vort_center, blocks_for(nC, 256), 256, c->geo, c->vel, vc);
  stage_launch(c, k_vort_omega_ops, k_vort_omega, blocks_for(nC, 256), 256, c->geo, vc, om, (float)(0.5 / c->dx));
  stage_launch(c, k_vort_force_ops, k_vort_force, blocks_for(nC, 256), 256, c->geo, c->cellType, om, fc,
               (float)c->vorticity, c->ds);
  stage_launch(c, k_vort_splat_ops, k_vort_splat, blocks_for(nfaces(c), 256), 256, c->geo, c->cellType, fc, c->vel);
  WS_LAUNCHED();
  return WS_OK;
}

int st_clamp(ws_ctx *c) {
  stage_launch(c, k_clamp_ops, k_clamp, reduce_grid(nfaces(c)), 256, nfaces(c), c->vel, (float)(c->maxSpeed * 1.1),
               c->ds);
  WS_LAUNCHED();
  return WS_OK;
}

// the row-pipelined extrapolation: flags (two launches), values (one),
// the check that every job ran
int extrapolate_rows(ws_ctx *c) {
  const int D1 = c->ny + 1, D2 = c->nz + 1, W = (c->nx + 1 + 63) / 64;
  if (!c->exFin) {
    c->exS = (D2 + EXR_TK - 1) / EXR_TK;
    c->exC = W;
    const size_t nxk = (size_t)3 * c->exS * c->exC * D1 * 64, nxi = (size_t)3 * c->exS * c->exC * EXR_TK * D1;
    WS_HIP(hipMalloc((void **)&c->exFin, sizeof(unsigned long long) * (size_t)3 * D2 * D1 * W));
    WS_HIP(hipMalloc((void **)&c->exXk, sizeof(unsigned long long) * nxk));
    WS_HIP(hipMalloc((void **)&c->exXi, sizeof(unsigned long long) * nxi));
    WS_HIP(hipMalloc((void **)&c->exCnt, sizeof(unsigned) * 8));
    // tags 0: below every launch epoch (the clears run on the context's
    // stream, ahead of the first launch)
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for((long)nxk * 2, 256)), dim3(256), 0, c->stream,
                       (unsigned *)c->exXk, (long)nxk * 2);
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for((long)nxi * 2, 256)), dim3(256), 0, c->stream,
                       (unsigned *)c->exXi, (long)nxi * 2);
    c->exEpoch = 0;
  }
  if (++c->exEpoch == 0) {
    const size_t nxk = (size_t)3 * c->exS * c->exC * D1 * 64, nxi = (size_t)3 * c->exS * c->exC * EXR_TK * D1;
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for((long)nxk * 2, 256)), dim3(256), 0, c->stream,
                       (unsigned *)c->exXk, (long)nxk * 2);
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for((long)nxi * 2, 256)), dim3(256), 0, c->stream,
                       (unsigned *)c->exXi, (long)nxi * 2);
    c->exEpoch = 1;
  }
  ExRun R;
  R.ticket = c->exCnt;
  R.done = c->exCnt + 4;
  R.xk = c->exXk;
  R.xi = c->exXi;
  R.S = c->exS;
  R.C = c->exC;
  R.epoch = c->exEpoch;
  hipLaunchKernelGGL(k_exf_rows, dim3(3 * D2), dim3(256), 0, c->stream, c->geo, c->valid, c->exFin, R);
  hipLaunchKernelGGL(k_exf_k, dim3((3 * D1 * W + 3) / 4), dim3(256), 0, c->stream, c->geo, c->exFin);
  // the most jobs of any component, on each of the 8 XCDs (components run on
  // XCDs 0, 1, 2; the other blocks return at once)
  const int jobs = c->exS * c->exC;
  hipLaunchKernelGGL(k_extrap_rows, dim3(8 * jobs), dim3(EXR_THREADS), 0, c->stream, c->geo, c->vel, c->valid,
                     c->exFin, R, c->ds);
  hipLaunchKernelGGL(k_extrap_check, dim3(1), dim3(64), 0, c->stream, c->geo, R, c->ds);
  WS_LAUNCHED();
  return WS_OK;
}

// failures the device reports in its scalars (read after each copy of them)
int device_failures(ws_ctx *c) {
  if (c->hs->scanFailed) {
    c->err = "a prefix sum's wait for an earlier tile timed out";
    return WS_ESTATE;
  }
  if (c->hs->exFailed) {
    c->err = c->hs->exFailed & 2 ? "extrapolation: a component's jobs were not all run (no workgroup on its XCD)"
                                 : "extrapolation: a wait for a neighbouring slab or chunk timed out";
    return WS_ESTATE;
  }
  return WS_OK;
}

int st_extrapolate(ws_ctx *c) {
  if (c->exRows) return extrapolate_rows(c);
  for (size_t d = 0; d + 1 < c->exDiagOff.size(); d++) {
    int n = c->exDiagOff[d + 1] - c->exDiagOff[d];
    if (n > 0)
      hipLaunchKernelGGL(k_extrap_diag, dim3(n), dim3(EX_THREADS), 0, c->stream, c->geo, c->exDiag[d], c->vel,
                         c->valid);
  }
  WS_LAUNCHED();
  return WS_OK;
}

int st_vapor(ws_ctx *c) {
  if (!c->hasVapor || c->nP <= 0) return WS_OK;
  int nb = blocks_for(c->nP, 256);
  World w = world_of(c);
  Par P = make_par(c);
  unsigned *needA = c->vneed, *offA = c->vneed + c->capacity, *needB = c->vneed + 2L * c->capacity,
           *offB = c->vneed + 3L * c->capacity;
  hipLaunchKernelGGL(k_vapor_a, dim3(nb), dim3(256), 0, c->stream, c->geo, w, P, c->nP, c->px, c->py, c->pz, c->pvx,
                     c->pvy, c->pvz, c->pflag, c->pAir, c->vscr, needA, c->ds);
  WS_LAUNCHED();
  if (int r = scan_excl(c, needA, offA, c->nP)) return r;
  hipLaunchKernelGGL(k_vapor_b, dim3(nb), dim3(256), 0, c->stream, c->geo, P, c->nP, c->px, c->py, c->pz, c->pvx,
                     c->pvy, c->pvz, c->pflag, c->vscr, needA, offA, needB, c->ds, c->rng);
  hipLaunchKernelGGL(k_rng_advance, dim3(1), dim3(64), 0, c->stream, c->rng, 0, needA, offA, c->nP);
  WS_LAUNCHED();
  if (int r = scan_excl(c, needB, offB, c->nP)) return r;
  hipLaunchKernelGGL(k_vapor_c, dim3(nb), dim3(256), 0, c->stream, c->geo, w, P, c->nP, c->px, c->py, c->pz, c->pvx,
                     c->pvy, c->pvz, c->pflag, c->pAir, c->pT, c->cellType, c->vscr, needB, offB, c->ds, c->rng);
  hipLaunchKernelGGL(k_rng_advance, dim3(1), dim3(64), 0, c->stream, c->rng, 1, needB, offB, c->nP);
  WS_LAUNCHED();
  return WS_OK;
}

int st_particles(ws_ctx *c, int mode) {
  if (c->nP <= 0) return WS_OK;
  stage_launch(c, k_g2p_advect_ops, k_g2p_advect, blocks_for(c->nP, 256), 256, c->geo, world_of(c), make_par(c), c->nP,
               mode, c->vel, c->velO, c->cellType, c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pflag, c->pcool,
               c->pAir, c->advU, c->advV, c->advW, c->ds);
  WS_LAUNCHED();
  if (mode & 2) {
    int r = st_vapor(c);
    if (r) return r;
    // ball collisions after advection (solver.js:1762-1763)
    return balls_push(c, false);
  }
  return WS_OK;
}

int substep_pre(ws_ctx *c, double dt, int use_law);
int substep_post(ws_ctx *c);

// the full substep; dt either from the host or from the CFL law on the device
int substep(ws_ctx *c, double dt, int use_law) {
  int r;
  if ((r = substep_pre(c, dt, use_law)) || (r = st_pressure(c)) || (r = substep_post(c))) return r;
  return WS_OK;
}

// the substep up to the pressure solve: dt, rasterization, P2G, boundary
// conditions, gravity and viscosity
int substep_pre(ws_ctx *c, double dt, int use_law) {
  int r;
  if ((r = need_world(c))) return r;
  double base = c->dtFrameBase > 0 ? c->dtFrameBase : 0.01;
  stage_launch(c, k_begin_substep_ops, k_begin_substep, 1, 64, c->ds, dt, use_law, c->cfl * c->dx, base,
               (float)c->gravity);
  stage_launch(c, k_reset_counters_ops, k_reset_counters, 1, 64, c->ds);
  WS_LAUNCHED();
  bool visc = c->viscosity > 1e-6;
  if ((r = st_rasterize(c))) return r;
  if ((r = balls_push(c, true))) return r;
  if ((r = st_p2g(c))) return r;
  if ((r = st_bc(c, c->velO, visc ? 0 : 1))) return r;
  if (visc) {
    if ((r = st_viscosity(c, use_law ? 0.01 : dt, c->velO))) return r;
    if ((r = st_gravity(c))) return r;
  }
  return WS_OK;
}

// the substep after the pressure solve: projection, confinement, clamp,
// extrapolation, G2P and advection, balls
int substep_post1(ws_ctx *c);
int substep_post2(ws_ctx *c);

int substep_post(ws_ctx *c) {
  int r;
  if ((r = substep_post1(c)) || (r = st_extrapolate(c)) || (r = substep_post2(c))) return r;
  return WS_OK;
}

// projection, confinement and clamp
int substep_post1(ws_ctx *c) {
  int r;
  bool vort = c->vorticity > 0;
  if ((r = st_project(c, vort ? 0 : 1))) return r;
  if (vort) {
    if ((r = st_vorticity(c))) return r;
    if ((r = st_clamp(c))) return r;
  }
  return WS_OK;
}

// after the extrapolation: G2P and advection, balls, the substep's end
int substep_post2(ws_ctx *c) {
  int r;
  if ((r = st_particles(c, 3))) return r;
  if ((r = balls_update(c))) return r;
  stage_launch(c, k_end_substep_ops, k_end_substep, 1, 64, c->ds);
  WS_LAUNCHED();
  return WS_OK;
}

int pull_scalars(ws_ctx *c) {
  if (int r = read_scalars(c)) return r;
  if (int r = device_failures(c)) return r;
  c->umax = __builtin_bit_cast(float, c->hs->umax_bits);
  c->airborne = (int)c->hs->airborne;
  c->gRamp = c->hs->gRamp;
  c->dtLast = c->hs->dtLast;
  return WS_OK;
}

int push_ramp(ws_ctx *c) {
  // gRamp lives on the device; write it without touching the other scalars
  WS_HIP(hipMemcpyAsync(&c->ds->gRamp, &c->gRamp, sizeof(double), hipMemcpyHostToDevice, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

struct Field {
  void *dev;
  size_t elem;
  long count;
  bool particle;
};

bool field_of(ws_ctx *c, const char *n, Field &f) {
  long F = c->uN + c->vN + c->wN;
  (void)F;
#define PF(nm, ptr, ty) if (!strcmp(n, nm)) { f = Field{(void *)(ptr), sizeof(ty), c->nP, true}; return true; }
#define GF(nm, ptr, ty, cnt) if (!strcmp(n, nm)) { f = Field{(void *)(ptr), sizeof(ty), (long)(cnt), false}; return true; }
  PF("px", c->px, float) PF("py", c->py, float) PF("pz", c->pz, float)
  PF("pvx", c->pvx, float) PF("pvy", c->pvy, float) PF("pvz", c->pvz, float)
  PF("pT", c->pT, float) PF("pAir", c->pAir, float)
  PF("pflag", c->pflag, uint8_t) PF("pcool", c->pcool, uint8_t)
  PF("advU", c->advU, float) PF("advV", c->advV, float) PF("advW", c->advW, float)
  PF("pLight", c->pLight, float) PF("pDepth", c->pDepth, uint8_t)
  PF("pWx", c->pWx, float) PF("pWy", c->pWy, float) PF("pWz", c->pWz, float) PF("pPh", c->pPh, float)
  PF("pcell", c->pcell, int)
  GF("terrainSolid", c->tsolid, uint8_t, (long)c->tn * c->tn * c->tn)
  GF("cellType", c->cellType, uint8_t, c->nC) GF("staticTypes", c->staticTypes, uint8_t, c->nC)
  GF("q", c->q, float, c->nC) GF("rhs", c->rhs, float, c->nC)
  GF("u", c->vel, float, c->uN) GF("v", c->vel + c->uN, float, c->vN) GF("w", c->vel + c->uN + c->vN, float, c->wN)
  GF("uO", c->velO, float, c->uN) GF("vO", c->velO + c->uN, float, c->vN)
  GF("wO", c->velO + c->uN + c->vN, float, c->wN)
  GF("uW", c->wgt, float, c->uN) GF("vW", c->wgt + c->uN, float, c->vN) GF("wW", c->wgt + c->uN + c->vN, float, c->wN)
  GF("validU", c->valid, uint8_t, c->uN) GF("validV", c->valid + c->uN, uint8_t, c->vN)
  GF("validW", c->valid + c->uN + c->vN, uint8_t, c->wN)
  GF("gu", c->gu, float, c->uN) GF("gv", c->gu + c->uN, float, c->vN) GF("gw", c->gu + c->uN + c->vN, float, c->wN)
  GF("dens", c->dens, float, c->nD)
  GF("terrainR", c->terrainR, float, (long)c->tn * c->tn * c->tn)
  if (c->sv) {
    GF("svx", c->sv, float, c->nC) GF("svy", c->sv + c->nC, float, c->nC) GF("svz", c->sv + 2 * c->nC, float, c->nC)
    GF("ballHist", c->ballHist, unsigned, (long)c->nBalls * 25) GF("ballSurf", c->ballSurf, float, c->nBalls)
  }
  if (c->mbField) GF("metaballField", c->mbField, float, (long)(c->mbM + 1) * (c->mbM + 1) * (c->mbM + 1))
#undef PF
#undef GF
  return false;
}

}  // namespace

// the halves of a device-dt-law substep around the pressure sweeps, for the
// batched driver (ws_batch.hip)
namespace ws {
int batch_substep_pre(ws_ctx *c) {
  int r = substep_pre(c, 0.01, 1);
  return r ? r : pressure_setup(c);
}
int batch_substep_post1(ws_ctx *c) { return substep_post1(c); }
int batch_substep_post2(ws_ctx *c) { return substep_post2(c); }
ExPlanet batch_ex_planet(ws_ctx *c) { return ExPlanet{c->geo, c->vel, c->valid}; }
SorJob batch_sor_job(ws_ctx *c) { return sor_job(c); }
void batch_solved(ws_ctx *c) { c->devStatsPending = true; }
}  // namespace ws

// the terrain view the frame passes use (ws_frame.hip)
World ws::make_world(ws_ctx *c) { return world_of(c); }

// ====================================================================== ABI
extern "C" {

int ws_create(ws_ctx **out, int nx, int ny, int nz, double dx, int capacity) {
  if (!out) return WS_EINVAL;
  *out = nullptr;
  if (nx < 4 || ny < 4 || nz < 4 || !(dx > 0) || capacity < 0) return WS_EINVAL;
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return WS_ENODEV;
  if (hipGetDevice(&dev) != hipSuccess) return WS_ENODEV;
  // the device address of this GPU's static scratch (ws_mt_build_noalloc)
  if (mt_scratch_setup(dev) != WS_OK) return WS_EHIP;
  ws_ctx *c = new ws_ctx();
  c->device = dev;
  c->nx = nx; c->ny = ny; c->nz = nz; c->dx = dx; c->capacity = capacity;
  // solver.js:84-91
  c->W = nx * dx; c->H = ny * dx; c->D = nz * dx;
  c->cx = c->W / 2; c->cy = c->H / 2; c->cz = c->D / 2;
  double m = c->W < c->H ? c->W : c->H;
  m = m < c->D ? m : c->D;
  c->domainR = 0.5 * m - dx;
  c->spacing = dx * 0.5;
  c->particleVolume = c->spacing * c->spacing * c->spacing;
  c->nC = (long)nx * ny * nz;
  c->uN = (long)(nx + 1) * ny * nz; c->vN = (long)nx * (ny + 1) * nz; c->wN = (long)nx * ny * (nz + 1);
  c->nD = (long)(nx + 1) * (ny + 1) * (nz + 1);
  Geo &g = c->geo;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.uN = (int)c->uN; g.vN = (int)c->vN; g.wN = (int)c->wN; g.nC = (int)c->nC;
  g.dx = (float)dx; g.inv_dx = (float)(1.0 / dx);
  g.dx_lo = (float)(dx - (double)g.dx);
  g.cx = (float)c->cx; g.cy = (float)c->cy; g.cz = (float)c->cz;
  g.cx_lo = (float)(c->cx - g.cx); g.cy_lo = (float)(c->cy - g.cy); g.cz_lo = (float)(c->cz - g.cz);
  long F = c->uN + c->vN + c->wN;
  int r = WS_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return WS_EHIP; }
  int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
  std::vector<float> thr = floor_thresholds(dx, nmax);
  long cap = capacity > 0 ? capacity : 1;
  if ((r = dalloc(c, &c->staticTypes, c->nC)) || (r = dalloc(c, &c->gu, F)) || (r = dalloc(c, &c->thr, thr.size())) ||
      (r = dalloc(c, &c->cellType, c->nC)) || (r = dalloc(c, &c->nbm, c->nC)) || (r = dalloc(c, &c->dg, c->nC)) ||
      (r = dalloc(c, &c->q, c->nC)) || (r = dalloc(c, &c->q2, c->nC)) || (r = dalloc(c, &c->rhs, c->nC)) || (r = dalloc(c, &c->pw, c->nC)) ||
      (r = dalloc(c, &c->cellCount, c->nC + 1)) || (r = dalloc(c, &c->cellStart, c->nC + 1)) ||
      (r = dalloc(c, &c->vort, 10 * c->nC)) || (r = dalloc(c, &c->vel, F)) || (r = dalloc(c, &c->velO, F)) ||
      (r = dalloc(c, &c->wgt, F)) || (r = dalloc(c, &c->scr, F)) || (r = dalloc(c, &c->scr2, F)) ||
      (r = dalloc(c, &c->valid, F)) || (r = dalloc(c, &c->dens, c->nD)) ||
      (r = dalloc(c, &c->px, cap)) || (r = dalloc(c, &c->py, cap)) || (r = dalloc(c, &c->pz, cap)) ||
      (r = dalloc(c, &c->pvx, cap)) || (r = dalloc(c, &c->pvy, cap)) || (r = dalloc(c, &c->pvz, cap)) ||
      (r = dalloc(c, &c->pT, cap)) || (r = dalloc(c, &c->pAir, cap)) || (r = dalloc(c, &c->advU, cap)) ||
      (r = dalloc(c, &c->advV, cap)) || (r = dalloc(c, &c->advW, cap)) || (r = dalloc(c, &c->pflag, cap)) ||
      (r = dalloc(c, &c->pcool, cap)) || (r = dalloc(c, &c->pcell, cap)) || (r = dalloc(c, &c->prank, cap)) ||
      (r = dalloc(c, &c->sortedIdx, cap)) || (r = dalloc(c, &c->vneed, 4 * cap)) ||
      (r = dalloc(c, &c->vscr, 4 * cap)) || (r = dalloc(c, &c->rng, 4)) || (r = dalloc(c, &c->ds, 1)) ||
      (r = dalloc(c, &c->pLight, cap)) || (r = dalloc(c, &c->pDepth, cap)) || (r = dalloc(c, &c->pWx, cap)) ||
      (r = dalloc(c, &c->pWy, cap)) || (r = dalloc(c, &c->pWz, cap)) || (r = dalloc(c, &c->pPh, cap)) ||
      (r = dalloc(c, &c->evBuf, 7 * cap + 4)) || (r = dalloc(c, &c->jmp, 2 * EV_WIN)) ||
      (r = dalloc(c, &c->mtCount, c->nC + 1))) {
    ws_destroy(c);
    return r;
  }
  c->sorTiles[0] = (nx + SOR_TX - 1) / SOR_TX;
  c->sorTiles[1] = (ny + SOR_TY - 1) / SOR_TY;
  c->sorTiles[2] = (nz + SOR_TZ - 1) / SOR_TZ;
  if (dalloc(c, &c->sorActive, (size_t)c->sorTiles[0] * c->sorTiles[1] * c->sorTiles[2])) { ws_destroy(c); return WS_EHIP; }
  // fluid-cell lists: red entries from 0, black from sorHalf; the list code
  // holds the cell index in 26 bits
  c->sorHalf = (long)(c->nC + 1) / 2 + 1;
  if (c->nC >= (1L << 26)) c->sorMode = 0;
  c->tbTiles[0] = (nx + SOR_TB_X - 1) / SOR_TB_X;
  c->tbTiles[1] = (ny + SOR_TB_Y - 1) / SOR_TB_Y;
  c->tbTiles[2] = (nz + SOR_TB_Z - 1) / SOR_TB_Z;
  if ((r = dalloc(c, &c->tbActive, (size_t)c->tbTiles[0] * c->tbTiles[1] * c->tbTiles[2]))) {
    ws_destroy(c);
    return r;
  }
  if ((r = dalloc(c, &c->sorOff, c->nC + 1)) || (r = dalloc(c, &c->sorCode, c->nC + 4)) ||
      (r = dalloc(c, &c->sorRhs, c->nC + 4)) || (r = dalloc(c, &c->sorPw, c->nC + 4)) ||
      (r = dalloc(c, &c->sorDg, c->nC + 4)) || (r = dalloc(c, &c->sorChunk, 2 * (c->nC / 2048 + 2))) ||
      (r = dalloc(c, &c->sorArrive, 2)) || (r = dalloc(c, &c->sorFlow, (size_t)kFlowStride * 256))) {
    ws_destroy(c);
    return r;
  }
  c->gv = c->gu + c->uN;
  c->gw = c->gu + c->uN + c->vN;
  // coherent: k_publish_scalars writes both while the host polls
  if (hipHostMalloc((void **)&c->hs, sizeof(DevScalars), hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&c->hsSeq, 64, hipHostMallocCoherent) != hipSuccess) {
    ws_destroy(c);
    return WS_EHIP;
  }
  *c->hsSeq = 0;
  c->nPartials = 1024;
  if (dalloc(c, &c->partials, 2 * c->nPartials + 2048)) { ws_destroy(c); return WS_EHIP; }
  // scan scratch: enough for the larger of the cell and particle scans
  size_t b1 = 0, b2 = 0;
  size_t b3 = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, b1, c->cellCount, c->cellStart, (int)c->nC + 1, c->stream) ||
      hipcub::DeviceScan::ExclusiveSum(nullptr, b2, c->vneed, c->vneed, (int)cap, c->stream) ||
      hipcub::DeviceScan::ExclusiveSum(nullptr, b3, c->sorOff, c->sorOff, (int)c->nC + 1, c->stream)) {
    ws_destroy(c);
    return WS_EHIP;
  }
  c->scanTmpBytes = b1 > b2 ? b1 : b2;
  if (b3 > c->scanTmpBytes) c->scanTmpBytes = b3;
  if (hipMalloc(&c->scanTmp, c->scanTmpBytes) != hipSuccess) { ws_destroy(c); return WS_EHIP; }
  c->scanTiles = scan_tiles_for(std::max((long)c->nC + 1, (long)cap));
  const unsigned epoch1 = 1;  // status words start at tag 0, below every epoch
  if (hipMalloc((void **)&c->scanStatus, sizeof(unsigned long long) * 2 * c->scanTiles) != hipSuccess ||
      hipMalloc((void **)&c->scanCtl, sizeof(unsigned) * 4) != hipSuccess ||
      hipMemset(c->scanStatus, 0, sizeof(unsigned long long) * 2 * c->scanTiles) != hipSuccess ||
      hipMemcpy(c->scanCtl, &epoch1, sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
    ws_destroy(c);
    return WS_EHIP;
  }
  // the zeroing of every buffer above (on the context's stream) is done
  // before the host copies below (on the null stream) write into some of them
  if (hipStreamSynchronize(c->stream) != hipSuccess) { ws_destroy(c); return WS_EHIP; }
  DevScalars init{};
  init.gRamp = c->gRamp;
  uint32_t rng[4] = {0x2f6e2b1u, 0x7f4a7c15u, 0x51ab3c77u, 0u};  // _cfS, _bfS, _evS (solver.js:151-153)
  if (hipMemcpy(c->thr, thr.data(), thr.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->ds, &init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->rng, rng, sizeof(rng), hipMemcpyHostToDevice) != hipSuccess) {
    ws_destroy(c);
    return WS_EHIP;
  }
  {
    // affine maps of 1..EV_WIN LCG steps: s -> A s + C (mod 2^32)
    std::vector<uint32_t> J(2 * EV_WIN);
    uint32_t A = 1u, C = 0u;
    for (int n = 0; n < EV_WIN; n++) {
      A = A * 1664525u;
      C = C * 1664525u + 1013904223u;
      J[n] = A;
      J[EV_WIN + n] = C;
    }
    if (hipMemcpy(c->jmp, J.data(), J.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
      ws_destroy(c);
      return WS_EHIP;
    }
  }
  // extrapolation tile schedule, grouped by tile diagonal I+J+K
  {
    int dims[3][3] = {{nx + 1, ny, nz}, {nx, ny + 1, nz}, {nx, ny, nz + 1}};
    std::vector<std::vector<int4>> byDiag;
    for (int comp = 0; comp < 3; comp++) {
      int ti = (dims[comp][0] + EX_TI - 1) / EX_TI, tj = (dims[comp][1] + EX_TJ - 1) / EX_TJ;
      int tk = (dims[comp][2] + EX_TK - 1) / EX_TK;
      for (int K = 0; K < tk; K++)
        for (int J = 0; J < tj; J++)
          for (int I = 0; I < ti; I++) {
            size_t d = (size_t)(I + J + K);
            if (byDiag.size() <= d) byDiag.resize(d + 1);
            byDiag[d].push_back(make_int4(comp, I, J, K));
          }
    }
    std::vector<int4> flat;
    c->exDiagOff.push_back(0);
    for (size_t d = 0; d < byDiag.size(); d++) {
      const auto &v = byDiag[d];
      flat.insert(flat.end(), v.begin(), v.end());
      c->exDiagOff.push_back((int)flat.size());
      ExDiag D{};
      D.d = (int)d;
      for (int comp = 0; comp < 3; comp++) {
        D.ti[comp] = (dims[comp][0] + EX_TI - 1) / EX_TI;
        D.tj[comp] = (dims[comp][1] + EX_TJ - 1) / EX_TJ;
        D.tk[comp] = (dims[comp][2] + EX_TK - 1) / EX_TK;
      }
      for (const int4 &t : v) D.cnt[t.x]++;
      c->exDiag.push_back(D);
    }
    if (hipMalloc(&c->exTiles, flat.size() * sizeof(int4)) != hipSuccess) { ws_destroy(c); return WS_EHIP; }
    if (hipMemcpy(c->exTiles, flat.data(), flat.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess) {
      ws_destroy(c);
      return WS_EHIP;
    }
    c->exTilesHost = flat;
  }
  if (hipDeviceSynchronize() != hipSuccess) { ws_destroy(c); return WS_EHIP; }
  *out = c;
  return WS_OK;
}

void ws_destroy(ws_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  net_free(c);
  contacts_release(c, true);
  balls_release(c);
  metaballs_release(c);
  void *ptrs[] = {c->staticTypes, c->gu, c->thr, c->terrainR, c->cellType, c->nbm, c->dg, c->q, c->q2, c->rhs, c->pw,
                  c->sorActive, c->tbActive, c->tbPk, c->tbRhs, c->tbCnt, c->sorOff, c->sorCode, c->sorRhs, c->sorPw, c->sorDg, c->sorChunk, c->sorArrive, c->sorFlow, c->xcdX, c->xcdFlow, c->exFin, c->exXk, c->exXi, c->exCnt,
                  c->cellCount, c->cellStart, c->vort, c->vel, c->velO, c->wgt, c->scr, c->scr2, c->valid, c->dens,
                  c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pT, c->pAir, c->advU, c->advV, c->advW, c->pflag,
                  c->pcool, c->pcell, c->prank, c->sortedIdx, c->vneed, c->vscr, c->rng, c->ds, c->partials,
                  c->scanTmp, c->exTiles, c->mtCount, c->mtPos, c->mtNrm, c->mtOff, c->tsolid, c->pDepth,
                  c->pLight, c->pWx, c->pWy, c->pWz, c->pPh, c->evBuf, c->jmp, c->currField, c->smpBuf};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->scanStatus) (void)hipFree(c->scanStatus);
  if (c->scanCtl) (void)hipFree(c->scanCtl);
  if (c->hs) (void)hipHostFree(c->hs);
  if (c->hsSeq) (void)hipHostFree(c->hsSeq);
  if (c->renderHost) (void)hipHostFree(c->renderHost);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &sp : c->spans) {
    (void)hipEventDestroy(sp.a);
    (void)hipEventDestroy(sp.b);
  }
  for (auto &e : c->evFree) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *ws_last_error(const ws_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ws_set_f(ws_ctx *c, const char *n, double v) {
  if (!c || !n) return WS_EINVAL;
#define SF(nm) if (!strcmp(n, #nm)) { c->nm = v; return WS_OK; }
#define SI(nm) if (!strcmp(n, #nm)) { c->nm = (int)v; return WS_OK; }
  SF(cfl) SF(pic) SF(gravity) SF(viscosity) SF(omega) SF(maxSpeed) SF(vorticity) SF(coreR) SF(oceanR)
  SF(atmosphereH) SF(spacing) SF(particleVolume) SF(iso) SF(substeps) SF(rainT) SF(cloudT) SF(rainLiftAge)
  SF(pressureTol) SF(dv) SF(sunX) SF(sunY) SF(sunZ) SF(dtFrameBase)
  SF(heatK) SF(Tamb) SF(shadeCool) SF(sunActivity) SF(evapIntensity) SF(snowT) SF(meltT) SF(cloudP)
  SF(currents) SF(Rlo) SF(Rhi) SF(simTime) SF(heatAcc) SF(netLossCoarse)
  SI(pressureIters) SI(maxPressureIters) SI(residualEvery) SI(ceilReflect) SI(hasSun) SI(measureResidual)
  SI(adaptiveChecks) SI(spinOn) SI(warmStart) SI(recordSamples) SI(pressureCompare) SI(netTiled) SI(netCoarse) SI(vaporSerialMax) SI(vaporQueueMax) SI(vaporLdsMax)
#undef SF
#undef SI
  if (!strcmp(n, "sorMode")) {
    int m = (int)v;
    if (m < 0 || m > 6 || m != v) { c->err = "sorMode must be 0..6"; return WS_EINVAL; }
    if (m >= 2 && c->nC >= (1L << 26)) { c->err = "sorMode 2..6 need fewer than 2^26 cells"; return WS_EINVAL; }
    c->sorMode = m;
    return WS_OK;
  }
  if (!strcmp(n, "sorGroups")) {
    int dev = 0, cus = 0;
    WS_HIP(hipGetDevice(&dev));
    WS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (!(v >= 1 && v <= cus) || v != (int)v) { c->err = "sorGroups must be 1 .. the compute units"; return WS_EINVAL; }
    c->sorGroups = (int)v;
    return WS_OK;
  }
  if (!strcmp(n, "xcdGroups")) {
    int dev = 0, cus = 0;
    WS_HIP(hipGetDevice(&dev));
    WS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (v != (int)v || v < 0 || v > std::min(kXcdMaxG, cus / 8)) {
      c->err = "xcdGroups must be 0 (one per compute unit of an XCD) .. the compute units of one XCD";
      return WS_EINVAL;
    }
    c->xcdGroups = (int)v;
    return WS_OK;
  }
  if (!strcmp(n, "exRows")) {
    if (v != 0 && v != 1) { c->err = "exRows must be 0 (tile diagonals) or 1 (row pipeline)"; return WS_EINVAL; }
    if (v == 1) {
      c->err = "exRows 1 (the row-pipelined extrapolation) is switched off: its first GPU run stalled, cause not found";
      return WS_EINVAL;
    }
    c->exRows = (int)v;
    return WS_OK;
  }
  if (!strcmp(n, "xcdUnits")) {
    if (v != (int)v || v < 0 || v > 3) { c->err = "xcdUnits must be 0 (chosen from the list lengths) .. 3"; return WS_EINVAL; }
    c->xcdUnits = (int)v;
    c->xcdMinUnits = 1;
    return WS_OK;
  }
  if (!strcmp(n, "flowGroups") || !strcmp(n, "flowXcds")) {
    int dev = 0, cus = 0;
    WS_HIP(hipGetDevice(&dev));
    WS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const bool g = n[4] == 'G';
    const int G = g ? (int)v : c->flowGroups, X = g ? c->flowXcds : (int)v;
    if (v != (int)v || X < 1 || X > 8 || G < 1 || G > 256 || G > cus / 8 * X) {
      c->err = "flowGroups must be 1..256 and at most the compute units of flowXcds (1..8) XCDs";
      return WS_EINVAL;
    }
    c->flowGroups = G;
    c->flowXcds = X;
    return WS_OK;
  }
  if (!strcmp(n, "p2gBin")) {
    if (v != (int)v || v < 0 || v >= P2G_NBINS) { c->err = "p2gBin must be 0..3 (8x8x8, 4x4x4, 16x8x8, 8x8x4)"; return WS_EINVAL; }
    c->p2gBin = (int)v;
    return WS_OK;
  }
  if (!strcmp(n, "sorThreads")) {
    if (v != 256 && v != 512 && v != 1024) { c->err = "sorThreads must be 256, 512 or 1024"; return WS_EINVAL; }
    c->sorThreads = (int)v;
    return WS_OK;
  }
  if (!strcmp(n, "gRamp")) { c->gRamp = v; return push_ramp(c); }
  if (!strcmp(n, "nP")) {
    if (v < 0 || v > c->capacity) { c->err = "nP exceeds capacity"; return WS_EINVAL; }
    c->nP = (int)v;
    return WS_OK;
  }
  if (!strcmp(n, "tn")) { c->tn = (int)v; return WS_OK; }
  if (!strcmp(n, "hasVapor")) { c->hasVapor = v != 0; return WS_OK; }
  if (!strcmp(n, "timing")) {
    int r = fold_dev_stats(c);  // solves before the reset stay out of the new totals
    if (r || (r = fold_spans(c))) return r;
    c->timing = v != 0;
    for (int i = 0; i < 4 && c->timing; i++)
      if (!c->ev[i]) WS_HIP(hipEventCreate(&c->ev[i]));
    while (c->timing && c->evFree.size() < 2 * kSpanMax + 2) {
      hipEvent_t e;
      WS_HIP(hipEventCreate(&e));
      c->evFree.push_back(e);
    }
    c->sorMs = c->pressureMs = 0;
    c->sorLaunches = c->pressureSolves = c->sweepsTotal = c->residualsTotal = c->ticksTotal = 0;
    c->sorCellSweeps = 0;
    c->cmpMsPlain = c->cmpMsLearned = c->netMs = 0;
    c->netCalls = 0;
    c->cmpSweepsPlain = c->cmpSweepsLearned = c->cmpSolves = 0;
    return WS_OK;
  }
  c->err = std::string("unknown scalar ") + n;
  return WS_EINVAL;
}

int ws_get_f(ws_ctx *c, const char *n, double *v) {
  if (!c || !n || !v) return WS_EINVAL;
  int r;
  if (!strcmp(n, "residual") || !strcmp(n, "sweepsTotal") || !strcmp(n, "residualsTotal") ||
      !strcmp(n, "pressureIters_done"))
    if ((r = fold_dev_stats(c))) return r;
  if (!strcmp(n, "sorMs") || !strcmp(n, "sorLaunches") || !strcmp(n, "pressureMs") ||
      !strcmp(n, "pressureSolves") || !strcmp(n, "netMs") || !strcmp(n, "netCalls"))
    if ((r = fold_spans(c))) return r;
  if (!strcmp(n, "umax") || !strcmp(n, "airborneCount") || !strcmp(n, "gRamp") || !strcmp(n, "dtLast") ||
      !strcmp(n, "dfRefined")) {
    if ((r = pull_scalars(c))) return r;
    if (!strcmp(n, "umax")) *v = c->umax;
    else if (!strcmp(n, "airborneCount")) *v = c->airborne;
    else if (!strcmp(n, "gRamp")) *v = c->gRamp;
    else if (!strcmp(n, "dtLast")) *v = c->dtLast;
    else *v = c->hs->df_refined;
    return WS_OK;
  }
#define GF(nm) if (!strcmp(n, #nm)) { *v = c->nm; return WS_OK; }
  GF(cfl) GF(pic) GF(gravity) GF(viscosity) GF(omega) GF(maxSpeed) GF(vorticity) GF(coreR) GF(oceanR)
  GF(atmosphereH) GF(spacing) GF(particleVolume) GF(iso) GF(substeps) GF(pressureTol) GF(pressureIters)
  GF(maxPressureIters) GF(nP) GF(capacity) GF(residual) GF(domainR) GF(dx) GF(tn) GF(dv)
  GF(sorMs) GF(pressureMs) GF(sorLaunches) GF(pressureSolves) GF(sweepsTotal) GF(residualsTotal) GF(ticksTotal) GF(sorMode) GF(sorCellSweeps) GF(xcdFallbacks) GF(xcdGroups) GF(xcdUnits) GF(exRows)
  GF(cmpMsPlain) GF(cmpMsLearned) GF(cmpSweepsPlain) GF(cmpSweepsLearned) GF(cmpSolves) GF(warmStart) GF(netMs) GF(netCalls)
  GF(vaporCount) GF(nightVaporCount) GF(contactOverflow) GF(nBalls) GF(mbM) GF(vaporBigGroups) GF(vaporGroupMax)
  GF(vaporBigParcels) GF(netTiled) GF(netCoarse) GF(p2gBin) GF(sorThreads) GF(sorGroups) GF(flowGroups) GF(flowXcds)
  GF(heatK) GF(Tamb) GF(shadeCool) GF(sunActivity) GF(evapIntensity) GF(snowT) GF(meltT) GF(cloudP) GF(currents)
  GF(simTime) GF(heatAcc) GF(evapCount) GF(cloudT) GF(rainT) GF(rainLiftAge)
#undef GF
  if (!strcmp(n, "pressureIters_done")) { *v = c->itersDone; return WS_OK; }
  if (!strcmp(n, "rngCfS") || !strcmp(n, "rngBfS") || !strcmp(n, "rngEvS")) {
    uint32_t rng[4];
    // on the context's stream: ordered after its queued substeps
    WS_HIP(hipMemcpyAsync(rng, c->rng, sizeof(rng), hipMemcpyDeviceToHost, c->stream));
    WS_HIP(hipStreamSynchronize(c->stream));
    *v = n[4] == 'C' ? rng[0] : (n[4] == 'B' ? rng[1] : rng[2]);
    return WS_OK;
  }
  c->err = std::string("unknown scalar ") + n;
  return WS_EINVAL;
}

int ws_upload(ws_ctx *c, const char *name, const void *host, size_t nbytes) {
  if (!c || !name || (!host && nbytes)) return WS_EINVAL;
  Field f;
  if (!strcmp(name, "pcell")) { c->err = "pcell is read-only (the marking pass writes it)"; return WS_EINVAL; }
  if (!strcmp(name, "terrainR")) {
    long n = (long)c->tn * c->tn * c->tn;
    if (n <= 0 || nbytes != (size_t)n * 4) { c->err = "terrainR: set tn first; size must be tn^3 floats"; return WS_EINVAL; }
    if (c->terrainR) (void)hipFree(c->terrainR);
    c->terrainR = nullptr;
    int r = dalloc(c, &c->terrainR, n);
    if (r) return r;
  }
  if (!strcmp(name, "terrainSolid")) {
    long n = (long)c->tn * c->tn * c->tn;
    if (n <= 0 || nbytes != (size_t)n) { c->err = "terrainSolid: set tn first; size must be tn^3 bytes"; return WS_EINVAL; }
    if (c->tsolid) (void)hipFree(c->tsolid);
    c->tsolid = nullptr;
    int r = dalloc(c, &c->tsolid, n);
    if (r) return r;
  }
  if (!strcmp(name, "rng")) {
    if (nbytes != 12) return WS_EINVAL;
    WS_HIP(hipMemcpyAsync(c->rng, host, 12, hipMemcpyHostToDevice, c->stream));
    WS_HIP(hipStreamSynchronize(c->stream));
    return WS_OK;
  }
  if (!field_of(c, name, f)) { c->err = std::string("unknown field ") + name; return WS_EINVAL; }
  if (nbytes != f.elem * (size_t)f.count) {
    c->err = std::string("size mismatch for ") + name + ": expected " + std::to_string(f.elem * (size_t)f.count);
    return WS_EINVAL;
  }
  WS_HIP(hipMemcpyAsync(f.dev, host, nbytes, hipMemcpyHostToDevice, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  if (!strcmp(name, "staticTypes")) c->haveStatic = true;
  if (!strcmp(name, "terrainR")) c->haveTerrain = true;
  if (!strcmp(name, "terrainSolid")) c->haveSolid = 1;
  if (!strcmp(name, "pflag")) {
    const uint8_t *fl = (const uint8_t *)host;
    for (long i = 0; i < f.count; i++)
      if (fl[i] == 2 || fl[i] == 3) { c->hasVapor = true; break; }
  }
  return WS_OK;
}

int ws_download(ws_ctx *c, const char *name, void *host, size_t nbytes) {
  if (!c || !name || (!host && nbytes)) return WS_EINVAL;
  if (!strcmp(name, "rng")) {
    if (nbytes != 12) return WS_EINVAL;
    WS_HIP(hipMemcpyAsync(host, c->rng, 12, hipMemcpyDeviceToHost, c->stream));
    WS_HIP(hipStreamSynchronize(c->stream));
    return WS_OK;
  }
  Field f;
  if (!field_of(c, name, f)) { c->err = std::string("unknown field ") + name; return WS_EINVAL; }
  if (nbytes != f.elem * (size_t)f.count) {
    c->err = std::string("size mismatch for ") + name + ": expected " + std::to_string(f.elem * (size_t)f.count);
    return WS_EINVAL;
  }
  WS_HIP(hipMemcpyAsync(host, f.dev, nbytes, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_substep(ws_ctx *c, double dt) {
  if (!c || !(dt > 0)) return WS_EINVAL;
  int r = substep(c, dt, 0);
  if (r) return r;
  if ((r = pull_scalars(c))) return r;
  c->dtLast = dt;
  return WS_OK;
}

int ws_tune(ws_ctx *c, int sweeps, char *report, size_t len) {
  if (!c || sweeps < 1) return WS_EINVAL;
  int r;
  if ((r = need_world(c))) return r;
  for (int i = 0; i < 4; i++)
    if (!c->ev[i]) WS_HIP(hipEventCreate(&c->ev[i]));
  if ((r = fold_dev_stats(c))) return r;
  // the state the trial solves must leave untouched
  const double keepRes = c->residual;
  const int keepMode = c->sorMode, keepIters = c->pressureIters, keepRec = c->recordSamples, keepWarm = c->warmStart;
  const int keepTiled = c->netTiled, keepTiming = c->timing;
  const double keepTol = c->pressureTol;
  const long keep[] = {c->sweepsTotal, c->residualsTotal, c->sorLaunches, c->sorCellSweeps, c->pressureSolves,
                       c->netCalls, (long)c->itersDone};
  const double keepMs[] = {c->sorMs, c->pressureMs, c->netMs};
  float *saved = nullptr;
  WS_HIP(hipMalloc((void **)&saved, sizeof(float) * c->nC));
  WS_HIP(hipMemcpyAsync(saved, c->q, sizeof(float) * c->nC, hipMemcpyDeviceToDevice, c->stream));
  c->recordSamples = 0;
  c->timing = 0;
  c->pressureTol = 0;
  c->pressureIters = sweeps;
  std::string rep;
  auto trial = [&](float *ms) -> int {
    WS_HIP(hipMemcpyAsync(c->q, saved, sizeof(float) * c->nC, hipMemcpyDeviceToDevice, c->stream));
    WS_HIP(hipEventRecord(c->ev[2], c->stream));
    int e = st_pressure_solve(c);
    if (e) return e;
    WS_HIP(hipEventRecord(c->ev[3], c->stream));
    WS_HIP(hipEventSynchronize(c->ev[3]));
    WS_HIP(hipEventElapsedTime(ms, c->ev[2], c->ev[3]));
    return WS_OK;
  };
  // SOR kernels: per-colour grid, LDS tile, fluid list, two sweeps per
  // launch, the one-workgroup solve at three workgroup sizes and the
  // neighbour-synchronised solve
  c->warmStart = 0;
  const int keepThreads = c->sorThreads;
  int best = keepMode, bestThreads = keepThreads;
  float bestMs = 1e30f;
  const int cand[][2] = {{0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 256}, {4, 512}, {4, 1024}, {5, 256}, {6, 0}};
  for (const auto &cm : cand) {
    if (r) break;
    if (cm[0] >= 2 && c->nC >= (1L << 26)) continue;
    c->sorMode = cm[0];
    if (cm[1]) c->sorThreads = cm[1];
    const long fallbacks = c->xcdFallbacks;
    float ms = 0;
    for (int rep2 = 0; rep2 < 2 && !r; rep2++) r = trial(&ms);  // the second run is timed
    if (r) break;
    char line[128];
    if (cm[0] == 6 && c->xcdFallbacks != fallbacks) {
      // the XCD-resident sweeps declined (placement or list length): not a candidate
      snprintf(line, sizeof line, "sorMode 6: declined, not a candidate\n");
      rep += line;
      continue;
    }
    if (cm[1])
      snprintf(line, sizeof line, "sorMode %d (%d threads): %d sweeps in %.3f ms\n", cm[0], cm[1], sweeps, ms);
    else
      snprintf(line, sizeof line, "sorMode %d: %d sweeps in %.3f ms\n", cm[0], sweeps, ms);
    rep += line;
    if (ms < bestMs) { bestMs = ms; best = cm[0]; bestThreads = cm[1] ? cm[1] : keepThreads; }
  }
  c->sorMode = best;
  c->sorThreads = bestThreads;
  // the warm start's convolutions: untiled, LDS-tiled where the grid fills
  // the GPU (split K elsewhere), or LDS-tiled on every level
  int bestTiled = keepTiled;
  if (!r && keepWarm && c->net) {
    float bt = 1e30f;
    for (int tiled = 0; tiled <= 2 && !r; tiled++) {
      c->netTiled = tiled;
      c->warmStart = 1;
      float ms = 0;
      for (int rep2 = 0; rep2 < 2 && !r; rep2++) r = trial(&ms);
      if (r) break;
      char line[128];
      snprintf(line, sizeof line, "netTiled %d: solve with inference %.3f ms\n", tiled, ms);
      rep += line;
      if (ms < bt) { bt = ms; bestTiled = tiled; }
    }
  }
  c->netTiled = bestTiled;
  // P2G bins: every bin gives the same sums; the faces and weights are kept
  // aside and put back
  int bestBin = c->p2gBin;
  if (!r && c->nP > 0) {
    const long nf = nfaces(c);
    float *keepVel = nullptr, *keepWgt = nullptr;
    if (hipMalloc((void **)&keepVel, sizeof(float) * nf) != hipSuccess ||
        hipMalloc((void **)&keepWgt, sizeof(float) * nf) != hipSuccess) {
      c->err = "ws_tune: no memory for the P2G trials";
      r = WS_EHIP;
    }
    if (!r && (hipMemcpyAsync(keepVel, c->vel, sizeof(float) * nf, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
               hipMemcpyAsync(keepWgt, c->wgt, sizeof(float) * nf, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)) {
      c->err = "ws_tune: copy failed";
      r = WS_EHIP;
    }
    const int keepBin = c->p2gBin;
    float bt = 1e30f;
    for (int bin = 0; bin < P2G_NBINS && !r; bin++) {
      c->p2gBin = bin;
      float ms = 0;
      for (int rep2 = 0; rep2 < 2 && !r; rep2++) {
        if (hipEventRecord(c->ev[2], c->stream) != hipSuccess) { r = WS_EHIP; break; }
        if ((r = st_p2g(c))) break;
        if (hipEventRecord(c->ev[3], c->stream) != hipSuccess || hipEventSynchronize(c->ev[3]) != hipSuccess ||
            hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) != hipSuccess) {
          c->err = "ws_tune: P2G timing failed";
          r = WS_EHIP;
        }
      }
      if (r) break;
      char line[128];
      snprintf(line, sizeof line, "p2gBin %d (%dx%dx%d): %.3f ms\n", bin, P2G_BINS[bin][0], P2G_BINS[bin][1],
               P2G_BINS[bin][2], ms);
      rep += line;
      if (ms < bt) { bt = ms; bestBin = bin; }
    }
    c->p2gBin = r ? keepBin : bestBin;
    if (keepVel && keepWgt && !r &&
        (hipMemcpyAsync(c->vel, keepVel, sizeof(float) * nf, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
         hipMemcpyAsync(c->wgt, keepWgt, sizeof(float) * nf, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess)) {
      c->err = "ws_tune: restoring the faces failed";
      r = WS_EHIP;
    }
    (void)hipStreamSynchronize(c->stream);
    if (keepVel) (void)hipFree(keepVel);
    if (keepWgt) (void)hipFree(keepWgt);
  }
  // put everything back
  hipError_t e1 = hipMemcpyAsync(c->q, saved, sizeof(float) * c->nC, hipMemcpyDeviceToDevice, c->stream);
  hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(saved);
  int rf = fold_dev_stats(c);  // the trial solves' device counts, absorbed and then undone below
  if (!r) r = rf;
  c->residual = keepRes;
  c->recordSamples = keepRec;
  c->warmStart = keepWarm;
  c->timing = keepTiming;
  c->pressureTol = keepTol;
  c->pressureIters = keepIters;
  c->sweepsTotal = keep[0]; c->residualsTotal = keep[1]; c->sorLaunches = keep[2]; c->sorCellSweeps = keep[3];
  c->pressureSolves = keep[4]; c->netCalls = keep[5]; c->itersDone = (int)keep[6];
  c->sorMs = keepMs[0]; c->pressureMs = keepMs[1]; c->netMs = keepMs[2];
  if (r) return r;
  if (e1 != hipSuccess || e2 != hipSuccess) { c->err = "ws_tune: restoring the pressure failed"; return WS_EHIP; }
  if (report && len) {
    snprintf(report, len, "%schosen: sorMode %d, sorThreads %d, netTiled %d, p2gBin %d", rep.c_str(), c->sorMode,
             c->sorThreads, c->netTiled, c->p2gBin);
  }
  return WS_OK;
}

int ws_run_substeps(ws_ctx *c, int n) {
  if (!c || n < 0) return WS_EINVAL;
  for (int i = 0; i < n; i++) {
    int r = substep(c, 0.01, 1);
    if (r) return r;
  }
  return WS_OK;
}

int ws_step_frame(ws_ctx *c, double dtFrame, int *n_sub) {
  if (!c || !(dtFrame > 0)) return WS_EINVAL;
  int r;
  if ((r = need_world(c))) return r;
  // _stepSync, solver.js:2670-2700: the 25 Hz thermal / evaporation /
  // currents tick, the CFL substep loop, then the frame-end phase pass
  c->simTime += dtFrame;
  c->heatAcc += dtFrame;
  if (c->heatAcc >= 1.0 / 25) {
    double dtTick = c->heatAcc < 1.0 / 12 ? c->heatAcc : 1.0 / 12;
    if ((r = frame_heat(c, dtTick))) return r;
    if (c->sunActivity > 0 && (r = frame_evaporation(c, dtTick))) return r;
    if (c->currents > 0 && (r = frame_currents(c, dtTick))) return r;
    c->heatAcc = 0;
    c->ticksTotal++;
  }
  if ((r = pull_scalars(c))) return r;
  double remaining = dtFrame, base = dtFrame / (c->substeps > 1 ? c->substeps : 1);
  int n = 0;
  while (remaining > 1e-6 && n < 10) {
    double um = c->umax > 0.08 ? c->umax : 0.08;
    double dt = c->cfl * c->dx / um;
    if (base < dt) dt = base;
    if (dt > 0.01) dt = 0.01;
    if (dt < 1e-5) dt = 1e-5;
    if ((r = substep(c, dt, 0))) return r;
    if ((r = pull_scalars(c))) return r;
    remaining -= dt;
    n++;
  }
  // frame-end passes in the reference's order (solver.js:2696-2699)
  if ((r = frame_vapor_collisions(c))) return r;
  if ((r = frame_phase(c, dtFrame))) return r;
  if ((r = frame_ice_contacts(c))) return r;
  if ((r = frame_push_surface(c))) return r;
  WS_HIP(hipStreamSynchronize(c->stream));
  if (n_sub) *n_sub = n;
  return WS_OK;
}

int ws_stage(ws_ctx *c, int stage, double dt) {
  if (!c) return WS_EINVAL;
  int r = need_world(c);
  if (r) return r;
  if (stage != WS_STAGE_GRAVITY)
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, c->stream, c->ds, dt);
  switch (stage) {
    case WS_STAGE_CLEAR:
      WS_HIP(hipMemsetAsync(c->vel, 0, sizeof(float) * nfaces(c), c->stream));
      WS_HIP(hipMemsetAsync(c->wgt, 0, sizeof(float) * nfaces(c), c->stream));
      break;
    case WS_STAGE_RASTERIZE: r = st_rasterize(c); break;
    case WS_STAGE_P2G: r = st_p2g(c); break;
    case WS_STAGE_BC: r = st_bc(c, c->scr, 0); break;
    case WS_STAGE_SNAPSHOT:
      WS_HIP(hipMemcpyAsync(c->velO, c->vel, sizeof(float) * nfaces(c), hipMemcpyDeviceToDevice, c->stream));
      break;
    case WS_STAGE_VISCOSITY:
      WS_HIP(hipMemcpyAsync(c->scr, c->vel, sizeof(float) * nfaces(c), hipMemcpyDeviceToDevice, c->stream));
      r = st_viscosity(c, dt, c->scr);
      break;
    case WS_STAGE_GRAVITY:
      stage_launch(c, k_begin_substep_ops, k_begin_substep, 1, 64, c->ds, dt, 0, 0.0, 0.0, (float)c->gravity);
      r = st_gravity(c);
      break;
    case WS_STAGE_PRESSURE: r = st_pressure(c); break;
    case WS_STAGE_PROJECT: r = st_project(c, 0); break;
    case WS_STAGE_VORTICITY: r = st_vorticity(c); break;
    case WS_STAGE_CLAMP:
      stage_launch(c, k_reset_counters_ops, k_reset_counters, 1, 64, c->ds);
      r = st_clamp(c);
      break;
    case WS_STAGE_EXTRAPOLATE:
      hipLaunchKernelGGL(k_valid, dim3(blocks_for(nfaces(c), 256)), dim3(256), 0, c->stream, c->geo, c->cellType,
                         c->wgt, c->valid);
      r = st_extrapolate(c);
      break;
    case WS_STAGE_G2P:
      stage_launch(c, k_reset_counters_ops, k_reset_counters, 1, 64, c->ds);
      r = st_particles(c, 1);
      break;
    case WS_STAGE_ADVECT: r = st_particles(c, 2); break;
    case WS_STAGE_HEAT: r = frame_heat(c, dt); break;
    case WS_STAGE_EVAPORATION: r = frame_evaporation(c, dt); break;
    case WS_STAGE_CURRENTS: r = frame_currents(c, dt); break;
    case WS_STAGE_PHASE: r = frame_phase(c, dt); break;
    case WS_STAGE_PUSH_SURFACE: r = frame_push_surface(c); break;
    case WS_STAGE_VAPOR_COLLISIONS: r = frame_vapor_collisions(c); break;
    case WS_STAGE_ICE_CONTACTS: r = frame_ice_contacts(c); break;
    case WS_STAGE_PUSH_BALLS: r = balls_push(c, true); break;
    case WS_STAGE_UPDATE_BALLS: r = balls_update(c); break;
    default: c->err = "unknown stage"; return WS_EINVAL;
  }
  if (r) return r;
  WS_HIP(hipStreamSynchronize(c->stream));
  return pull_scalars(c);
}

int ws_read_render(ws_ctx *c, unsigned mask, ws_render_view *view) {
  if (!c || !view) return WS_EINVAL;
  struct Src {
    unsigned bit;
    const void *dev;
    size_t elem;
    const void **dst;
  };
  const Src src[] = {
      {WS_RF_PX, c->px, 4, (const void **)&view->px},       {WS_RF_PY, c->py, 4, (const void **)&view->py},
      {WS_RF_PZ, c->pz, 4, (const void **)&view->pz},       {WS_RF_PFLAG, c->pflag, 1, (const void **)&view->pflag},
      {WS_RF_PT, c->pT, 4, (const void **)&view->pT},       {WS_RF_PLIGHT, c->pLight, 4, (const void **)&view->pLight},
      {WS_RF_PDEPTH, c->pDepth, 1, (const void **)&view->pDepth}, {WS_RF_PPH, c->pPh, 4, (const void **)&view->pPh},
      {WS_RF_PWX, c->pWx, 4, (const void **)&view->pWx},    {WS_RF_PWY, c->pWy, 4, (const void **)&view->pWy},
      {WS_RF_PWZ, c->pWz, 4, (const void **)&view->pWz}};
  size_t n = (size_t)(c->nP > 0 ? c->nP : 0), need = 0;
  for (const Src &f : src)
    if (mask & f.bit) need += (f.elem * n + 255) & ~(size_t)255;
  if (need > c->renderCap) {
    if (c->renderHost) (void)hipHostFree(c->renderHost);
    c->renderHost = nullptr;
    c->renderCap = 0;
    size_t cap = need + need / 4;
    WS_HIP(hipHostMalloc((void **)&c->renderHost, cap, 0));
    c->renderCap = cap;
  }
  size_t off = 0;
  view->nP = (int)n;
  for (const Src &f : src) {
    *f.dst = nullptr;
    if (!(mask & f.bit)) continue;
    if (!f.dev) { c->err = "render readback: field not allocated"; return WS_ESTATE; }
    if (n) WS_HIP(hipMemcpyAsync(c->renderHost + off, f.dev, f.elem * n, hipMemcpyDeviceToHost, c->stream));
    *f.dst = c->renderHost + off;
    off += (f.elem * n + 255) & ~(size_t)255;
  }
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_sample_vel(ws_ctx *c, const float *xyz, long n, float *out) {
  if (!c || n < 0 || (n && (!xyz || !out))) return WS_EINVAL;
  if (!n) return WS_OK;
  if (n > c->smpCap) {
    if (c->smpBuf) (void)hipFree(c->smpBuf);
    c->smpBuf = nullptr;
    c->smpCap = 0;
    int r = dalloc(c, &c->smpBuf, 6 * n);
    if (r) return r;
    c->smpCap = n;
  }
  WS_HIP(hipMemcpyAsync(c->smpBuf, xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_sample_vel, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, c->geo, n, c->smpBuf, c->vel,
                     c->smpBuf + 3 * n);
  WS_LAUNCHED();
  WS_HIP(hipMemcpyAsync(out, c->smpBuf + 3 * n, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_mark_fluid(ws_ctx *c, int p0, int p1) {
  if (!c || p0 < 0 || p1 < p0 || p1 > c->nP) return WS_EINVAL;
  int r;
  if ((r = need_world(c))) return r;
  if (p1 > p0)
    stage_launch(c, k_mark_ops, k_mark, blocks_for(p1 - p0, 256), 256, c->geo, world_of(c), p1 - p0, c->px + p0,
                 c->py + p0, c->pz + p0, c->pflag + p0, c->cellType, c->cellCount, c->pcell + p0, c->prank + p0, c->ds);
  WS_LAUNCHED();
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

int ws_sync(ws_ctx *c) {
  if (!c) return WS_EINVAL;
  WS_HIP(hipStreamSynchronize(c->stream));
  return WS_OK;
}

}  // extern "C"
