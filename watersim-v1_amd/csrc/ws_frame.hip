// Host orchestration of the frame-level passes: the 25 Hz thermal tick
// (_updateHeat, _updateEvaporation, _updateCurrents; js/solver.js:2672-2681)
// and the phase pass at the end of a frame (_updatePhaseChanges, :2697).
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "ws_frame.h"

namespace ws {

namespace {

__global__ void k_zero_words(unsigned *p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

int zero_words(ws_ctx *c, void *p, long words) {
  hipLaunchKernelGGL(k_zero_words, dim3(blocks_for(words, 256)), dim3(256), 0, c->stream, (unsigned *)p, words);
  WS_LAUNCHED();
  return WS_OK;
}

ThermalPar thermal_par(ws_ctx *c, double dt) {
  ThermalPar T;
  T.dt = (float)dt;
  T.hasSun = c->hasSun;
  T.sunX = (float)c->sunX; T.sunY = (float)c->sunY; T.sunZ = (float)c->sunZ;
  T.sunPow = (float)(c->sunActivity / 3);
  T.shadeCool = (float)c->shadeCool;
  T.Tamb = (float)c->Tamb;
  double a = c->heatK * dt * 0.15;
  T.alpha = (float)(a < 0.14 ? a : 0.14);
  T.Rlo2 = (float)(c->Rlo * c->Rlo); T.Rhi2 = (float)(c->Rhi * c->Rhi); T.Rhi = (float)c->Rhi;
  T.coreR2 = (float)(c->coreR * c->coreR);
  T.inv_dv = c->dv > 0 ? (float)(1.0 / c->dv) : 0.f;
  T.hasTerrain = c->haveSolid && c->tsolid;
  T.tn = c->tn;
  T.conduct = c->heatK > 0;
  return T;
}

}  // namespace

// _updateHeat, solver.js:2715-2728
int frame_heat(ws_ctx *c, double dt) {
  if (c->nP < 1) return WS_OK;
  ThermalPar T = thermal_par(c, dt);
  int nb = blocks_for(c->nP, 256);
  unsigned long long *csum = (unsigned long long *)c->vort;
  unsigned *ccnt = (unsigned *)(c->vort + 2 * c->nC);
  float *mean = c->vort + 3 * c->nC, *dT = c->vort + 4 * c->nC;
  int r;
  if (T.conduct && (r = zero_words(c, c->vort, 3 * c->nC))) return r;
  hipLaunchKernelGGL(k_heat_particle, dim3(nb), dim3(256), 0, c->stream, c->geo, T, c->nP, c->px, c->py, c->pz, c->pT,
                     c->pflag, c->pLight, c->pDepth, c->cellType, c->dens, c->tsolid, c->thr, csum, ccnt);
  WS_LAUNCHED();
  if (T.conduct) {
    hipLaunchKernelGGL(k_heat_cells, dim3(blocks_for(c->nC, 256)), dim3(256), 0, c->stream, c->geo, T.alpha, csum, ccnt,
                       mean, dT);
    WS_LAUNCHED();
  }
  hipLaunchKernelGGL(k_heat_apply, dim3(nb), dim3(256), 0, c->stream, c->geo, T, c->nP, c->px, c->py, c->pz, c->pT,
                     c->pflag, c->pvx, c->pvy, c->pvz, mean, dT, c->thr);
  WS_LAUNCHED();
  return WS_OK;
}

// _updateEvaporation, solver.js:3008-3116
int frame_evaporation(ws_ctx *c, double dt) {
  if (c->sunActivity <= 0 || c->evapIntensity <= 0 || c->nP < 1) return WS_OK;
  EvapPar E;
  E.dt = dt;
  E.rate = c->sunActivity * c->evapIntensity;
  E.freezeT = c->meltT > c->snowT * 1.05 ? c->meltT : c->snowT * 1.05;
  E.dx = c->dx;
  E.cx = c->cx;
  E.cy = c->cy;
  E.cz = c->cz;
  long cap = c->capacity;
  unsigned *cand = c->evBuf, *off = c->evBuf + cap, *thresh = c->evBuf + 2 * cap, *list = c->evBuf + 3 * cap;
  unsigned *thrC = c->evBuf + 4 * cap, *evList = c->evBuf + 5 * cap, *evState = c->evBuf + 6 * cap;
  unsigned *evCount = c->evBuf + 7 * cap;
  int nb = blocks_for(c->nP, 256);
  hipLaunchKernelGGL(k_evap_candidates, dim3(nb), dim3(256), 0, c->stream, c->geo, E, c->nP, c->px, c->py, c->pz,
                     c->pT, c->pflag, c->cellType, c->thr, cand, thresh);
  WS_LAUNCHED();
  if (int r = scan_excl(c, cand, off, c->nP)) return r;
  hipLaunchKernelGGL(k_evap_compact, dim3(nb), dim3(256), 0, c->stream, c->nP, cand, off, thresh, list, thrC);
  hipLaunchKernelGGL(k_evap_chooser, dim3(1), dim3(EV_THREADS), 0, c->stream, thrC, off, cand, c->nP, c->jmp,
                     c->jmp + EV_WIN, c->rng, evList, evState, evCount);
  hipLaunchKernelGGL(k_evap_apply, dim3(nb), dim3(256), 0, c->stream, c->geo, E, list, evList, evState, evCount, c->px,
                     c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pT, c->pflag, c->pAir, c->vel, c->thr);
  WS_LAUNCHED();
  unsigned n = 0;
  WS_HIP(hipMemcpyAsync(&n, evCount, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  c->evapCount = (int)n;
  if (n) c->hasVapor = true;
  return WS_OK;
}

// _buildCurrWaves, solver.js:3635-3656: three octaves of three seeded waves
static void curr_waves(CurrPar &Q) {
  double seed = 1234567;
  auto rnd = [&]() {
    seed = std::fmod(seed * 1664525 + 1013904223, 4294967296.0);
    return seed / 4294967296.0;
  };
  const double K[3] = {1.5, 4.0, 9.0}, amp[3] = {0.10, 0.085, 0.05}, drift[3] = {0.021, 0.05, 0.11};
  for (int o = 0; o < 3; o++) {
    for (int cmp = 0; cmp < 3; cmp++) {
      double rx = rnd() * 2 - 1, ry = rnd() * 2 - 1, rz = rnd() * 2 - 1;
      double L = std::sqrt(rx * rx + ry * ry + rz * rz);
      if (!(L != 0)) L = 1;
      Q.wave[o][cmp][0] = (float)(rx / L * K[o]);
      Q.wave[o][cmp][1] = (float)(ry / L * K[o]);
      Q.wave[o][cmp][2] = (float)(rz / L * K[o]);
      Q.wave[o][cmp][3] = (float)(amp[o] / K[o]);
      Q.wave[o][cmp][4] = (float)(rnd() * 6.283);
    }
    Q.drift[o] = (float)drift[o];
  }
}

// _updateCurrents, solver.js:3658-3665
int frame_currents(ws_ctx *c, double dt) {
  double shell = c->oceanR - c->coreR;
  if (!(c->currents > 0) || !(shell > 0) || c->nP < 1) return WS_OK;
  CurrPar Q;
  curr_waves(Q);
  const int stride = 4;
  Q.gnx = c->nx / stride + 2; Q.gny = c->ny / stride + 2; Q.gnz = c->nz / stride + 2;
  Q.kMax = c->nz / stride;
  Q.sx = (float)(c->dx * stride);
  Q.invS = (float)(1.0 / (c->dx * stride));
  Q.t = (float)c->simTime;
  Q.amp = (float)(c->currents * 2);
  double relax = 2.2 * dt;
  Q.relax = (float)(relax > 0.35 ? 0.35 : relax);
  Q.oceanR = (float)c->oceanR;
  Q.shell = (float)shell;
  long gN = (long)Q.gnx * Q.gny * Q.gnz;
  if (c->currN < gN) {
    if (c->currField) (void)hipFree(c->currField);
    c->currField = nullptr;
    WS_HIP(hipMalloc((void **)&c->currField, sizeof(float) * 3 * gN));
    c->currN = gN;
  }
  hipLaunchKernelGGL(k_currents_lattice, dim3(blocks_for(gN, 256)), dim3(256), 0, c->stream, Q, c->currField);
  hipLaunchKernelGGL(k_currents_relax, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->geo, Q, c->nP, c->px,
                     c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pflag, c->currField);
  WS_LAUNCHED();
  return WS_OK;
}

// _updatePhaseChanges, solver.js:3152-3281
int frame_phase(ws_ctx *c, double dt) {
  if (c->nP < 1) return WS_OK;
  PhasePar P;
  double rainT = c->rainT < c->cloudT - 0.02 ? c->rainT : c->cloudT - 0.02;
  if (rainT < 0) rainT = 0;
  double snowT = c->snowT < rainT - 0.02 ? c->snowT : rainT - 0.02;
  if (snowT < 0) snowT = 0;
  double meltT = c->meltT > snowT * 1.05 ? c->meltT : snowT * 1.05;
  P.snowT = (float)snowT; P.meltT = (float)meltT; P.rainT = (float)rainT; P.cloudT = (float)c->cloudT;
  P.cloudP = (float)c->cloudP; P.rainLiftAge = (float)c->rainLiftAge;
  P.cloudBase = (float)(c->oceanR + 0.3 * c->atmosphereH);
  P.dt = (float)dt;
  double damp = 1 - 0.12 * dt;
  P.damp = (float)(dt > 0 ? (damp > 0 ? damp : 0) : 1);
  P.tick = (uint32_t)(int)(c->simTime * 60);
  P.spinOn = c->spinOn;
  unsigned *steam = (unsigned *)c->vort;
  int r;
  if ((r = zero_words(c, steam, c->nC))) return r;
  int nb = blocks_for(c->nP, 256);
  hipLaunchKernelGGL(k_steam_count, dim3(nb), dim3(256), 0, c->stream, c->geo, c->nP, c->px, c->py, c->pz, c->pflag,
                     c->thr, steam);
  hipLaunchKernelGGL(k_phase, dim3(nb), dim3(256), 0, c->stream, c->geo, P, c->nP, c->px, c->py, c->pz, c->pvx, c->pvy,
                     c->pvz, c->pT, c->pflag, c->pAir, c->pWx, c->pWy, c->pWz, c->pPh, c->cellType, steam, c->thr);
  WS_LAUNCHED();
  return WS_OK;
}

// _pushSurfaceParticles, solver.js:3486-3488
int frame_push_surface(ws_ctx *c) {
  if (c->nP < 1 || !c->haveTerrain) return WS_OK;
  float skyHi = (float)(c->oceanR + c->atmosphereH - 1e-3 * c->dx);
  float outer = (float)(c->domainR - c->dx * 0.02);
  hipLaunchKernelGGL(k_push_surface, dim3(blocks_for(c->nP, 256)), dim3(256), 0, c->stream, c->geo, make_world(c),
                     c->nP, (float)c->domainR, skyHi, outer, c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pflag);
  WS_LAUNCHED();
  return WS_OK;
}

}  // namespace ws
