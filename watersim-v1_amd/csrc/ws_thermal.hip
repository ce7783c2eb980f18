// The 25 Hz frame tick and frame-end passes of the planet (js/solver.js):
// sun exposure + radiative relaxation + conservative cell conduction +
// radial convection (_updateHeat, :2715-2994), evaporation with the
// reference's sequential LCG stream (_updateEvaporation, :3008-3125),
// curl-noise currents (_updateCurrents, :3658-3782) and the phase state
// machine (_updatePhaseChanges, :3152-3281).
#include "ws_kernels.h"

namespace ws {

__device__ __forceinline__ int trunc_cell(float x, const Geo &g, const float *thr, int n) {
  // (x / dx) | 0 then clamped to [0, n-1]; truncation equals floor for x >= 0
  return clampi(cell_floor(x, g.inv_dx, thr, n), 0, n - 1);
}

// ---------------------------------------------------------------- heat
// Exposure, pDepth and the per-flag radiative relaxation of one particle
// (_heatShadowChunk, solver.js:2773-2918), then the cell sums for the
// conduction (exact 64-bit fixed point, so the means do not depend on the
// order particles are summed in).
__global__ __launch_bounds__(256) void k_heat_particle(Geo g, ThermalPar T, int nP, const float *__restrict__ px,
                                                       const float *__restrict__ py, const float *__restrict__ pz,
                                                       float *__restrict__ pT, const uint8_t *__restrict__ pflag,
                                                       float *__restrict__ pLight, uint8_t *__restrict__ pDepth,
                                                       const uint8_t *__restrict__ ct, const float *__restrict__ dens,
                                                       const uint8_t *__restrict__ tsolid, const float *__restrict__ thr,
                                                       unsigned long long *__restrict__ csum,
                                                       unsigned *__restrict__ ccnt) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  float x = px[p], y = py[p], z = pz[p];
  const float dx = g.dx, cx = g.cx, cy = g.cy, cz = g.cz;
  // water cells stacked over the particle along the local up (solver.js:2800-2810)
  float dwx = x - cx, dwy = y - cy, dwz = z - cz;
  float dwl = sqrtf(dwx * dwx + dwy * dwy + dwz * dwz);
  if (!(dwl != 0.f)) dwl = 1e-9f;
  float stpU = dx / dwl;
  int above = 0;
  for (int st = 1; st <= 4; st++) {
    float qx = x + dwx * stpU * st, qy = y + dwy * stpU * st, qz = z + dwz * stpU * st;
    int qi = (int)(qx * g.inv_dx), qj = (int)(qy * g.inv_dx), qk = (int)(qz * g.inv_dx);  // "| 0"
    if (qi < 0 || qj < 0 || qk < 0 || qi >= g.nx || qj >= g.ny || qk >= g.nz) break;
    if (ct[(qk * g.ny + qj) * g.nx + qi] != FLUID) break;
    above++;
  }
  pDepth[p] = (uint8_t)above;
  float lit = T.hasSun ? 1.f : 0.f;
  if (T.hasSun) {
    if (above >= 4) lit = 0.f;
    else {
      float lx = T.sunX - x, ly = T.sunY - y, lz = T.sunZ - z;
      float ld = sqrtf(lx * lx + ly * ly + lz * lz);
      if (!(ld != 0.f)) ld = 1e-9f;
      float dxn = lx / ld, dyn = ly / ld, dzn = lz / ld;
      float ex = cx - x, ey = cy - y, ez = cz - z;
      float b = ex * dxn + ey * dyn + ez * dzn;
      if (T.hasTerrain) {
        float d2c = ex * ex + ey * ey + ez * ez - b * b;
        if (b > 0.f && d2c < T.Rlo2 * 0.96f) lit = 0.f;
        else if (b > 0.f || d2c < T.Rhi2 * 1.1f) {
          float stp = dx * 1.5f, mx = x, my = y, mz = z, esc2 = (T.Rhi + 0.1f) * (T.Rhi + 0.1f);
          for (int st2 = 0; st2 < 26; st2++) {
            mx += dxn * stp; my += dyn * stp; mz += dzn * stp;
            float rex = mx - cx, rey = my - cy, rez = mz - cz;
            if (rex * rex + rey * rey + rez * rez > esc2) break;
            int ivx = (int)(mx * T.inv_dv), ivy = (int)(my * T.inv_dv), ivz = (int)(mz * T.inv_dv);
            if (ivx < 0 || ivx >= T.tn || ivy < 0 || ivy >= T.tn || ivz < 0 || ivz >= T.tn) break;
            if (tsolid[((long)ivz * T.tn + ivy) * T.tn + ivx]) { lit = 0.f; break; }
          }
        }
      } else if (b > 0.f) {
        float d2c = ex * ex + ey * ey + ez * ez - b * b;
        if (d2c < T.coreR2) lit = 0.f;
      }
      if (lit > 0.f) {
        // optical depth through the density field toward the sun
        float od = 0.f, sx = x, sy = y, sz = z, half = dx;
        int sj = g.nx + 1, sk = (g.ny + 1) * sj;
        for (int st = 0; st < 6 && st * half < 0.55f; st++) {
          sx += dxn * half; sy += dyn * half; sz += dzn * half;
          int i0 = (int)(sx * g.inv_dx), j0 = (int)(sy * g.inv_dx), k0 = (int)(sz * g.inv_dx);
          if (i0 < 0 || j0 < 0 || k0 < 0 || i0 >= g.nx || j0 >= g.ny || k0 >= g.nz) break;
          od += dens[k0 * sk + j0 * sj + i0] * half * 1.1f;
          if (od > 3.5f) break;
        }
        lit = od < 3.5f ? expf(-od) : 0.f;
      }
    }
  }
  pLight[p] = lit;
  float Tp = pT[p], Tn;
  uint8_t fl = pflag[p];
  float sunPow = T.sunPow;
  if (fl == 2) {
    float amb = lit > 0.05f ? T.Tamb + 0.18f : 0.06f;
    float k = T.hasSun ? 0.16f + (1.f - lit) * 0.20f : 1.2f;
    Tn = Tp + (amb - Tp) * k * T.dt;
    if (lit > 0.f) Tn += sunPow * 0.6f * lit * T.dt;
  } else if (fl == 3) {
    float amb = lit > 0.05f ? T.Tamb + 0.18f : 0.06f;
    float k = T.hasSun ? 0.10f + (1.f - lit) * 0.12f : 0.9f;
    Tn = Tp + (amb - Tp) * k * T.dt;
    if (lit > 0.f) Tn += sunPow * 0.35f * lit * T.dt;
  } else if (fl == 4) {
    float amb = lit > 0.05f ? T.Tamb + 0.18f : 0.06f;
    float k = T.hasSun ? 0.14f + (1.f - lit) * 0.16f : 1.0f;
    Tn = Tp + (amb - Tp) * k * T.dt;
    if (lit > 0.f) Tn += sunPow * 0.45f * lit * T.dt;
  } else {
    float amb = lit > 0.05f ? T.Tamb : 0.f;
    Tn = Tp + (amb - Tp) * (0.045f + (1.f - lit) * T.shadeCool) * T.dt;
    if (lit > 0.f) Tn += sunPow * lit * T.dt;
  }
  Tn = fminf(fmaxf(Tn, 0.f), 1.15f);
  pT[p] = Tn;
  if (T.conduct) {
    int c = (trunc_cell(z, g, thr, g.nz) * g.ny + trunc_cell(y, g, thr, g.ny)) * g.nx + trunc_cell(x, g, thr, g.nx);
    atomicAdd(&csum[c], (unsigned long long)(long long)__builtin_rintf(Tn * kFixScaleF));
    atomicAdd(&ccnt[c], 1u);
  }
}

// cell means and the net conductive exchange of each cell with its six
// neighbours (_heatMeansSlice + _heatConductFlux, solver.js:2921-2964),
// gathered per cell in a fixed order
__global__ void k_heat_cells(Geo g, float alpha, const unsigned long long *__restrict__ csum,
                             const unsigned *__restrict__ ccnt, float *__restrict__ mean,
                             float *__restrict__ dT) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nC) return;
  unsigned n = ccnt[c];
  if (!n) { mean[c] = 0.f; dT[c] = 0.f; return; }
  int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny), nxy = g.nx * g.ny;
  float m = (float)(long long)csum[c] * kFixInvF / (float)n;
  float ch = 0.f;
  int nb[6] = {i < g.nx - 1 ? c + 1 : -1, j < g.ny - 1 ? c + g.nx : -1, k < g.nz - 1 ? c + nxy : -1,
               i > 0 ? c - 1 : -1, j > 0 ? c - g.nx : -1, k > 0 ? c - nxy : -1};
#pragma unroll
  for (int d = 0; d < 6; d++) {
    int o = nb[d];
    if (o < 0) continue;
    unsigned no = ccnt[o];
    if (!no) continue;
    float mo = (float)(long long)csum[o] * kFixInvF / (float)no;
    float e = alpha * (float)min(n, no) * (mo - m);  // flux into c from o
    ch += e;
  }
  mean[c] = m;
  dT[c] = ch / (float)n;
}

// _heatApplySlice (solver.js:2966-2979) and _heatConvectChunk (:2982-2994)
__global__ void k_heat_apply(Geo g, ThermalPar T, int nP, const float *__restrict__ px, const float *__restrict__ py,
                             const float *__restrict__ pz, float *__restrict__ pT, const uint8_t *__restrict__ pflag,
                             float *__restrict__ pvx, float *__restrict__ pvy, float *__restrict__ pvz,
                             const float *__restrict__ mean, const float *__restrict__ dT,
                             const float *__restrict__ thr) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  float x = px[p], y = py[p], z = pz[p], t = pT[p];
  if (T.conduct) {
    int c = (trunc_cell(z, g, thr, g.nz) * g.ny + trunc_cell(y, g, thr, g.ny)) * g.nx + trunc_cell(x, g, thr, g.nx);
    t = t + (T.alpha * (mean[c] - t) + dT[c]);
    pT[p] = t;
  }
  if (pflag[p] != 0) return;
  float rx = x - g.cx, ry = y - g.cy, rz = z - g.cz;
  float rl = sqrtf(rx * rx + ry * ry + rz * rz);
  if (!(rl != 0.f)) rl = 1e-9f;
  float ab = 3.0f * (t - T.Tamb) * T.dt / rl;
  pvx[p] = pvx[p] + rx * ab;
  pvy[p] = pvy[p] + ry * ab;
  pvz[p] = pvz[p] + rz * ab;
}

// ---------------------------------------------------------- evaporation
// Candidates of _evapChunk (solver.js:3041-3069): grid-coupled, above the
// freezing point, in a FLUID cell with no FLUID cell along the local up.
// Each candidate draws once from the _evS stream in particle order; the
// escape threshold is stored as the largest stream state that ejects. The
// freezing gate, the up cell and the probability are evaluated in double
// from the float fields, as the reference does: a draw ejects iff
// state / 2^32 <= prob, i.e. state <= floor(prob * 2^32), exact in double.
__global__ void k_evap_candidates(Geo g, EvapPar E, int nP, const float *__restrict__ px,
                                  const float *__restrict__ py, const float *__restrict__ pz,
                                  const float *__restrict__ pT, const uint8_t *__restrict__ pflag,
                                  const uint8_t *__restrict__ ct, const float *__restrict__ thr,
                                  unsigned *__restrict__ cand, unsigned *__restrict__ thresh) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  cand[p] = 0;
  if (pflag[p] != 0) return;
  const double T = pT[p];
  if (T <= E.freezeT) return;
  const float x = px[p], y = py[p], z = pz[p];
  int i = clampi(cell_floor(x, g.inv_dx, thr, g.nx), 1, g.nx - 2);
  int j = clampi(cell_floor(y, g.inv_dx, thr, g.ny), 1, g.ny - 2);
  int k = clampi(cell_floor(z, g.inv_dx, thr, g.nz), 1, g.nz - 2);
  if (ct[(k * g.ny + j) * g.nx + i] != FLUID) return;
  const double ex = (double)x - E.cx, ey = (double)y - E.cy, ez = (double)z - E.cz;
  double el = sqrt(ex * ex + ey * ey + ez * ez);
  if (!(el != 0.0)) el = 1e-9;
  // (x + e/|e| dx) / dx truncated toward zero, as `| 0` does
  const int ui = (int)(((double)x + ex / el * E.dx) / E.dx), uj = (int)(((double)y + ey / el * E.dx) / E.dx);
  const int uk = (int)(((double)z + ez / el * E.dx) / E.dx);
  if (ui < 0 || uj < 0 || uk < 0 || ui >= g.nx || uj >= g.ny || uk >= g.nz) return;
  if (ct[(uk * g.ny + uj) * g.nx + ui] == FLUID) return;
  double prob = E.rate * E.dt * exp(3.5 * (T - E.freezeT));
  if (prob > 1.0) prob = 1.0;
  const double lim = floor(prob * 4294967296.0);
  thresh[p] = lim >= 4294967295.0 ? 0xffffffffu : (unsigned)lim;
  cand[p] = 1;
}

__global__ void k_evap_compact(int nP, const unsigned *__restrict__ cand, const unsigned *__restrict__ off,
                               const unsigned *__restrict__ thresh, unsigned *__restrict__ list,
                               unsigned *__restrict__ thrC) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nP && cand[p]) {
    list[off[p]] = (unsigned)p;
    thrC[off[p]] = thresh[p];
  }
}

__device__ __forceinline__ uint32_t lcg1(uint32_t s) { return s * 1664525u + 1013904223u; }

// Walks the candidates in particle order through windows of EV_WIN. Without
// an ejection each candidate takes one draw; an ejection then takes its three
// Gaussian normals (two draws each, the first repeated while it is below
// 1e-9, solver.js:3120), almost always six draws. So if m candidates of a
// window before candidate j ejected, j draws state A_(j+1+6m) s + C_(j+1+6m)
// from the window's start state s. Every thread tests its candidates of the
// window under m = 0..EV_HYP-1 at once (bit rows per m in LDS); one
// wavefront then follows the real sequence: the first ejection under m = 0,
// the next one after it under m = 1, and so on. The window ends early after
// the EV_HYP-th ejection or after an ejection whose normals took other than
// six draws, and the next one starts behind it. Records the candidate
// index of each ejection (k_evap_apply maps it to the particle) with the
// stream state its Gaussian draws start from.
__global__ __launch_bounds__(EV_THREADS) void k_evap_chooser(const unsigned *__restrict__ thrC,
                                                             const unsigned *__restrict__ off,
                                                             const unsigned *__restrict__ cand, int nP,
                                                             const uint32_t *__restrict__ jmpA,
                                                             const uint32_t *__restrict__ jmpC,
                                                             uint32_t *__restrict__ rng,
                                                             unsigned *__restrict__ evList,
                                                             unsigned *__restrict__ evState,
                                                             unsigned *__restrict__ evCount) {
  constexpr int U = EV_WIN / EV_THREADS, WORDS = EV_WIN / 64;
  constexpr unsigned STAGE = 4 * EV_WIN;  // thresholds staged in LDS per global load
  static_assert(WORDS == 64, "one bit-row word per lane of the resolving wavefront");
  static_assert(EV_HYP <= 64, "one lane per ejection of a window");
  __shared__ unsigned long long rows[EV_HYP][WORDS];
  __shared__ uint32_t sA[EV_WIN], sC[EV_WIN];
  __shared__ unsigned sthr[STAGE];
  __shared__ uint32_t shS;
  __shared__ unsigned shK, shN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const unsigned nCand = nP > 0 ? off[nP - 1] + cand[nP - 1] : 0;
  for (int e = tid; e < EV_WIN; e += EV_THREADS) { sA[e] = jmpA[e]; sC[e] = jmpC[e]; }
  // the thread's candidates of a window: j = u * EV_THREADS + tid
  uint32_t myA[U], myC[U];
#pragma unroll
  for (int u = 0; u < U; u++) { myA[u] = jmpA[u * EV_THREADS + tid]; myC[u] = jmpC[u * EV_THREADS + tid]; }
  // maps of 1, 3, 5 and 6 draws
  const uint32_t A1 = jmpA[0], C1 = jmpC[0], A3 = jmpA[2], C3 = jmpC[2], A5 = jmpA[4], C5 = jmpC[4];
  const uint32_t A6 = jmpA[5], C6 = jmpC[5];
  if (tid == 0) { shS = rng[2]; shK = 0; shN = 0; }
  unsigned sb = 0, sEnd = 0;  // staged thresholds: candidates [sb, sEnd)
  __syncthreads();
  while (true) {
    const unsigned kw = shK;
    const uint32_t s = shS;
    if (kw >= nCand) break;  // uniform: every thread read the same kw after the barrier
    const unsigned nw = nCand - kw < (unsigned)EV_WIN ? nCand - kw : (unsigned)EV_WIN;
    if (kw + nw > sEnd) {  // uniform: restage from kw (no thread still reads the old stage)
      sb = kw;
      sEnd = nCand - kw < STAGE ? nCand : kw + STAGE;
      for (unsigned e = tid; e < sEnd - sb; e += EV_THREADS) sthr[e] = thrC[sb + e];
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const unsigned j = (unsigned)(u * EV_THREADS + tid);
      const bool valid = j < nw;
      const unsigned lim = valid ? sthr[kw - sb + j] : 0u;
      uint32_t st = myA[u] * s + myC[u];
#pragma unroll
      for (int m = 0; m < EV_HYP; m++) {
        unsigned long long b = __ballot(valid && st <= lim);
        if (lane == 0) rows[m][u * (EV_THREADS / 64) + wv] = b;
        st = A6 * st + C6;
      }
    }
    __syncthreads();
    if (wv == 0) {
      unsigned long long wr[EV_HYP];
#pragma unroll
      for (int m = 0; m < EV_HYP; m++) wr[m] = rows[m][lane];
      // the real sequence by bit scans alone: lane h keeps the candidate of
      // the window's ejection h
      unsigned pos = 0, myj = 0;
      int m = 0;
      for (int h = 0; h < EV_HYP; h++) {
        // first candidate at or after pos that ejects when h ejected before it
        unsigned long long w = wr[h];
        const unsigned base = (unsigned)lane * 64u;
        if (base + 64u <= pos) w = 0;
        else if (base < pos) w &= ~0ull << (pos - base);
        const unsigned long long any = __ballot(w != 0);
        if (!any) break;
        const int fl = __ffsll((long long)any) - 1;
        const unsigned long long wf =
            ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(w >> 32), fl) << 32) |
            (unsigned)__builtin_amdgcn_readlane((int)(unsigned)w, fl);
        const unsigned j = (unsigned)fl * 64u + (unsigned)(__ffsll((long long)wf) - 1);
        if (lane == h) myj = j;
        pos = j + 1u;
        m = h + 1;
      }
      // each ejection's state (j + 1 + 6h draws from s) and its normals, one
      // lane each: six draws unless a first draw is <= 4 (u1 = g / 2^32 <
      // 1e-9 is drawn again)
      const bool mine = lane < m;
      uint32_t se = 0, gs = 0;
      bool six = true;
      if (mine) {
        se = sA[myj] * s + sC[myj];
        if (lane > 0) se = sA[6 * lane - 1] * se + sC[6 * lane - 1];
        six = A1 * se + C1 > 4u && A3 * se + C3 > 4u && A5 * se + C5 > 4u;
        if (six) {
          gs = A6 * se + C6;
        } else {
          gs = se;
          for (int n = 0; n < 3; n++) {
            do { gs = lcg1(gs); } while (gs <= 4u);
            gs = lcg1(gs);
          }
        }
      }
      // the window ends behind the first ejection with other than six draws
      // (later ones assumed six), or behind the EV_HYP-th
      const unsigned long long bad = __ballot(mine && !six);
      int keep = m;
      bool cut = m == EV_HYP;
      if (bad) { keep = __ffsll((long long)bad); cut = true; }
      uint32_t next;
      unsigned end;
      if (cut) {
        next = (uint32_t)__builtin_amdgcn_readlane((int)gs, keep - 1);
        end = (unsigned)__builtin_amdgcn_readlane((int)myj, keep - 1) + 1u;
      } else {  // all nw candidates with m ejections among them
        next = sA[nw - 1] * s + sC[nw - 1];
        if (m > 0) next = sA[6 * m - 1] * next + sC[6 * m - 1];
        end = nw;
      }
      const unsigned nev = shN;
      if (lane < keep) { evList[nev + lane] = kw + myj; evState[nev + lane] = se; }
      if (lane == 0) { shS = next; shK = kw + end; shN = nev + (unsigned)keep; }
    }
    __syncthreads();
  }
  if (tid == 0) { rng[2] = shS; *evCount = shN; }
}

// Maxwell-Boltzmann kick, radial mirror and no-slip tangential adoption of
// the ejected particles (solver.js:3083-3105)
__global__ void k_evap_apply(Geo g, EvapPar E, const unsigned *__restrict__ list, const unsigned *__restrict__ evList,
                             const unsigned *__restrict__ evState,
                             const unsigned *__restrict__ evCount, float *__restrict__ px, float *__restrict__ py,
                             float *__restrict__ pz, float *__restrict__ pvx, float *__restrict__ pvy,
                             float *__restrict__ pvz, const float *__restrict__ pT, uint8_t *__restrict__ pflag,
                             float *__restrict__ pAir, const float *__restrict__ vel, const float *__restrict__ thr) {
  unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= *evCount) return;
  unsigned p = list[evList[e]];
  uint32_t gst = evState[e];
  float nrm[3];
  for (int n = 0; n < 3; n++) {
    do { gst = lcg1(gst); } while (gst <= 4u);
    float u1 = (float)gst * 2.3283064365386963e-10f;
    gst = lcg1(gst);
    float u2 = (float)gst * 2.3283064365386963e-10f;
    nrm[n] = sqrtf(-2.f * logf(u1)) * cosf(6.283185306f * u2);
  }
  float T = pT[p];
  const float fT = (float)E.freezeT;
  float tau = fminf(fmaxf((T - fT) / (1.15f - fT), 0.f), 1.f);
  float sig = 0.30f + 0.32f * sqrtf(tau);
  float kx = sig * nrm[0], ky = sig * nrm[1], kz = sig * nrm[2];
  pflag[p] = 2;
  pAir[p] = 0.f;
  float x = px[p], y = py[p], z = pz[p];
  float ex = x - g.cx, ey = y - g.cy, ez = z - g.cz;
  float el2 = ex * ex + ey * ey + ez * ez;
  float elL = sqrtf(el2);
  if (!(elL != 0.f)) elL = 1e-9f;
  float kUp = (kx * ex + ky * ey + kz * ez) / el2;
  if (kUp < 0.f) { kx -= 2.f * kUp * ex / elL; ky -= 2.f * kUp * ey / elL; kz -= 2.f * kUp * ez / elL; }
  int i = clampi(cell_floor(x, g.inv_dx, thr, g.nx), 1, g.nx - 2);
  int j = clampi(cell_floor(y, g.inv_dx, thr, g.ny), 1, g.ny - 2);
  int k = clampi(cell_floor(z, g.inv_dx, thr, g.nz), 1, g.nz - 2);
  float guS = vel[(k * g.ny + j) * (g.nx + 1) + i];
  float gvS = vel[g.uN + (k * (g.ny + 1) + j) * g.nx + i];
  float gwS = vel[g.uN + g.vN + (k * g.ny + j) * g.nx + i];
  float nl2 = 1.f / el2, vN = (pvx[p] * ex + pvy[p] * ey + pvz[p] * ez) * nl2;
  float gN = (guS * ex + gvS * ey + gwS * ez) * nl2;
  pvx[p] = vN * ex + (guS - gN * ex) + kx;
  pvy[p] = vN * ey + (gvS - gN * ey) + ky;
  pvz[p] = vN * ez + (gwS - gN * ez) + kz;
}

// --------------------------------------------------------------- currents
// _buildCurrentField, solver.js:3682-3722: curl of three octaves of
// sine-wave potentials on a stride-4 lattice
__global__ void k_currents_lattice(CurrPar Q, float *__restrict__ cf) {
  int gi = blockIdx.x * blockDim.x + threadIdx.x;
  int gN = Q.gnx * Q.gny * Q.gnz;
  if (gi >= gN) return;
  int ii = gi % Q.gnx, jj = (gi / Q.gnx) % Q.gny, kk = gi / (Q.gnx * Q.gny);
  float wx2 = ii * Q.sx, wy2 = jj * Q.sx, wz2 = kk * Q.sx;
  float ux = 0.f, uy = 0.f, uz = 0.f;
  for (int o = 0; o < 3; o++) {
    float dph = Q.drift[o] * Q.t;
    float th[3];
    for (int c = 0; c < 3; c++) {
      const float *w = Q.wave[o][c];
      th[c] = w[0] * wx2 + w[1] * wy2 + w[2] * wz2 + dph + w[4];
    }
    const float *w0 = Q.wave[o][0], *w1 = Q.wave[o][1], *w2 = Q.wave[o][2];
    ux += w2[3] * w2[1] * cosf(th[2]) - w1[3] * w1[2] * cosf(th[1]);
    uy += w0[3] * w0[2] * cosf(th[0]) - w2[3] * w2[0] * cosf(th[2]);
    uz += w1[3] * w1[0] * cosf(th[1]) - w0[3] * w0[1] * cosf(th[0]);
  }
  cf[gi] = ux * Q.amp;
  cf[gN + gi] = uy * Q.amp;
  cf[2 * gN + gi] = uz * Q.amp;
}

// _currentsChunk, solver.js:3726-3782: deep particles relax toward the
// tangential part of the lattice current
__global__ void k_currents_relax(Geo g, CurrPar Q, int nP, const float *__restrict__ px, const float *__restrict__ py,
                                 const float *__restrict__ pz, float *__restrict__ pvx, float *__restrict__ pvy,
                                 float *__restrict__ pvz, const uint8_t *__restrict__ pflag,
                                 const float *__restrict__ cf) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP || pflag[p] != 0) return;
  float x = px[p], y = py[p], z = pz[p];
  float rx = x - g.cx, ry = y - g.cy, rz = z - g.cz;
  float rl = sqrtf(rx * rx + ry * ry + rz * rz);
  if (!(rl != 0.f)) rl = 1e-9f;
  float dpt = (Q.oceanR - rl) / Q.shell;
  float wgt = (dpt - 0.06f) * 4.0f;
  if (wgt <= 0.f) return;
  if (wgt > 1.f) wgt = 1.f;
  float qa = x * Q.invS, qb = y * Q.invS, qc = z * Q.invS;
  int i0 = (int)qa, j0 = (int)qb, k0 = (int)qc;
  if (qa < 0.f) i0 = 0;
  if (qb < 0.f) j0 = 0;
  if (qc < 0.f) k0 = 0;
  if (i0 > Q.gnx - 2) i0 = Q.gnx - 2; else if (i0 < 0) i0 = 0;
  if (j0 > Q.gny - 2) j0 = Q.gny - 2; else if (j0 < 0) j0 = 0;
  if (k0 > Q.kMax) k0 = Q.kMax; else if (k0 < 0) k0 = 0;
  float ta = fminf(fmaxf(qa - i0, 0.f), 1.f), tb = fminf(fmaxf(qb - j0, 0.f), 1.f), tc = fminf(fmaxf(qc - k0, 0.f), 1.f);
  int gN = Q.gnx * Q.gny * Q.gnz, gxy = Q.gnx * Q.gny;
  int g000 = (k0 * Q.gny + j0) * Q.gnx + i0;
  float a0 = 1.f - ta, a1 = ta, b0 = 1.f - tb, b1 = tb, c0 = 1.f - tc, c1 = tc;
  float U[3];
  for (int d = 0; d < 3; d++) {
    const float *f = cf + (long)d * gN;
    U[d] = (f[g000] * a0 + f[g000 + 1] * a1) * b0 * c0 + (f[g000 + Q.gnx] * a0 + f[g000 + Q.gnx + 1] * a1) * b1 * c0 +
           (f[g000 + gxy] * a0 + f[g000 + gxy + 1] * a1) * b0 * c1 +
           (f[g000 + gxy + Q.gnx] * a0 + f[g000 + gxy + Q.gnx + 1] * a1) * b1 * c1;
  }
  float ir = 1.f / rl;
  float Ur = (U[0] * rx + U[1] * ry + U[2] * rz) * ir;
  float Tx = (U[0] - rx * ir * Ur) * wgt, Ty = (U[1] - ry * ir * Ur) * wgt, Tz = (U[2] - rz * ir * Ur) * wgt;
  float vx = pvx[p], vy = pvy[p], vz = pvz[p];
  float vr = (vx * rx + vy * ry + vz * rz) * ir;
  float tvx = vx - rx * ir * vr, tvy = vy - ry * ir * vr, tvz = vz - rz * ir * vr;
  float ex = Q.relax * (Tx - tvx), ey = Q.relax * (Ty - tvy), ez = Q.relax * (Tz - tvz);
  float e2 = ex * ex + ey * ey + ez * ez;
  if (e2 > 0.15f * 0.15f) { float es = 0.15f / sqrtf(e2); ex *= es; ey *= es; ez *= es; }
  pvx[p] = vx + ex;
  pvy[p] = vy + ey;
  pvz[p] = vz + ez;
}

// ----------------------------------------------------------- phase changes
__global__ void k_steam_count(Geo g, int nP, const float *__restrict__ px, const float *__restrict__ py,
                              const float *__restrict__ pz, const uint8_t *__restrict__ pflag,
                              const float *__restrict__ thr, unsigned *__restrict__ cnt) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP || pflag[p] != 2) return;
  int c = (trunc_cell(pz[p], g, thr, g.nz) * g.ny + trunc_cell(py[p], g, thr, g.ny)) * g.nx +
          trunc_cell(px[p], g, thr, g.nx);
  atomicAdd(&cnt[c], 1u);
}

// (a * b) >>> 0 in JavaScript: the product is a double, rounded to 53 bits
// when it exceeds 2^53, then taken modulo 2^32
__device__ __forceinline__ uint32_t js_mul_u32(uint32_t a, uint32_t b) {
  unsigned long long p = (unsigned long long)a * b;
  if (p >> 53) {
    int sh = (64 - __clzll((long long)p)) - 53;
    unsigned long long mask = (1ull << sh) - 1ull, half = 1ull << (sh - 1), rem = p & mask;
    p >>= sh;
    if (rem > half || (rem == half && (p & 1ull))) p++;
    p <<= sh;
  }
  return (uint32_t)p;
}

// _spinHash, solver.js:3145-3150
__device__ __forceinline__ uint32_t spin_hash(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t s = a ^ js_mul_u32(b, 0x9e3779b9u) ^ js_mul_u32(c, 0x85ebca6bu) ^ js_mul_u32(0x51ab3c77u, 0xc2b2ae35u);
  s = s * 1664525u + 1013904223u;
  s ^= s >> 13;
  s = js_mul_u32(s, 0x5bd1e995u);
  s ^= s >> 15;
  return s;
}

// _updatePhaseChanges, solver.js:3152-3281
__global__ void k_phase(Geo g, PhasePar P, int nP, const float *__restrict__ px, const float *__restrict__ py,
                        const float *__restrict__ pz, float *__restrict__ pvx, float *__restrict__ pvy,
                        float *__restrict__ pvz, const float *__restrict__ pT, uint8_t *__restrict__ pflag,
                        float *__restrict__ pAir, float *__restrict__ pWx, float *__restrict__ pWy,
                        float *__restrict__ pWz, float *__restrict__ pPh, const uint8_t *__restrict__ ct,
                        const unsigned *__restrict__ steam, const float *__restrict__ thr) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  uint8_t f = pflag[p];
  float T = pT[p];
  float x = px[p], y = py[p], z = pz[p];
  int i = trunc_cell(x, g, thr, g.nx), j = trunc_cell(y, g, thr, g.ny), k = trunc_cell(z, g, thr, g.nz);
  int cB = (k * g.ny + j) * g.nx + i, nxy = g.nx * g.ny;
  if (f == 0) {
    if (T <= P.snowT && T > 1e-3f) {
      bool atSurf = (k > 0 && ct[cB - nxy] != FLUID) || (k < g.nz - 1 && ct[cB + nxy] != FLUID) ||
                    (j > 0 && ct[cB - g.nx] != FLUID) || (j < g.ny - 1 && ct[cB + g.nx] != FLUID) ||
                    (i > 0 && ct[cB - 1] != FLUID) || (i < g.nx - 1 && ct[cB + 1] != FLUID);
      if (atSurf) {
        pflag[p] = 5;
        pvx[p] = pvy[p] = pvz[p] = 0.f;
        pAir[p] = 1.f;
      }
    }
    return;
  }
  if (f == 5) {
    if (T >= P.meltT) {
      pflag[p] = ct[cB] == FLUID ? 0 : 4;
      pAir[p] = 0.f;
      pWx[p] = pWy[p] = pWz[p] = 0.f;
    }
    return;
  }
  if (T > 1e-3f) {
    if (f == 2) {
      if (T <= P.rainT && pAir[p] >= P.rainLiftAge) {
        f = 4;
        pflag[p] = 4;
        pAir[p] = 0.f;
        pvx[p] = pvy[p] = pvz[p] = 0.f;
        pWx[p] = pWy[p] = pWz[p] = 0.f;
      } else if (T <= P.cloudT) {
        float rr = sqrtf((x - g.cx) * (x - g.cx) + (y - g.cy) * (y - g.cy) + (z - g.cz) * (z - g.cz));
        if (rr >= P.cloudBase) {
          int i0 = i > 0 ? i - 1 : 0, i1 = i < g.nx - 1 ? i + 1 : g.nx - 1;
          int j0 = j > 0 ? j - 1 : 0, j1 = j < g.ny - 1 ? j + 1 : g.ny - 1;
          int k0 = k > 0 ? k - 1 : 0, k1 = k < g.nz - 1 ? k + 1 : g.nz - 1;
          unsigned sum = 0;
          for (int kk = k0; kk <= k1; kk++)
            for (int jj = j0; jj <= j1; jj++)
              for (int ii = i0; ii <= i1; ii++) sum += steam[(kk * g.ny + jj) * g.nx + ii];
          if ((float)sum / 27.f >= P.cloudP) { f = 3; pflag[p] = 3; pAir[p] = 0.f; }
        }
      }
    } else if (f == 3) {
      if (T >= P.cloudT + 0.05f) { f = 2; pflag[p] = 2; }
    } else if (f == 4) {
      if (T <= P.snowT) {
        pflag[p] = 5;
        pvx[p] = pvy[p] = pvz[p] = 0.f;
        pAir[p] = 0.f;
        pWx[p] = pWy[p] = pWz[p] = 0.f;
        return;
      }
    }
  }
  // spin bookkeeping of the airborne family (solver.js:3259-3280); the flag
  // here is the one after this pass, as in the reference's second loop
  if (P.spinOn && P.dt > 0.f && (f == 2 || f == 3)) {
    float wx = pWx[p], wy = pWy[p], wz = pWz[p];
    if (f == 2 && pAir[p] <= 1e-6f && wx == 0.f && wy == 0.f && wz == 0.f) {
      float h1 = (float)spin_hash(0x51ab3c77u, (uint32_t)p, P.tick) * 2.3283064365386963e-10f;
      float h2 = (float)spin_hash(0x51ab3c78u, (uint32_t)p, P.tick) * 2.3283064365386963e-10f;
      float h3 = (float)spin_hash(0x51ab3c79u, (uint32_t)p, P.tick) * 2.3283064365386963e-10f;
      float zz = 2.f * h2 - 1.f, rr = sqrtf(fmaxf(0.f, 1.f - zz * zz)), th = 6.283185306f * h1;
      float Tp = T;
      if (!(Tp > 0.05f)) Tp = 0.05f; else if (Tp > 1.15f) Tp = 1.15f;
      float wSpd = 3.0f * sqrtf(Tp) * (0.5f + h3);
      wx = rr * cosf(th) * wSpd; wy = zz * wSpd; wz = rr * sinf(th) * wSpd;
    }
    float wm = sqrtf(wx * wx + wy * wy + wz * wz);
    if (wm > 1e-6f) pPh[p] = pPh[p] + wm * P.dt;
    pWx[p] = wx * P.damp; pWy[p] = wy * P.damp; pWz[p] = wz * P.damp;
  }
}

// ------------------------------------------------------- surface rescue
// _pushSurfaceChunk, solver.js:3496-3587: airborne water capped under the
// atmosphere top, particles in rasterized rock walked out radially and
// parked on the smooth surface, grid-coupled water buried under the surface
// lifted back, fluid kept inside the domain; the inward velocity of a
// rescued particle is removed. Per particle, float arithmetic.
__global__ void k_push_surface(Geo g, World w, int nP, float domainR, float skyHi, float outer,
                               float *__restrict__ px, float *__restrict__ py, float *__restrict__ pz,
                               float *__restrict__ pvx, float *__restrict__ pvy, float *__restrict__ pvz,
                               const uint8_t *__restrict__ pflag) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nP) return;
  const float dx = g.dx, cx = g.cx, cy = g.cy, cz = g.cz;
  float x = px[p], y = py[p], z = pz[p];
  float ex = x - cx, ey = y - cy, ez = z - cz;
  float er = sqrtf(ex * ex + ey * ey + ez * ez);
  if (!(er != 0.f)) er = 1e-9f;
  bool fixed = false, rockFixed = false;
  uint8_t fl = pflag[p];
  if (er > skyHi) {
    x = cx + ex / er * skyHi; y = cy + ey / er * skyHi; z = cz + ez / er * skyHi;
    float vx = pvx[p], vy = pvy[p], vz = pvz[p];
    float outward = (vx * ex + vy * ey + vz * ez) / er;
    if (outward > 0.f) {
      pvx[p] = vx - ex / er * outward; pvy[p] = vy - ey / er * outward; pvz[p] = vz - ez / er * outward;
    }
    px[p] = x; py[p] = y; pz[p] = z;
    ex = x - cx; ey = y - cy; ez = z - cz; er = skyHi;
    fixed = true;
  }
  if (er <= domainR && rock_cell(w, g, x, y, z)) {
    rockFixed = true;
    float sx = ex / er * dx * 0.5f, sy = ey / er * dx * 0.5f, sz = ez / er * dx * 0.5f;
    for (int tries = 0; tries < 2 * g.nx && rock_cell(w, g, x, y, z); tries++) { x += sx; y += sy; z += sz; }
    float exR = x - cx, eyR = y - cy, ezR = z - cz;
    float erR = sqrtf(exR * exR + eyR * eyR + ezR * ezR);
    if (!(erR != 0.f)) erR = 1e-9f;
    float wantR = terrain_radius(w, x, y, z) + dx * 0.06f;
    if (erR > wantR) { x = cx + exR / erR * wantR; y = cy + eyR / erR * wantR; z = cz + ezR / erR * wantR; }
    float tx = exR / erR * dx * 0.5f, ty = eyR / erR * dx * 0.5f, tz = ezR / erR * dx * 0.5f;
    for (int kR = 0; kR < 4 && rock_cell(w, g, x, y, z); kR++) { x += tx; y += ty; z += tz; }
    fixed = true;
  } else if (fl == 0 && er <= domainR) {
    float Rt = terrain_radius(w, x, y, z);
    if (er < Rt - dx * 0.5f) {
      rockFixed = true;
      float lift = Rt + dx * 0.05f;
      x = cx + ex / er * lift; y = cy + ey / er * lift; z = cz + ez / er * lift;
      float sx = ex / er * dx * 0.5f, sy = ey / er * dx * 0.5f, sz = ez / er * dx * 0.5f;
      for (int t2 = 0; t2 < 8 && rock_cell(w, g, x, y, z); t2++) { x += sx; y += sy; z += sz; }
      fixed = true;
    }
  }
  float fx = x - cx, fy = y - cy, fz = z - cz;
  float finalR = sqrtf(fx * fx + fy * fy + fz * fz);
  float capR = fl == 0 ? outer : skyHi;
  if (finalR > capR) {
    x = cx + fx * capR / finalR; y = cy + fy * capR / finalR; z = cz + fz * capR / finalR;
    fixed = true;
  }
  if (fixed) {
    px[p] = x; py[p] = y; pz[p] = z;
    float vx = pvx[p], vy = pvy[p], vz = pvz[p];
    float vr = (vx * ex + vy * ey + vz * ez) / er;
    if (rockFixed && vr < 0.f) { pvx[p] = vx - ex / er * vr; pvy[p] = vy - ey / er * vr; pvz[p] = vz - ez / er * vr; }
  }
}

}  // namespace ws
