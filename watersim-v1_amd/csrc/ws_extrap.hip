// Velocity extrapolation (solver.js:1562-1636) as one pipeline over rows.
//
// The reference visits the faces of each component in index order (i, then
// j, then k): an invalid face with at least one valid neighbour takes the
// mean of the valid ones and counts as valid from then on. Its -i/-j/-k
// neighbours have been visited (their final state), its +i/+j/+k
// neighbours not (their original state).
//
// Flags. With seed(f) = orig(f) | orig(+i) | orig(+j) | orig(+k), the final
// flag is final(f) = seed(f) | final(-i) | final(-j) | final(-k), a prefix OR
// of the seeds over the lower octant. k_exf_rows computes, per row (j, k),
// the seeds as bit masks (64 faces a word) and their prefix OR along i, then
// the prefix OR along j of each plane; k_exf_k the prefix OR along k. The
// result (one bit per face) is exact.
//
// Values. A filled face is x = A x(-i) + Pj x(-j) + Pk x(-k) + Q: r = 1/cnt,
// A, Pj, Pk = r where that neighbour is valid, Q = r (the sum of its valid
// +i, +j, +k neighbours' values, in the reference's order). Every
// coefficient follows from the flags and the original values. Along a row
// the -i dependence is a scan of the affine maps x -> A x + B across the 64
// lanes of a wavefront (DPP row shifts and broadcasts), the multiplicative
// half of which is value-free.
//
// Pipeline. A workgroup takes one job (component, slab of 8 k-planes, chunk
// of 64 faces along i) and walks all j rows of it: at step s, wavefront lk
// (plane k0 + lk) relaxes row j = s - lk, whose -j row it did itself one step
// earlier (a register) and whose -k row the wavefront below did one step
// earlier (LDS). The slab below hands its top plane's rows, and the chunk
// before each wavefront's last face, over as 64-bit words (value bits, tag)
// in exchange arrays. Each word is written once per launch, so the tag is
// the launch's epoch and a consumer polls the word itself until the tag
// matches. The jobs of a component run on one XCD, the one whose id equals
// the component: values written by one workgroup are seen by another's
// agent-scope loads through that XCD's shared L2 (one hop about 0.5 us), and
// the critical path is about d1 + 10 (d2 / 8) row steps instead of a chain
// of launches over tile diagonals.
//
// Scheduling and progress. Blocks dealt to XCDs 3..7 return at once. A
// block on XCD c takes jobs of component c from a ticket counter in an order
// that puts a job's -k and -i producers before it, so a job only waits for
// jobs that a running workgroup has already taken. Polls are bounded; a poll
// that gives up, or a component whose jobs were not all done (no workgroup
// ran on its XCD), sets DevScalars::exFailed, which the host reports.
//
// Status: switched off (the context refuses exRows 1). Its first GPU run, a
// 64^3 substep test, went silent for three minutes and was killed. Reading
// the code found no unbounded wait (every poll gives up after 2^16 tries,
// every barrier is reached by all wavefronts, tickets order each job after
// its producers), so no cause was shown and the path stays off until one
// is. The tile-diagonal extrapolation (k_extrap_diag) runs instead.
#include "ws_kernels.h"

namespace ws {

namespace {

constexpr int kExPolls = 1 << 16;

__device__ __forceinline__ unsigned long long ex_ld(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long ex_tagged(float v, unsigned tag) {
  return (unsigned long long)__float_as_uint(v) | (unsigned long long)tag << 32;
}

// inclusive OR scan across the 64 lanes
__device__ __forceinline__ unsigned long long or_scan64(unsigned long long v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long u = __shfl_up(v, o, 64);
    v |= lane >= o ? u : 0ull;
  }
  return v;
}

// 1/c, correctly rounded, for the neighbour counts 1..6
__device__ __forceinline__ float ex_rcp(int c) {
  float r = 1.f;
  r = c == 2 ? 0.5f : r;
  r = c == 3 ? 1.f / 3.f : r;
  r = c == 4 ? 0.25f : r;
  r = c == 5 ? 0.2f : r;
  r = c == 6 ? 1.f / 6.f : r;
  return r;
}

}  // namespace

// lane-shifted copy of v (DPP): lanes with no source lane get `old`
#define EXR_DPP(v, CTRL, RMASK, OLD) \
  __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(OLD), __float_as_int(v), CTRL, RMASK, 0xf, false))
// B += a * (B of the source lane), one DPP FMA; a lane with no source lane,
// or in a row the mask leaves out, keeps B. The s_nop gives the two wait
// states a DPP read needs after B's last VALU write.
#define EXR_FMAC_DPP(B, a, CTRL) \
  asm("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 " CTRL " bank_mask:0xf" : "+v"(B) : "v"(B), "v"(a))

// the three components' dimensions and offsets
__device__ __forceinline__ void ex_dims(const Geo &g, int comp, int &d0, int &d1, int &d2, long &base) {
  d0 = comp == 0 ? g.nx + 1 : g.nx;
  d1 = comp == 1 ? g.ny + 1 : g.ny;
  d2 = comp == 2 ? g.nz + 1 : g.nz;
  base = comp == 0 ? 0 : (comp == 1 ? g.uN : (long)g.uN + g.vN);
}

// per row: seed bits and their prefix OR along i; then, per plane, the
// prefix OR along j. One block (256 threads) per (component, plane k);
// fin[((comp * D2 + k) * D1 + j) * W + w] with D1 = ny + 1, D2 = nz + 1,
// W = ceil((nx + 1) / 64). Block 0 also clears the value pass's tickets.
__global__ __launch_bounds__(256) void k_exf_rows(Geo g, const uint8_t *__restrict__ valid,
                                                  unsigned long long *__restrict__ fin, ExRun R) {
  const int D1 = g.ny + 1, D2 = g.nz + 1, W = (g.nx + 1 + 63) / 64;
  const int comp = blockIdx.x / D2, k = blockIdx.x % D2;
  if (blockIdx.x == 0 && threadIdx.x < 3) {
    R.ticket[threadIdx.x] = 0u;
    R.done[threadIdx.x] = 0u;
  }
  int d0, d1, d2;
  long base;
  ex_dims(g, comp, d0, d1, d2, base);
  if (k >= d2) return;
  const uint8_t *V = valid + base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the plane's row words in LDS where they fit (grids up to 128^3), else in
  // fin itself
  constexpr int kRowsLds = 129 * 3;
  __shared__ unsigned long long rows[kRowsLds];
  const bool inLds = d1 * W <= kRowsLds;
  for (int j = wave; j < d1; j += 4) {
    unsigned long long carry = 0;
    for (int w = 0; w < W; w++) {
      const int i = w * 64 + lane;
      bool s = false;
      if (i < d0) {
        const long f = ((long)k * d1 + j) * d0 + i;
        s = V[f] || (i + 1 < d0 && V[f + 1]) || (j + 1 < d1 && V[f + d0]) || (k + 1 < d2 && V[f + (long)d0 * d1]);
      }
      const unsigned long long m = __ballot(s);
      const int nin = d0 - w * 64;
      const unsigned long long inMask = nin >= 64 ? ~0ull : (nin > 0 ? (1ull << nin) - 1ull : 0ull);
      const unsigned long long p = (carry ? ~0ull : (m ? ~0ull << __builtin_ctzll(m) : 0ull)) & inMask;
      carry |= m;
      if (lane == 0) {
        if (inLds) rows[j * W + w] = p;
        else fin[(((long)comp * D2 + k) * D1 + j) * W + w] = p;
      }
    }
  }
  __syncthreads();
  // prefix OR along j: wavefront w scans word w of every row, 64 rows at a time
  for (int w = wave; w < W; w += 4) {
    unsigned long long carry = 0;
    for (int j0 = 0; j0 < d1; j0 += 64) {
      const int j = j0 + lane;
      unsigned long long v = 0;
      if (j < d1) v = inLds ? rows[j * W + w] : fin[(((long)comp * D2 + k) * D1 + j) * W + w];
      v = or_scan64(v) | carry;
      carry = __shfl(v, 63, 64);
      if (j < d1) fin[(((long)comp * D2 + k) * D1 + j) * W + w] = v;
    }
  }
}

// the prefix OR along k: one wavefront per (component, row j, word w), 64
// planes at a time
__global__ __launch_bounds__(256) void k_exf_k(Geo g, unsigned long long *__restrict__ fin) {
  const int D1 = g.ny + 1, D2 = g.nz + 1, W = (g.nx + 1 + 63) / 64;
  const int id = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (id >= 3 * D1 * W) return;
  const int comp = id / (D1 * W), j = (id / W) % D1, w = id % W;
  int d0, d1, d2;
  long base;
  ex_dims(g, comp, d0, d1, d2, base);
  if (j >= d1 || w * 64 >= d0) return;
  unsigned long long carry = 0;
  for (int k0 = 0; k0 < d2; k0 += 64) {
    const int k = k0 + lane;
    const long at = (((long)comp * D2 + k) * D1 + j) * W + w;
    unsigned long long v = k < d2 ? fin[at] : 0ull;
    v = or_scan64(v) | carry;
    carry = __shfl(v, 63, 64);
    if (k < d2) fin[at] = v;
  }
}

// job t of a component with S slabs and C chunks: jobs ordered by slab +
// chunk, then by slab
__device__ __forceinline__ void ex_job(int t, int S, int C, int &m, int &I) {
  for (int dsum = 0;; dsum++) {
    const int lo = max(0, dsum - (C - 1)), hi = min(S - 1, dsum);
    const int n = hi >= lo ? hi - lo + 1 : 0;
    if (t < n) {
      m = lo + t;
      I = dsum - m;
      return;
    }
    t -= n;
  }
}

__global__ __launch_bounds__(EXR_THREADS) void k_extrap_rows(Geo g, float *__restrict__ vel,
                                                             uint8_t *__restrict__ valid,
                                                             const unsigned long long *__restrict__ fin, ExRun R,
                                                             DevScalars *ds) {
  const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);  // HW_REG_XCC_ID, bits 3:0
  if (xcc >= 3) return;
  const int comp = (int)xcc;
  int d0, d1, d2;
  long base;
  ex_dims(g, comp, d0, d1, d2, base);
  const int D1 = g.ny + 1, D2 = g.nz + 1, W = (g.nx + 1 + 63) / 64;
  const int S = (d2 + EXR_TK - 1) / EXR_TK, C = (d0 + 63) / 64, jobs = S * C;
  float *Av = vel + base;
  uint8_t *Vf = valid + base;
  const unsigned E = R.epoch;
  __shared__ float xfin[EXR_TK][2][64];  // a finished row, for the wavefront above
  __shared__ int sjob, sok;
  const int lane = threadIdx.x & 63;
  const int lk = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // this wavefront's plane in the slab
  for (;;) {
    if (threadIdx.x == 0) {
      sjob = (int)atomicAdd(&R.ticket[comp], 1u);
      sok = 1;
    }
    __syncthreads();
    const int t = sjob;
    if (t >= jobs) return;
    int m, I;
    ex_job(t, S, C, m, I);
    const int k = m * EXR_TK + lk, i0 = I * 64, i = i0 + lane;
    const bool kin = k < d2, in = kin && i < d0;
    // exchange slots: the -k rows this slab's wave 0 reads, the +k rows its
    // top wave writes, and the -i / +i faces of each wavefront's rows
    unsigned long long *xkIn = R.xk + (((long)comp * R.S + (m > 0 ? m - 1 : 0)) * R.C + I) * (long)D1 * 64;
    unsigned long long *xkOut = R.xk + (((long)comp * R.S + m) * R.C + I) * (long)D1 * 64;
    unsigned long long *xiIn = R.xi + ((((long)comp * R.S + m) * R.C + (I > 0 ? I - 1 : 0)) * EXR_TK + lk) * D1;
    unsigned long long *xiOut = R.xi + ((((long)comp * R.S + m) * R.C + I) * EXR_TK + lk) * D1;
    const unsigned long long *F = fin + (long)comp * D2 * D1 * W;
    // a row's inputs: final flags (this row, the row below, the face before
    // the chunk) and the original flags and values of the face and its +i,
    // +j, +k neighbours (not yet visited). Loaded one row ahead.
    struct RowIn {
      unsigned long long Fr, Fk;
      int cin, o0, ob, od, of;
      float v0, vb, vd, vf;
    };
    auto load_row = [&](int j) {
      RowIn q;
      q.Fr = q.Fk = 0ull;
      q.cin = q.o0 = q.ob = q.od = q.of = 0;
      q.v0 = q.vb = q.vd = q.vf = 0.f;
      if (kin && j >= 0 && j < d1) {
        q.Fr = F[((long)k * D1 + j) * W + I];
        q.Fk = k > 0 ? F[((long)(k - 1) * D1 + j) * W + I] : 0ull;
        q.cin = I > 0 ? (int)((F[((long)k * D1 + j) * W + I - 1] >> 63) & 1ull) : 0;
        if (in) {
          const long f = ((long)k * d1 + j) * d0 + i;
          q.o0 = Vf[f];
          q.v0 = Av[f];
          if (i + 1 < d0) { q.ob = Vf[f + 1]; q.vb = Av[f + 1]; }
          if (j + 1 < d1) { q.od = Vf[f + d0]; q.vd = Av[f + d0]; }
          if (k + 1 < d2) { q.of = Vf[f + (long)d0 * d1]; q.vf = Av[f + (long)d0 * d1]; }
        }
      }
      return q;
    };
    float xj = 0.f;                // this wave's previous row
    unsigned long long Fprev = 0;  // its final flags
    bool good = true;
    RowIn cur = load_row(-lk);
    for (int s = 0; s < d1 + EXR_TK - 1; s++) {
      const int j = s - lk;
      const bool act = kin && j >= 0 && j < d1;
      // the -k row: the wavefront below (LDS), or the slab below; the face
      // before the chunk: the chunk before, same plane
      float xk = 0.f, xc = 0.f;
      if (act) {
        if (lk > 0) {
          xk = xfin[lk - 1][j & 1][lane];
        } else if (m > 0) {
          unsigned long long x = ex_ld(xkIn + (long)j * 64 + lane);
          int polls = 0;
          while ((unsigned)(x >> 32) != E && good) {
            if (++polls > kExPolls) good = false;
            __builtin_amdgcn_s_sleep(1);
            x = ex_ld(xkIn + (long)j * 64 + lane);
          }
          xk = __uint_as_float((unsigned)x);
        }
        if (I > 0 && lane == 0) {
          unsigned long long x = ex_ld(xiIn + j);
          int polls = 0;
          while ((unsigned)(x >> 32) != E && good) {
            if (++polls > kExPolls) good = false;
            __builtin_amdgcn_s_sleep(1);
            x = ex_ld(xiIn + j);
          }
          xc = __uint_as_float((unsigned)x);
        }
      }
      // the next row's inputs, in flight during this row's recurrence
      const RowIn nxt = load_row(j + 1);
      if (act) {
        const long f = ((long)k * d1 + j) * d0 + i;
        const bool V = in && ((cur.Fr >> lane) & 1ull);
        const int ua = lane > 0 ? (int)((cur.Fr >> (lane > 0 ? lane - 1 : 0)) & 1ull) : cur.cin;
        const int uc = j > 0 ? (int)((Fprev >> lane) & 1ull) : 0;
        const int ue = k > 0 ? (int)((cur.Fk >> lane) & 1ull) : 0;
        const int cnt = ua + (cur.ob ? 1 : 0) + uc + (cur.od ? 1 : 0) + ue + (cur.of ? 1 : 0);
        const bool fill = V && !cur.o0;
        const float r = ex_rcp(cnt);
        // the +i, +j, +k neighbours' values, in the reference's order
        float Rs = 0.f;
        Rs = cur.ob ? Rs + cur.vb : Rs;
        Rs = cur.od ? Rs + cur.vd : Rs;
        Rs = cur.of ? Rs + cur.vf : Rs;
        const float Q = fill ? Rs * r : (V ? cur.v0 : 0.f);
        const float Pj = (fill && uc) ? r : 0.f;
        const float Pk = (fill && ue) ? r : 0.f;
        const float A0 = (fill && lane == 0 && ua) ? r : 0.f;  // the face before the chunk
        float A = (fill && lane > 0 && ua) ? r : 0.f;
        float As[6];
        As[0] = A;
        A *= EXR_DPP(A, 0x111, 0xf, 1.f);  // row_shr:1
        As[1] = A;
        A *= EXR_DPP(A, 0x112, 0xf, 1.f);  // row_shr:2
        As[2] = A;
        A *= EXR_DPP(A, 0x114, 0xf, 1.f);  // row_shr:4
        As[3] = A;
        A *= EXR_DPP(A, 0x118, 0xf, 1.f);  // row_shr:8
        As[4] = A;
        A *= EXR_DPP(A, 0x142, 0xa, 1.f);  // row_bcast:15 into rows 1 and 3
        As[5] = A;
        float B = fmaf(Pk, xk, fmaf(Pj, xj, Q));
        B = fmaf(A0, xc, B);
        EXR_FMAC_DPP(B, As[0], "row_shr:1 row_mask:0xf");
        EXR_FMAC_DPP(B, As[1], "row_shr:2 row_mask:0xf");
        EXR_FMAC_DPP(B, As[2], "row_shr:4 row_mask:0xf");
        EXR_FMAC_DPP(B, As[3], "row_shr:8 row_mask:0xf");
        EXR_FMAC_DPP(B, As[4], "row_bcast:15 row_mask:0xa");  // into rows 1 and 3
        EXR_FMAC_DPP(B, As[5], "row_bcast:31 row_mask:0xc");  // into rows 2 and 3
        if (lk < EXR_TK - 1) xfin[lk][j & 1][lane] = B;
        else if (m + 1 < S) xkOut[(long)j * 64 + lane] = ex_tagged(B, E);
        if (lane == 63 && I + 1 < C) xiOut[j] = ex_tagged(B, E);
        if (fill) {
          Av[f] = B;
          Vf[f] = 1;
        }
        xj = B;
        Fprev = cur.Fr;
      }
      cur = nxt;
      __syncthreads();
    }
    if (!good) sok = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (!sok) atomicOr(&ds->exFailed, 1u);
      atomicAdd(&R.done[comp], 1u);
    }
  }
}
#undef EXR_DPP
#undef EXR_FMAC_DPP

// every job of every component done (a component whose XCD ran no
// workgroup would leave its jobs undone)
__global__ void k_extrap_check(Geo g, ExRun R, DevScalars *ds) {
  if (threadIdx.x >= 3) return;
  const int comp = threadIdx.x;
  int d0, d1, d2;
  long base;
  ex_dims(g, comp, d0, d1, d2, base);
  const unsigned jobs = (unsigned)(((d2 + EXR_TK - 1) / EXR_TK) * ((d0 + 63) / 64));
  if (R.done[comp] != jobs) atomicOr(&ds->exFailed, 2u);
}

}  // namespace ws
