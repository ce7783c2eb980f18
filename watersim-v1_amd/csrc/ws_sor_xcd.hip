// XCD-resident red-black SOR sweeps (sorMode 6).
//
// n sweeps of the fluid lists in one launch, with the arithmetic of
// k_sor_list (nb_sum, then sor_relax), so the bits equal n pairs of list
// launches. The pressure relaxation is solver.js:1397-1404; red-black
// ordering and the lists are described with k_sor_list.
//
// Why one XCD. A 128^3 planet has about 90 000 fluid cells of each colour:
// one colour is about a megabyte of traffic, a fraction of a microsecond at
// HBM rate, while a kernel launch costs several. The sweeps are therefore
// bound by how fast half-sweep h + 1 can start after half-sweep h, not by
// bytes. The compute units of one XCD share one L2: a value that one
// workgroup stores is seen by another's agent-scope load (it bypasses the
// per-CU L1 and hits that L2) once the store has completed, with no L2
// write-back or invalidation, and waiting for a neighbour costs an L2 round
// trip. Across XCDs the same exchange needs write-through stores and loads
// from memory. So the launch runs on the 32 compute units of one XCD.
//
// Placement. Workgroups are dealt to the 8 XCDs in turn, so blocks
// b = 0, 8, 16, ... of a grid of 8 (G - 1) + 1 blocks are G workgroups of
// one XCD; the other blocks return at once. Each workgroup reads its XCD id
// (hardware register XCC_ID) and the launch checks, before anything is
// written, that all G agree and that every workgroup's cells fit its LDS.
// If not, every workgroup returns and the launch counts itself as declined
// in DevScalars::xcdDeclined (and in xcdTooBig when a range or its halo did
// not fit); q is untouched and the host runs the same sweeps with
// k_sor_list.
//
// Work split. Workgroup w owns one contiguous range of cell indices: the
// red list cut into G equal parts, and the black entries of the same cells.
// It keeps the q of those cells in LDS for the whole launch, in list order,
// and each thread keeps the code word, rhs, weight and six neighbour
// references (LDS slots) of its cells in registers. A cell's neighbours lie
// within nx*ny cells of it, so only the cells within nx*ny of the range ends
// ("boundary" cells) are read by other workgroups or read other
// workgroups' cells; the others ("interior") touch LDS only.
//
// Exchange. A boundary cell publishes each new value together with the
// number of its half-sweep, as one 64-bit word (value bits, tag) in an
// exchange array indexed by cell (written with plain stores: they reach the
// XCD's L2 through the write-through L1). A launch given epoch E tags the
// values it starts from E and the value computed in half-sweep h (0-based)
// E + 1 + h. Each foreign neighbour of a boundary cell has a halo slot in
// LDS; before half-sweep h a workgroup loads its halo of the other colour
// with agent-scope loads and waits, slot by slot, for tag E + h. A consumer
// thus waits on the data itself, with no separate progress flag. A
// workgroup reads all of its halo before it publishes any value of the same
// half-sweep (a workgroup barrier separates the two), so a producer, which
// needs that value to compute its next value of the same cell, can never
// overwrite a value its neighbour has yet to read.
//
// Order within a half-sweep: halo loads issued, interior cells relaxed
// while they are in flight, halo tags checked (stale slots polled again),
// halo written to LDS, barrier, boundary cells relaxed and published,
// barrier. Every owned cell reaches q once, at the end of the launch.
//
// Tags rise monotonically across launches (the host hands each launch an
// epoch above every tag of the launches before it), and a launch rewrites
// every slot its neighbours read, with tag E, before the arrival barrier;
// so no array needs clearing between launches. Waits poll a bounded number
// of times; a wait that gives up sets DevScalars::xcdTimedOut and the
// workgroups leave, and the host reports the solve as failed.
#include "ws_kernels.h"

namespace ws {

namespace {

constexpr int kXcdPolls = 1 << 20;   // arrival
constexpr int kHaloPolls = 1 << 16;  // one halo slot

__device__ __forceinline__ unsigned ld_agent(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_agent(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long tagged(float v, unsigned tag) {
  return (unsigned long long)__float_as_uint(v) | (unsigned long long)tag << 32;
}

}  // namespace

// NT threads, U cells of each colour per thread, H halo slots per colour
template <int NT, int U, int H>
__global__ __launch_bounds__(NT) void k_sor_xcd(XcdSweep J) {
  if (blockIdx.x % 8) return;
  const int w = blockIdx.x / 8, G = J.G;
  constexpr int CAP = U * NT;      // cells of one colour a workgroup can hold
  constexpr int HB = 2 * CAP;      // halo slots of colour col start at HB + col H
  constexpr int HU = (H + NT - 1) / NT;
  static_assert(2 * CAP + 2 * H <= 0xffff, "LDS slots in 16 bits");
  __shared__ float sq[2 * CAP + 2 * H];  // red cells of the range, black cells, red halo, black halo
  __shared__ unsigned hsrc[2][H];        // cell index each halo slot copies
  __shared__ int hcount[2];
  __shared__ int sclo[kXcdMaxG + 1];
  __shared__ int sok;
  const int nx = J.g.nx, nxy = nx * J.g.ny;
  const long nR = J.ds->nRed;
  for (int v = threadIdx.x; v <= G; v += NT) {
    const long r = nR * v / G;
    sclo[v] = v == 0 ? 0 : (r < nR ? (int)(J.code[r] & 0x3ffffffu) : J.g.nC);
  }
  if (threadIdx.x == 0) {
    sok = 1;
    hcount[0] = hcount[1] = 0;
  }
  __syncthreads();
  const int clo = sclo[w], chi = sclo[w + 1];
  const long lo[2] = {nR * w / G, (long)(J.off[clo] >> 32)};
  const long cnt[2] = {nR * (w + 1) / G - lo[0], (long)(J.off[chi] >> 32) - lo[1]};
  const bool fits = cnt[0] <= CAP && cnt[1] <= CAP;
  // list positions of each colour: [0, ba) and [bb, cnt) are boundary cells
  // (within nx*ny of the range ends), [ba, bb) interior ones
  long ba[2], bb[2];
  {
    const int xa = min(clo + nxy, chi), xb = max(chi - nxy, xa);
    const unsigned long long oa = J.off[xa], ob = J.off[xb];
    ba[0] = (long)(oa & 0xffffffffull) - lo[0];
    bb[0] = (long)(ob & 0xffffffffull) - lo[0];
    ba[1] = (long)(oa >> 32) - lo[1];
    bb[1] = (long)(ob >> 32) - lo[1];
  }
  // thread t's cells: k = t + u NT in an order with the boundary cells
  // first (positions [0, ba), then [bb, cnt), then [ba, bb)); k >= cnt: none
  int pos[2][U];
  int nbd[2];
#pragma unroll
  for (int col = 0; col < 2; col++) {
    const long nhi = cnt[col] - bb[col];
    nbd[col] = (int)(ba[col] + nhi);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long k = threadIdx.x + u * NT;
      pos[col][u] = (int)(!fits || k >= cnt[col] ? CAP : (k < ba[col] ? k : (k < nbd[col] ? bb[col] + (k - ba[col])
                                                                                      : k - nhi)));
    }
  }
  const unsigned E = J.epoch;
  unsigned e[2][U], ref[2][U][3];
  float rh[2][U], pw[2][U];
  const int step[6] = {-1, 1, -nx, nx, -nxy, nxy};
#pragma unroll
  for (int col = 0; col < 2; col++) {
    const long base = col ? J.half + lo[1] : lo[0];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = pos[col][u];
      const bool in = t < CAP;
      e[col][u] = in ? J.code[base + t] : 0u;
      rh[col][u] = in ? J.lrhs[base + t] : 0.f;
      pw[col][u] = in ? J.lpw[base + t] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = pos[col][u];
      const bool in = t < CAP;
      const int c = (int)(e[col][u] & 0x3ffffffu);
      const unsigned m = e[col][u] >> 26;
      const unsigned own = (unsigned)(col * CAP + (in ? t : 0));
      unsigned r6[6];
#pragma unroll
      for (int d = 0; d < 6; d++) {
        const int n = c + step[d];
        unsigned r = own;  // absent neighbours: any slot (the sum adds +0)
        if (in && ((m >> d) & 1u)) {
          if (n >= clo && n < chi) {
            // a cell of the other colour in the range: its slot
            const unsigned long long o = J.off[n];
            r = col ? (unsigned)((long)(o & 0xffffffffull) - lo[0]) : (unsigned)(CAP + (long)(o >> 32) - lo[1]);
          } else {
            const int hs = atomicAdd(&hcount[1 - col], 1);
            if (hs < H) {
              hsrc[1 - col][hs] = (unsigned)n;
              r = (unsigned)(HB + (1 - col) * H + hs);
            }
          }
        }
        r6[d] = r;
      }
      ref[col][u][0] = r6[0] | r6[1] << 16;
      ref[col][u][1] = r6[2] | r6[3] << 16;
      ref[col][u][2] = r6[4] | r6[5] << 16;
      if (in) {
        const float v = J.q[c];
        sq[own] = v;
        if (t < ba[col] || t >= bb[col]) J.xq[c] = tagged(v, E);
      }
    }
  }
  __syncthreads();
  // arrival: this workgroup's XCD and whether its cells and halo fit; the
  // starting values above are complete before the counter says E + 1
  if (threadIdx.x == 0) {
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);  // HW_REG_XCC_ID, bits 3:0
    const unsigned ok = fits && hcount[0] <= H && hcount[1] <= H;
    st_agent(J.flow + (long)w * kFlowStride + 1, xcc | ok << 8 | (ok ? 0u : 1u << 9));
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) st_agent(J.flow + (long)w * kFlowStride, E + 1);
  if (threadIdx.x < 64) {
    bool good = true;
    for (int v = threadIdx.x; v < G && good; v += 64) {
      int polls = 0;
      while (ld_agent(J.flow + (long)v * kFlowStride) < E + 1) {
        if (++polls > kXcdPolls || ld_agent(&J.ds->xcdTimedOut)) {
          good = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (!good) {
      st_agent(&J.ds->xcdTimedOut, 1u);
      sok = 0;
    }
  }
  __syncthreads();
  if (!sok) return;
  __shared__ int sbig;
  if (threadIdx.x == 0) sbig = 0;
  __syncthreads();
  if (threadIdx.x < 64) {
    const unsigned mine = ld_agent(J.flow + 1) & 0xffu;
    bool good = true, big = false;
    for (int v = threadIdx.x; v < G; v += 64) {
      const unsigned s = ld_agent(J.flow + (long)v * kFlowStride + 1);
      good = good && (s & 0xffu) == mine && ((s >> 8) & 1u);
      big = big || ((s >> 9) & 1u);
    }
    if (!good) sok = 0;
    if (big) sbig = 1;
  }
  __syncthreads();
  if (!sok) {
    if (w == 0 && threadIdx.x == 0) {
      if (sbig) atomicAdd(&J.ds->xcdTooBig, 1u);
      atomicAdd(&J.ds->xcdDeclined, 1u);
    }
    return;
  }
  const int nh[2] = {hcount[0], hcount[1]};

  auto relax = [&](int col, int u) {
    const int own = col * CAP + pos[col][u];
    const unsigned m = e[col][u] >> 26;
    const unsigned r0 = ref[col][u][0], r1 = ref[col][u][1], r2 = ref[col][u][2];
    const float a0 = sq[r0 & 0xffffu], a1 = sq[r0 >> 16], a2 = sq[r1 & 0xffffu];
    const float a3 = sq[r1 >> 16], a4 = sq[r2 & 0xffffu], a5 = sq[r2 >> 16];
    // nb_sum's order: the six directions in turn, absent ones adding +0
    float sum = 0.f;
    sum += (m & 1u) ? a0 : 0.f;
    sum += (m & 2u) ? a1 : 0.f;
    sum += (m & 4u) ? a2 : 0.f;
    sum += (m & 8u) ? a3 : 0.f;
    sum += (m & 16u) ? a4 : 0.f;
    sum += (m & 32u) ? a5 : 0.f;
    const float v = sor_relax(sq[own], sum, rh[col][u], pw[col][u], J.om1);
    sq[own] = v;
    return v;
  };

  unsigned h = 0;  // half-sweeps done
  for (int s = 0; s < J.sweeps; s++) {
#pragma unroll
    for (int col = 0; col < 2; col++) {
      const int oc = 1 - col;
      const unsigned want = E + h;  // the other colour's values of half-sweep h - 1 (h = 0: the start)
      unsigned long long hx[HU];
#pragma unroll
      for (int i = 0; i < HU; i++) {
        const int sl = threadIdx.x + i * NT;
        hx[i] = sl < nh[oc] ? ld_agent(J.xq + hsrc[oc][sl]) : 0ull;
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if ((int)threadIdx.x + u * NT >= nbd[col] && pos[col][u] < CAP) relax(col, u);
      unsigned stale = 0;
#pragma unroll
      for (int i = 0; i < HU; i++)
        if ((int)threadIdx.x + i * NT < nh[oc] && (unsigned)(hx[i] >> 32) != want) stale |= 1u << i;
      int polls = 0;
      while (stale) {
        if (++polls > kHaloPolls || ((polls & 255) == 0 && ld_agent(&J.ds->xcdTimedOut))) {
          st_agent(&J.ds->xcdTimedOut, 1u);
          sok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int i = 0; i < HU; i++)
          if ((stale >> i) & 1u) {
            hx[i] = ld_agent(J.xq + hsrc[oc][threadIdx.x + i * NT]);
            if ((unsigned)(hx[i] >> 32) == want) stale &= ~(1u << i);
          }
      }
#pragma unroll
      for (int i = 0; i < HU; i++) {
        const int sl = threadIdx.x + i * NT;
        if (sl < nh[oc]) sq[HB + oc * H + sl] = __uint_as_float((unsigned)hx[i]);
      }
      // the halo is in LDS, and all of it was read before any value of this
      // half-sweep is published
      __syncthreads();
      if (!sok) return;
#pragma unroll
      for (int u = 0; u < U; u++)
        if ((int)threadIdx.x + u * NT < nbd[col] && pos[col][u] < CAP)
          J.xq[e[col][u] & 0x3ffffffu] = tagged(relax(col, u), E + 1 + h);
      h++;
      // this colour complete in LDS before the other colour reads it
      __syncthreads();
    }
  }
#pragma unroll
  for (int col = 0; col < 2; col++)
#pragma unroll
    for (int u = 0; u < U; u++)
      if (pos[col][u] < CAP) J.q[e[col][u] & 0x3ffffffu] = sq[col * CAP + pos[col][u]];
}

// the arrival counters and placement reports: zero before the first launch
// and before a launch whose epoch would wrap
__global__ void k_xcd_arm(int G, unsigned *__restrict__ flow) {
  for (int v = threadIdx.x; v < G; v += blockDim.x) {
    flow[(long)v * kFlowStride] = 0u;
    flow[(long)v * kFlowStride + 1] = 0u;
  }
}

template __global__ void k_sor_xcd<1024, 1, 1024>(XcdSweep);
template __global__ void k_sor_xcd<1024, 2, 2048>(XcdSweep);
template __global__ void k_sor_xcd<1024, 3, 3072>(XcdSweep);

}  // namespace ws
