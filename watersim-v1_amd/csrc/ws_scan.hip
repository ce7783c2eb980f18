// Exclusive prefix sums of unsigned 32- and 64-bit counts in one launch.
//
// The substep scans per-cell and per-particle counts into offsets: the
// particle bins (cell counts), the fluid lists (red count | black count << 32
// per cell) and the random-number offsets of the vapor passes. A library scan
// is two launches (an initialisation of its look-back state, then the scan)
// with a shared counter; this one is a single launch without atomics:
//
// - block b scans tile b (NT x IPT entries): the tile is read striped
//   (coalesced) into LDS (padded so that neither the striped nor the
//   per-thread accesses conflict on banks), each thread scans IPT consecutive
//   entries, the workgroup scans the thread totals;
// - it publishes its tile total as status words tagged with the launch's
//   epoch (value | epoch << 32; a 64-bit total as two such words), then one
//   wavefront sums the totals of every earlier tile, reading 256 of them per
//   round with all reads in flight, and waiting on those whose tag is not
//   yet this launch's. No tile waits for another's prefix, so there is no
//   chain of look-backs;
// - blocks are dispatched in index order on each XCD, so the unfinished
//   tile of lowest index is always running: every wait ends;
// - the epoch lives in device memory: the last tile, which has seen every
//   other tile's word (so every block has read the epoch), moves it on for
//   the next launch. No launch needs a memset and a captured graph replays
//   correctly;
// - status words are written with agent-scope atomic stores and read with
//   agent-scope atomic loads, which bypass the compute unit's L1 and see the
//   stores of every XCD.
//
// Waits poll a bounded number of times; a wait that gives up sets
// DevScalars::scanFailed, which the host reports. Integer sums are exact, so
// the result does not depend on the tiling.
#include <string>

#include "ws_frame.h"

namespace ws {

namespace {

constexpr int kScanNT = 256;
constexpr int kScanPolls = 1 << 20;

__device__ __forceinline__ unsigned long long scan_read(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void scan_publish(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// status words per tile: one for a 32-bit total, two (low, high half) for 64
template <typename T>
constexpr int kWords = sizeof(T) / 4;

// the total of tile q from its words, waiting until they carry epoch E
template <typename T>
__device__ __forceinline__ T scan_total(const unsigned long long *status, long q, unsigned E, int &polls,
                                        unsigned *failed) {
  T v = 0;
#pragma unroll
  for (int h = 0; h < kWords<T>; h++) {
    unsigned long long s = scan_read(status + q * kWords<T> + h);
    while ((unsigned)(s >> 32) != E) {
      if (++polls > kScanPolls) {
        atomicOr(failed, 1u);
        s = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      s = scan_read(status + q * kWords<T> + h);
    }
    v |= (T)(unsigned)s << (32 * h);
  }
  return v;
}

// IPT entries per thread, NT threads: a tile of NT * IPT entries. ctl[0] is
// the epoch (never 0: freshly cleared words carry tag 0).
template <typename T, int IPT>
__device__ __forceinline__ void k_scan_excl_body(long bx, long nbx, const T *in, T *out, long n,
                                                 unsigned long long *__restrict__ status, unsigned *__restrict__ ctl,
                                                 unsigned *__restrict__ failed) {
  constexpr int TILE = kScanNT * IPT;
  // entry e of the tile at e + e / IPT: thread t's IPT entries are
  // contiguous and consecutive threads' runs start IPT + 1 words apart
  __shared__ T buf[TILE + TILE / IPT];
  __shared__ T wtot[kScanNT / 64];
  __shared__ T tilePrefix;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long tile = bx, nTiles = nbx;
  const long base = tile * TILE;
  const unsigned E = __hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int u = 0; u < IPT; u++) {
    const int e = u * kScanNT + threadIdx.x;
    const long at = base + e;
    buf[e + e / IPT] = at < n ? in[at] : (T)0;
  }
  __syncthreads();
  T v[IPT];
  T run = 0;
#pragma unroll
  for (int u = 0; u < IPT; u++) {
    const T x = buf[threadIdx.x * (IPT + 1) + u];
    v[u] = run;
    run += x;
  }
  T incl = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wtot[wave] = incl;
  __syncthreads();
  T before = incl - run;
  T agg = 0;
#pragma unroll
  for (int w = 0; w < kScanNT / 64; w++) {
    if (w < wave) before += wtot[w];
    agg += wtot[w];
  }
  if (threadIdx.x < kWords<T>)
    scan_publish(status + tile * kWords<T> + threadIdx.x,
                 (unsigned long long)(unsigned)(agg >> (32 * threadIdx.x)) | (unsigned long long)E << 32);
  if (wave == 0) {
    T prefix = 0;
    int polls = 0;
    for (long q0 = 0; q0 < tile; q0 += 64 * 4) {
      T t[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long q = q0 + r * 64 + lane;
        t[r] = q < tile ? scan_total<T>(status, q, E, polls, failed) : (T)0;
      }
#pragma unroll
      for (int r = 0; r < 4; r++) prefix += t[r];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) prefix += __shfl_xor(prefix, o, 64);
    if (lane == 0) {
      tilePrefix = prefix;
      // the last tile has seen every tile's word: the next launch's epoch
      if (tile == nTiles - 1)
        __hip_atomic_store(ctl, E + 1 == 0 ? 1u : E + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  const T add = tilePrefix + before;
#pragma unroll
  for (int u = 0; u < IPT; u++) buf[threadIdx.x * (IPT + 1) + u] = v[u] + add;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < IPT; u++) {
    const int e = u * kScanNT + threadIdx.x;
    const long at = base + e;
    if (at < n) out[at] = buf[e + e / IPT];
  }
}

template <typename T, int IPT>
__global__ __launch_bounds__(kScanNT) void k_scan_excl(const T *in, T *out, long n,
                                                       unsigned long long *__restrict__ status,
                                                       unsigned *__restrict__ ctl, unsigned *__restrict__ failed) {
  k_scan_excl_body<T, IPT>(blockIdx.x, gridDim.x, in, out, n, status, ctl, failed);
}
WS_PLANETS_OPS(k_scan_u32_ops, (k_scan_excl<unsigned, 16>), (k_scan_excl_body<unsigned, 16>))
WS_PLANETS_OPS(k_scan_u64_ops, (k_scan_excl<unsigned long long, 8>), (k_scan_excl_body<unsigned long long, 8>))

template <typename T, int IPT>
int launch_scan(ws_ctx *c, const T *in, T *out, long n) {
  if (n <= 0) return WS_OK;
  constexpr long TILE = (long)kScanNT * IPT;
  const long tiles = (n + TILE - 1) / TILE;
  if (tiles * kWords<T> > c->scanTiles * 2) {
    c->err = "scan of " + std::to_string(n) + " entries exceeds the scan state";
    return WS_EINVAL;
  }
  const StageOps &ops = sizeof(T) == 4 ? k_scan_u32_ops : k_scan_u64_ops;
  stage_launch(c, ops, k_scan_excl<T, IPT>, tiles, kScanNT, in, out, n, c->scanStatus, c->scanCtl,
               &c->ds->scanFailed);
  return WS_OK;
}

}  // namespace

int scan_excl(ws_ctx *c, const unsigned *in, unsigned *out, long n) { return launch_scan<unsigned, 16>(c, in, out, n); }

int scan_excl(ws_ctx *c, const unsigned long long *in, unsigned long long *out, long n) {
  return launch_scan<unsigned long long, 8>(c, in, out, n);
}

// tiles of the smallest tile size (64-bit entries) for n entries; the state
// holds two words per such tile
long scan_tiles_for(long n) { return (n + (long)kScanNT * 8 - 1) / ((long)kScanNT * 8); }

}  // namespace ws
