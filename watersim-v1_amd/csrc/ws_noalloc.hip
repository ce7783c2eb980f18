// MarchingTetrahedra.build (js/surface.js:66-177) for callers that may not
// allocate device memory: the same kernels as ws_mt_build, with the field and
// the mesh in the caller's device buffers and every piece of scratch (the
// per-cell vertex counts and their exclusive scan, the scan's block sums, the
// volume partial sums) in one static device array of fixed size.
//
// The array is a host-side symbol in host code; its device address (one per
// GPU) is looked up once per device when a context is created
// (mt_scratch_setup, called by ws_create with the device current) and only that
// address is handed to kernels and copies. Every call on a GPU shares that
// GPU's array: calls that overlap on one GPU race.
#include <mutex>

#include "ws_frame.h"

using namespace ws;

namespace {

constexpr long kMaxCells = WS_MT_NOALLOC_MAX_CELLS;
constexpr int kScanBlock = 1024;  // elements per block of the scan
constexpr long kScanBlocks = (kMaxCells + 1 + kScanBlock - 1) / kScanBlock;
constexpr long kVolParts = (kMaxCells + 255) / 256;
// layout: offsets (kMaxCells + 1 u32), scan block sums (kScanBlocks u32),
// volume partials (kVolParts + 1 doubles), each start 256-byte aligned
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kOffBytes = align256(sizeof(unsigned) * (kMaxCells + 1));
constexpr size_t kBlkBytes = align256(sizeof(unsigned) * kScanBlocks);
constexpr size_t kVolBytes = align256(sizeof(double) * (kVolParts + 1));
constexpr size_t kScratchBytes = kOffBytes + kBlkBytes + kVolBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// exclusive scan of in[0..n) (in[n] taken as 0) into out[0..n], in place
// allowed, in three passes: block totals, a scan of the totals, the
// block-local scan plus the block's offset
__device__ __forceinline__ unsigned block_scan_excl(unsigned v, unsigned *total) {
  __shared__ unsigned wsum[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned x = v;
  for (int o = 1; o < 64; o <<= 1) {
    unsigned y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned run = 0;
    for (int w = 0; w < kScanBlock / 64; w++) {
      unsigned t = wsum[w];
      wsum[w] = run;
      run += t;
    }
    *total = run;
  }
  __syncthreads();
  return x - v + wsum[wv];
}

__global__ __launch_bounds__(kScanBlock) void k_scan_totals(const unsigned *__restrict__ in, long n,
                                                            unsigned *__restrict__ blk) {
  __shared__ unsigned tot;
  const long i = (long)blockIdx.x * kScanBlock + threadIdx.x;
  const unsigned v = i < n ? in[i] : 0u;
  block_scan_excl(v, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_top(unsigned *__restrict__ blk, long nb) {
  __shared__ unsigned tot;
  unsigned carry = 0;
  for (long b0 = 0; b0 < nb; b0 += kScanBlock) {
    const long b = b0 + threadIdx.x;
    const unsigned v = b < nb ? blk[b] : 0u;
    const unsigned e = block_scan_excl(v, &tot);
    if (b < nb) blk[b] = carry + e;
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(unsigned *__restrict__ data, long n,
                                                           const unsigned *__restrict__ blk) {
  __shared__ unsigned tot;
  const long i = (long)blockIdx.x * kScanBlock + threadIdx.x;
  const unsigned v = i < n ? data[i] : 0u;
  const unsigned e = block_scan_excl(v, &tot);
  if (i <= n) data[i] = blk[blockIdx.x] + e;
}

}  // namespace

namespace ws {

// the device address of this GPU's scratch array, looked up once per device
int mt_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return WS_EINVAL;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return WS_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return WS_EHIP;
  g_scratch[device] = p;
  return WS_OK;
}

}  // namespace ws

extern "C" {

int ws_mt_build_noalloc(ws_ctx *c, const float *dens_dev, int nx, int ny, int nz, double dx, double iso,
                        float *pos_dev, float *nrm_dev, long cap, long *count, double *volume) {
  if (!c || nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || cap < 0 || (pos_dev && !nrm_dev)) return WS_EINVAL;
  const long nC = (long)nx * ny * nz;
  if (nC > kMaxCells) {
    c->err = "ws_mt_build_noalloc: at most " + std::to_string(kMaxCells) + " cells (" + std::to_string(nC) +
             " asked)";
    return WS_EINVAL;
  }
  const float *dd = dens_dev;
  if (!dd) {
    if (nx != c->nx || ny != c->ny || nz != c->nz) {
      c->err = "ws_mt_build_noalloc on the context field needs the context dimensions";
      return WS_EINVAL;
    }
    dd = c->dens;
  }
  void *base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
  if (!base) {
    c->err = "ws_mt_build_noalloc: no scratch address for this context's device";
    return WS_ESTATE;
  }
  unsigned *off = (unsigned *)base;
  unsigned *blk = (unsigned *)((unsigned char *)base + kOffBytes);
  double *vpart = (double *)((unsigned char *)base + kOffBytes + kBlkBytes);
  // vertex counts (k_mt_count writes every cell), their exclusive scan with
  // the total at off[nC]
  hipLaunchKernelGGL(k_mt_count, dim3(blocks_for(nC, 256)), dim3(256), 0, c->stream, nx, ny, nz, dd, iso, off);
  const long nb = (nC + 1 + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_totals, dim3(nb), dim3(kScanBlock), 0, c->stream, off, nC, blk);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, c->stream, blk, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, c->stream, off, nC, blk);
  // emit into the caller's buffers (whole mesh up to cap vertices) and the
  // FP64 volume, summed as ws_mt_build sums it
  const int nv = blocks_for(nC, 256);
  hipLaunchKernelGGL(k_mt_emit, dim3(nv), dim3(256), 0, c->stream, nx, ny, nz, dx, dd, iso, off, pos_dev,
                     pos_dev ? nrm_dev : nullptr, pos_dev ? cap : 0L, vpart);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, c->stream, vpart, nv, vpart + nv);
  WS_LAUNCHED();
  // the two results into the context's pinned scalar mirror
  c->hs->mt_count = 0;
  WS_HIP(hipMemcpyAsync(&c->hs->vol, vpart + nv, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipMemcpyAsync(&c->hs->mt_count, off + nC, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  if (count) *count = (long)c->hs->mt_count;
  if (volume) *volume = std::fabs(c->hs->vol);
  return WS_OK;
}

}  // extern "C"
