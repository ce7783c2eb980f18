// Frame-end particle contacts of the planet (js/solver.js:3290-3479):
// vapor-vapor collisions with the vapor-liquid contacts, and snow contacts.
//
// The reference resolves contacts in one serial sweep over the parcels
// sorted by x (then index), updating as it goes. Here the parcels are split
// into groups that cannot reach each other: connected components of the
// "closer than a threshold" graph (union-find over cell-binned
// neighbours). Each group is then resolved serially, in the reference's
// order, by one thread; groups run in parallel. Snow grains do not move in
// their pass, so with the contact radius as threshold the result is the
// reference's exactly. Vapor parcels move while they are resolved; the
// threshold is twice the radius, which holds every contact unless a chain
// of pushes carries a parcel more than one radius (or breaks the x order
// the reference's early exit relies on).
//
// The vapor-liquid contacts move only the vapor parcel and never the water,
// so every parcel is independent: one thread per parcel visits the liquid
// particles near it in index order, as the reference's outer loop does.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "ws_frame.h"

namespace ws {

namespace {

// subset flags: bit 0 vapor (2/3), bit 1 snow (5), bit 2 liquid contact
// partner (spray 1, rain 4, or grid fluid in the surface band)
__global__ void k_ct_flags(Geo g, int nP, const float *__restrict__ px, const float *__restrict__ py,
                           const float *__restrict__ pz, const uint8_t *__restrict__ pflag, float bandLo2,
                           unsigned *__restrict__ fv, unsigned *__restrict__ fi, unsigned *__restrict__ fl) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > nP) return;
  if (p == nP) { fv[p] = fi[p] = fl[p] = 0; return; }
  uint8_t f = pflag[p];
  fv[p] = (f == 2 || f == 3);
  fi[p] = f == 5;
  unsigned liq = 0;
  if (f == 1 || f == 4) liq = 1;
  else if (f == 0) {
    float ex = px[p] - g.cx, ey = py[p] - g.cy, ez = pz[p] - g.cz;
    liq = ex * ex + ey * ey + ez * ez >= bandLo2;
  }
  fl[p] = liq;
}

__global__ void k_ct_compact(int nP, const unsigned *__restrict__ flag, const unsigned *__restrict__ off,
                             unsigned *__restrict__ list) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nP && flag[p]) list[off[p]] = (unsigned)p;
}

// sort keys: x bits (positions are non-negative, so float order = bit
// order), then the particle index
__global__ void k_ct_keys(int n, const unsigned *__restrict__ list, const float *__restrict__ px,
                          unsigned long long *__restrict__ key) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  unsigned p = list[e];
  key[e] = ((unsigned long long)__float_as_uint(px[p]) << 32) | p;
}

__device__ __forceinline__ int cell_of(const Geo &g, float x, float y, float z) {
  int i = clampi((int)(x * g.inv_dx), 0, g.nx - 1), j = clampi((int)(y * g.inv_dx), 0, g.ny - 1);
  int k = clampi((int)(z * g.inv_dx), 0, g.nz - 1);
  return (k * g.ny + j) * g.nx + i;
}

// counting sort of a subset by cell (order inside a cell is not used)
__global__ void k_ct_count(Geo g, int n, const unsigned *__restrict__ list, const float *__restrict__ px,
                           const float *__restrict__ py, const float *__restrict__ pz, unsigned *__restrict__ cnt,
                           unsigned *__restrict__ cellOf, unsigned *__restrict__ rankIn) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  unsigned p = list[e];
  int c = cell_of(g, px[p], py[p], pz[p]);
  cellOf[e] = c;
  rankIn[e] = atomicAdd(&cnt[c], 1u);
}
// a binned member with its position copied beside it, so that a
// neighbour search reads one 16-byte record per candidate
struct __align__(16) Slot {
  float x, y, z;
  unsigned o;  // subset position
};

__global__ void k_ct_scatter(int n, const unsigned *__restrict__ list, const float *__restrict__ px,
                             const float *__restrict__ py, const float *__restrict__ pz,
                             const unsigned *__restrict__ cellOf, const unsigned *__restrict__ rankIn,
                             const unsigned *__restrict__ start, unsigned *__restrict__ binned, Slot *__restrict__ slot) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  unsigned b = start[cellOf[e]] + rankIn[e];
  binned[b] = (unsigned)e;  // subset position
  unsigned p = list[e];
  slot[b] = Slot{px[p], py[p], pz[p], (unsigned)e};
}

// Parents only ever point to smaller indices, so every walk ends. Reads and
// the path-halving writes are device-scope atomics: the links other
// workgroups make (on other XCDs) are seen without stale cache lines.
__device__ __forceinline__ unsigned ld_parent(unsigned *parent, unsigned x) {
  return __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned uf_find(unsigned *parent, unsigned x) {
  unsigned p = ld_parent(parent, x);
  while (p != x) {
    unsigned q = ld_parent(parent, p);
    if (q != p) __hip_atomic_store(&parent[x], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // halving
    x = p;
    p = ld_parent(parent, x);
  }
  return x;
}

// union of subset members closer than thr (the smaller index becomes the
// root); with a mask only masked members take part, the others stay alone.
// The cells of one row of the search box are adjacent in the cell order, so
// their members form one run of slots.
// one wavefront per member: the lanes test the slots of the neighbouring
// cells 64 at a time and link each member found closer than the threshold.
// Links always put the larger root under the smaller, so every component
// ends with its smallest member as root whatever the order of the links.
__global__ __launch_bounds__(256) void k_ct_union(Geo g, int n, const unsigned *__restrict__ list,
                                                  const float *__restrict__ px, const float *__restrict__ py,
                                                  const float *__restrict__ pz, const unsigned *__restrict__ cellOf,
                                                  const unsigned *__restrict__ start, const Slot *__restrict__ slot,
                                                  float thr2, int span, const unsigned *__restrict__ mask,
                                                  unsigned *__restrict__ parent) {
  const unsigned lane = threadIdx.x & 63;
  const int e = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (e >= n) return;
  if (mask && !mask[e]) return;
  const unsigned p = list[e];
  const float x = px[p], y = py[p], z = pz[p];
  const int c = cellOf[e];
  const int ci = c % g.nx, cj = (c / g.nx) % g.ny, ck = c / (g.nx * g.ny);
  const int i0 = max(ci - span, 0), i1 = min(ci + span, g.nx - 1);
  for (int kk = max(ck - span, 0); kk <= min(ck + span, g.nz - 1); kk++)
    for (int jj = max(cj - span, 0); jj <= min(cj + span, g.ny - 1); jj++) {
      const int row = (kk * g.ny + jj) * g.nx;
      const unsigned b0 = start[row + i0], be = start[row + i1 + 1];
      for (unsigned b1 = b0; b1 < be; b1 += 64) {
        const unsigned b = b1 + lane;
        if (b >= be) continue;
        const Slot s = slot[b];
        const unsigned o = s.o;
        if (o <= (unsigned)e) continue;
        if (mask && !mask[o]) continue;
        float ddx = s.x - x, ddy = s.y - y, ddz = s.z - z;
        if (ddx * ddx + ddy * ddy + ddz * ddz >= thr2) continue;
        // link the larger root under the smaller one
        for (int guard = 0; guard < 1 << 20; guard++) {
          unsigned ra = uf_find(parent, (unsigned)e), rb = uf_find(parent, o);
          if (ra == rb) break;
          unsigned hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
          if (atomicCAS(&parent[hi], hi, lo) == hi) break;
        }
      }
    }
}

// component key: (root, position in the x order); members outside the
// mask get the all-ones key and sort after every group. The key is also
// written at the member's slot.
__global__ void k_ct_groupkeys(int n, unsigned *__restrict__ parent, const unsigned *__restrict__ rankOf,
                               const unsigned *__restrict__ mask, const unsigned *__restrict__ cellOf,
                               const unsigned *__restrict__ rankIn, const unsigned *__restrict__ start,
                               unsigned long long *__restrict__ key, unsigned long long *__restrict__ slotKey) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  unsigned long long kv = ~0ull;
  if (!mask || mask[e]) kv = ((unsigned long long)uf_find(parent, (unsigned)e) << 32) | rankOf[e];
  key[e] = kv;
  slotKey[start[cellOf[e]] + rankIn[e]] = kv;
}

// position of each member in the x order
__global__ void k_ct_rank(int n, const unsigned long long *__restrict__ sortedKey, const unsigned *__restrict__ posOf,
                          unsigned *__restrict__ rankOf) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  unsigned p = (unsigned)(sortedKey[r] & 0xffffffffull);
  rankOf[posOf[p]] = (unsigned)r;
}
__global__ void k_ct_pos(int n, const unsigned *__restrict__ list, unsigned *__restrict__ posOf) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) posOf[list[e]] = (unsigned)e;
}

// the members of one group in x order: particle indices from the sorted
// group keys (low word = x rank, which indexes the x-sorted particle list)
__device__ __forceinline__ unsigned member(const unsigned long long *gk, const unsigned long long *xk, int t) {
  return (unsigned)(xk[(unsigned)(gk[t] & 0xffffffffull)] & 0xffffffffull);
}

// one past the last position of every group in the sorted group keys,
// indexed by the group's root (all-ones keys belong to no group)
__global__ void k_ct_group_ends(int n, const unsigned long long *__restrict__ gk, unsigned *__restrict__ ends) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || gk[t] == ~0ull) return;
  unsigned root = (unsigned)(gk[t] >> 32);
  if (t == n - 1 || gk[t + 1] == ~0ull || (unsigned)(gk[t + 1] >> 32) != root) ends[root] = (unsigned)(t + 1);
}

// One pair of _vaporCollisions' parcel-parcel sweep (solver.js:3380-3470)
// on two parcels held in registers: I is the earlier parcel in x order.
// Returns false where the x window of I ends (the sweep moves on to the
// next parcel); *touched tells whether J changed.
struct VaporState {
  float *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz;
};
struct Parcel {
  float x, y, z, vx, vy, vz, wx, wy, wz;
};

__device__ __forceinline__ Parcel parcel_load(const VaporState &P, unsigned i, int spin) {
  Parcel a;
  a.x = P.px[i]; a.y = P.py[i]; a.z = P.pz[i];
  a.vx = P.vx[i]; a.vy = P.vy[i]; a.vz = P.vz[i];
  a.wx = a.wy = a.wz = 0.f;
  if (spin) { a.wx = P.wx[i]; a.wy = P.wy[i]; a.wz = P.wz[i]; }
  return a;
}

__device__ __forceinline__ void parcel_store(const VaporState &P, unsigned i, const Parcel &a, int spin) {
  P.px[i] = a.x; P.py[i] = a.y; P.pz[i] = a.z;
  P.vx[i] = a.vx; P.vy[i] = a.vy; P.vz[i] = a.vz;
  if (spin) { P.wx[i] = a.wx; P.wy[i] = a.wy; P.wz[i] = a.wz; }
}

__device__ __forceinline__ bool vapor_pair(Parcel &I, Parcel &J, float rad, float rad2, int spin, bool *touched) {
  *touched = false;
  if (J.x - I.x >= rad) return false;
  float ddx = J.x - I.x, ddy = J.y - I.y, ddz = J.z - I.z;
  float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
  if (d2 >= rad2) return true;
  *touched = true;
  float d = sqrtf(d2);
  if (!(d != 0.f)) d = 1e-9f;
  float nX = ddx / d, nY = ddy / d, nZ = ddz / d;
  float vRel = (I.vx - J.vx) * nX + (I.vy - J.vy) * nY + (I.vz - J.vz) * nZ;
  if (vRel > 0.f) {
    I.vx -= vRel * nX; I.vy -= vRel * nY; I.vz -= vRel * nZ;
    J.vx += vRel * nX; J.vy += vRel * nY; J.vz += vRel * nZ;
  }
  float half = (rad - d) * 0.5f;
  I.x -= nX * half; I.y -= nY * half; I.z -= nZ * half;
  J.x += nX * half; J.y += nY * half; J.z += nZ * half;
  float kick = (rad - d) * 2.f;
  I.vx -= nX * kick; I.vy -= nY * kick; I.vz -= nZ * kick;
  J.vx += nX * kick; J.vy += nY * kick; J.vz += nZ * kick;
  if (!spin) return true;
  float R = rad * 0.5f;
  float wSx = I.wx + J.wx, wSy = I.wy + J.wy, wSz = I.wz + J.wz;
  float usx = (I.vx - J.vx) + R * (wSy * nZ - wSz * nY);
  float usy = (I.vy - J.vy) + R * (wSz * nX - wSx * nZ);
  float usz = (I.vz - J.vz) + R * (wSx * nY - wSy * nX);
  float uN = usx * nX + usy * nY + usz * nZ;
  float utx = usx - uN * nX, uty = usy - uN * nY, utz = usz - uN * nZ;
  float utM = sqrtf(utx * utx + uty * uty + utz * utz);
  if (utM < 1e-9f) return true;
  float jt = utM / 7.f;
  if (vRel > 0.f) { float mu = 0.25f * vRel; if (jt > mu) jt = mu; }
  float jx = -utx / utM * jt, jy = -uty / utM * jt, jz = -utz / utM * jt;
  I.vx += jx; I.vy += jy; I.vz += jz;
  J.vx -= jx; J.vy -= jy; J.vz -= jz;
  float invI = 1.f / (0.4f * R * R);
  float dwx = R * (nY * jz - nZ * jy) * invI, dwy = R * (nZ * jx - nX * jz) * invI;
  float dwz = R * (nX * jy - nY * jx) * invI;
  I.wx += dwx; I.wy += dwy; I.wz += dwz;
  J.wx += dwx; J.wy += dwy; J.wz += dwz;
  return true;
}

// vapor_pair's first two tests alone: 2 the pass over this parcel's partners
// stops here (the reference's early exit), 1 the pair touches, 0 neither
// (the pair leaves both parcels as they are). The same expressions as
// vapor_pair, so the decisions agree with it.
__device__ __forceinline__ int vapor_pair_test(const Parcel &I, const Parcel &J, float rad, float rad2) {
  if (J.x - I.x >= rad) return 2;
  float ddx = J.x - I.x, ddy = J.y - I.y, ddz = J.z - I.z;
  float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
  return d2 >= rad2 ? 0 : 1;
}

__device__ __forceinline__ float bcast(float v, unsigned k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)k));
}

__device__ __forceinline__ Parcel parcel_bcast(const Parcel &J, unsigned k) {
  Parcel K;
  K.x = bcast(J.x, k); K.y = bcast(J.y, k); K.z = bcast(J.z, k);
  K.vx = bcast(J.vx, k); K.vy = bcast(J.vy, k); K.vz = bcast(J.vz, k);
  K.wx = bcast(J.wx, k); K.wy = bcast(J.wy, k); K.wz = bcast(J.wz, k);
  return K;
}

// the reference's serial sweep of one group by one thread
__device__ void vapor_group_serial(int t0, int t1, const unsigned long long *__restrict__ gk,
                                   const unsigned long long *__restrict__ xk, float rad, float rad2, int spin,
                                   const VaporState &P) {
  for (int a = t0; a < t1; a++) {
    unsigned i = member(gk, xk, a);
    Parcel I = parcel_load(P, i, spin);
    unsigned jn = a + 1 < t1 ? member(gk, xk, a + 1) : 0;
    Parcel Jn = a + 1 < t1 ? parcel_load(P, jn, spin) : I;
    for (int b = a + 1; b < t1; b++) {
      // the next member's state is fetched while this pair is resolved
      unsigned j = jn;
      Parcel J = Jn;
      if (b + 1 < t1) {
        jn = member(gk, xk, b + 1);
        Jn = parcel_load(P, jn, spin);
      }
      bool touched;
      bool go = vapor_pair(I, J, rad, rad2, spin, &touched);
      if (touched) parcel_store(P, j, J, spin);
      if (!go) break;
    }
    parcel_store(P, i, I, spin);
  }
}

// The sweeps of the small groups (groups larger than serialMax and not
// marked for the serial path go to the dependency-ordered sweep). A
// wavefront takes the groups that start among its 64 positions, one at a
// time: lane l holds member l's state in registers for the whole sweep. Step
// a broadcasts member a, tests every later member against it at once (a pair
// that neither touches nor stops the pass changes nothing) and resolves the
// touching pairs in member order, each on the parcel as the pairs before it
// left it: the reference's serial sweep, with the states never leaving the
// registers. A group of more than 64 members runs the serial sweep on one
// lane.
__global__ __launch_bounds__(256) void k_ct_vapor_groups(int n, const unsigned long long *__restrict__ gk,
                                                         const unsigned long long *__restrict__ xk,
                                                         const unsigned *__restrict__ ends, float rad, int spin,
                                                         VaporState P, int serialMax,
                                                         const unsigned *__restrict__ serialOnly) {
  const unsigned lane = threadIdx.x & 63;
  const int base = (int)((blockIdx.x * blockDim.x + threadIdx.x) & ~63u);
  const int t = base + (int)lane;
  // the first member of a group (its key's root differs from the previous)
  const bool first = t < n && (t == 0 || (gk[t] >> 32) != (gk[t - 1] >> 32));
  unsigned long long starts = __ballot(first);
  const float rad2 = rad * rad;
  while (starts) {
    const int t0 = base + __builtin_ctzll(starts);
    starts &= starts - 1;
    const unsigned root = (unsigned)(gk[t0] >> 32);
    const int t1 = (int)ends[root];
    const int m = t1 - t0;
    if (m < 2) continue;
    if (m > serialMax && !serialOnly[root]) continue;
    if (m > 64) {
      if (lane == 0) vapor_group_serial(t0, t1, gk, xk, rad, rad2, spin, P);
      continue;
    }
    const bool in = (int)lane < m;
    const unsigned me = in ? member(gk, xk, t0 + (int)lane) : 0u;
    Parcel S{};
    if (in) S = parcel_load(P, me, spin);
    for (int a = 0; a + 1 < m; a++) {
      Parcel I = parcel_bcast(S, (unsigned)a);
      for (int k = a + 1; k < m;) {
        const bool act = (int)lane >= k && in;
        const int tst = act ? vapor_pair_test(I, S, rad, rad2) : 0;
        const unsigned long long hit = __ballot(tst != 0);
        if (!hit) break;
        const unsigned f = (unsigned)__builtin_ctzll(hit);
        if (__builtin_amdgcn_readlane(tst, (int)f) == 2) break;
        Parcel K = parcel_bcast(S, f);
        bool touched;
        bool go = vapor_pair(I, K, rad, rad2, spin, &touched);
        if (touched && lane == f) S = K;
        if (!go) break;
        k = (int)f + 1;
      }
      if ((int)lane == a) S = I;
    }
    if (in) parcel_store(P, me, S, spin);
  }
}

// ---- large vapor groups: the same sweep, run in dependency order.
// Parcel a's step of the sweep reads and writes a and its later partners,
// the parcels after it in x order within twice the contact radius (the
// group threshold); the others it only passes over. Two steps that share a
// parcel must run in sweep order; steps that share none commute. Each
// parcel x keeps the sweep-ordered list of the steps that touch it (its
// own last), and a counter of how many of them are done; step a may run
// when, for each parcel it touches, the counter equals a's position in that
// parcel's list. One workgroup sweeps one group (the steps of a group only
// touch its members), each thread taking every blockDim-th step in order
// and polling for readiness a bounded number of times; the earliest pending
// step is always ready, so every poll round completes at least one step.
constexpr int kVcK = 63;  // partners kept per parcel (more: the group is swept serially); the wavefront step holds kVcK + 1 successors in its 64 lanes

// per group root: whether the group is larger than serialMax (only those
// groups' steps can run in dependency order)
__global__ void k_vc_mark_big(int n, const unsigned long long *__restrict__ gk, const unsigned *__restrict__ ends,
                              int serialMax, unsigned *__restrict__ gbig) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (t > 0 && (gk[t] >> 32) == (gk[t - 1] >> 32)) return;
  const unsigned root = (unsigned)(gk[t] >> 32);
  gbig[root] = (int)ends[root] - t > serialMax;
}

// later partners of each parcel: same group, later x rank, closer than thr
// at the start of the pass, sorted by rank (at most kVcK; a parcel with more
// marks its group for the serial sweep). Parcels of groups no larger than
// serialMax get none: their groups are swept whole by one wavefront, and
// their steps then only take part in the pair keys as themselves.
// One wavefront per parcel: the lanes test the slots of the neighbouring
// cells 64 at a time, the partners found are gathered in LDS in the order
// found, and a bitonic sort across the lanes puts them in rank order (ranks
// are distinct).
__global__ __launch_bounds__(256) void k_vc_cand(Geo g, int n, const unsigned *__restrict__ list,
                                                 const unsigned *__restrict__ gbig, const float *__restrict__ px,
                                                 const float *__restrict__ py, const float *__restrict__ pz,
                                                 const unsigned *__restrict__ cellOf, const unsigned *__restrict__ start,
                                                 const Slot *__restrict__ slot,
                                                 const unsigned long long *__restrict__ slotKey, int span,
                                                 const unsigned long long *__restrict__ key, float thr2,
                                                 unsigned *__restrict__ cand, unsigned *__restrict__ ncand,
                                                 unsigned *__restrict__ serialOnly) {
  __shared__ unsigned sRank[4][64], sO[4][64];
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int e = (int)(blockIdx.x * 4 + wave);
  if (e >= n) return;
  const unsigned long long ke = key[e];
  const unsigned p = list[e], re = (unsigned)ke, root = (unsigned)(ke >> 32);
  if (!gbig[root]) {
    if (lane == 0) ncand[e] = 0u;
    return;
  }
  const float x = px[p], y = py[p], z = pz[p];
  const int c = cellOf[e];
  const int ci = c % g.nx, cj = (c / g.nx) % g.ny, ck = c / (g.nx * g.ny);
  unsigned m = 0;  // partners found (wave-uniform)
  const int i0 = max(ci - span, 0), i1 = min(ci + span, g.nx - 1);
  for (int kk = max(ck - span, 0); kk <= min(ck + span, g.nz - 1); kk++)
    for (int jj = max(cj - span, 0); jj <= min(cj + span, g.ny - 1); jj++) {
      const int row = (kk * g.ny + jj) * g.nx;
      const unsigned b0 = start[row + i0], be = start[row + i1 + 1];
      for (unsigned b1 = b0; b1 < be; b1 += 64) {
        const unsigned b = b1 + lane;
        bool ok = false;
        unsigned ro = 0, so = 0;
        if (b < be) {
          const unsigned long long ko = slotKey[b];
          ro = (unsigned)ko;
          if (ro > re && (unsigned)(ko >> 32) == root) {
            const Slot sl = slot[b];
            float ddx = sl.x - x, ddy = sl.y - y, ddz = sl.z - z;
            ok = !(ddx * ddx + ddy * ddy + ddz * ddz >= thr2);
            so = sl.o;
          }
        }
        const unsigned long long bal = __ballot(ok);
        const unsigned at = m + (unsigned)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
        if (ok && at < 64) {
          sRank[wave][at] = ro;
          sO[wave][at] = so;
        }
        m += (unsigned)__popcll(bal);
      }
    }
  __builtin_amdgcn_wave_barrier();
  const unsigned kept = m < (unsigned)kVcK ? m : (unsigned)kVcK;
  unsigned rk = lane < kept ? sRank[wave][lane] : 0xffffffffu;
  unsigned o = lane < kept ? sO[wave][lane] : 0u;
  // bitonic sort of the 64 lanes by rank, ascending
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const unsigned prk = __shfl_xor(rk, j, 64), po = __shfl_xor(o, j, 64);
      const bool up = (lane & k) == 0, lower = (lane & j) == 0;
      const bool take = lower == up ? prk < rk : prk > rk;
      if (take) {
        rk = prk;
        o = po;
      }
    }
  }
  if (lane < kept) cand[(long)e * kVcK + lane] = o;
  if (lane == 0) {
    ncand[e] = kept;
    if (m > (unsigned)kVcK) __hip_atomic_store(&serialOnly[root], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void k_vc_pair_count(int n, const unsigned *__restrict__ ncand, unsigned *__restrict__ cnt) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) cnt[e] = 1 + ncand[e];
  if (e == n) cnt[e] = 0;
}

// (touched parcel, step rank, slot): slot 0 is the step's own parcel,
// slot k its k-th partner
__global__ void k_vc_pair_emit(int n, const unsigned *__restrict__ rankOf, const unsigned *__restrict__ cand,
                               const unsigned *__restrict__ ncand, const unsigned *__restrict__ off,
                               unsigned long long *__restrict__ keys) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  unsigned long long r = (unsigned long long)rankOf[e] << 6;
  unsigned long long *o = keys + off[e];
  o[0] = ((unsigned long long)e << 32) | r;
  for (unsigned k = 0; k < ncand[e]; k++) o[1 + k] = ((unsigned long long)cand[(long)e * kVcK + k] << 32) | r | (k + 1);
}

__global__ void k_vc_chain_heads(int np, const unsigned long long *__restrict__ keys, unsigned *__restrict__ chainStart) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  unsigned x = (unsigned)(keys[s] >> 32);
  if (s == 0 || (unsigned)(keys[s - 1] >> 32) != x) chainStart[x] = (unsigned)s;
}

// the position of each step in the list of every parcel it touches
__global__ void k_vc_need(int np, const unsigned long long *__restrict__ keys, const unsigned *__restrict__ chainStart,
                          const unsigned long long *__restrict__ xk, const unsigned *__restrict__ posOf,
                          unsigned *__restrict__ need) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  unsigned long long k = keys[s];
  unsigned x = (unsigned)(k >> 32), rank = (unsigned)((k >> 6) & 0x3ffffffull), slot = (unsigned)(k & 63ull);
  unsigned e = posOf[(unsigned)(xk[rank] & 0xffffffffull)];
  need[(long)e * (kVcK + 1) + slot] = (unsigned)s - chainStart[x];
}

// the groups for the dependency-ordered sweep, as [t0, t1) ranges of the
// sorted group keys
__global__ void k_vc_big(int n, const unsigned long long *__restrict__ gk, const unsigned *__restrict__ ends,
                         int serialMax, const unsigned *__restrict__ serialOnly, unsigned *__restrict__ big,
                         unsigned *__restrict__ nBig) {
  int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= n) return;
  if (t0 > 0 && (gk[t0] >> 32) == (gk[t0 - 1] >> 32)) return;
  unsigned root = (unsigned)(gk[t0] >> 32);
  int t1 = (int)ends[root];
  if (t1 - t0 <= serialMax || serialOnly[root]) return;
  atomicMax(&nBig[2], (unsigned)(t1 - t0));   // largest group
  atomicAdd(&nBig[3], (unsigned)(t1 - t0));   // parcels in these groups
  unsigned slot = atomicAdd(nBig, 1u);
  big[2 * slot] = (unsigned)t0;
  big[2 * slot + 1] = (unsigned)t1;
}

__device__ __forceinline__ unsigned ld_wg_relaxed(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// one group's steps in dependency order; the per-parcel counts live in LDS
// (indexed by group position) when the group fits, else in global memory
// (indexed by subset element)
constexpr int kVcLds = 16384;

template <bool LDS>
struct VcStep {
  unsigned pi, e, nc, own;
  unsigned x[4], need[4], p[4];  // slots 0..3: counter index, position in its list, partner particle
  const unsigned *ne, *ce, *cp;  // the full lists (slots beyond 3)
};

// the metadata of step `cur` (its parcel, partners, list positions); the
// first four slots are kept in registers, the rest read from memory
template <bool LDS>
__device__ __forceinline__ void vc_step_load(VcStep<LDS> &S, int cur, int t0, const unsigned long long *__restrict__ gk,
                                             const unsigned long long *__restrict__ xk,
                                             const unsigned *__restrict__ posOf, const unsigned *__restrict__ cand,
                                             const unsigned *__restrict__ candT, const unsigned *__restrict__ candP,
                                             const unsigned *__restrict__ ncand, const unsigned *__restrict__ need) {
  unsigned r = (unsigned)(gk[cur] & 0xffffffffull);
  S.pi = (unsigned)(xk[r] & 0xffffffffull);
  S.e = posOf[S.pi];
  S.nc = ncand[S.e];
  S.ne = need + (long)S.e * (kVcK + 1);
  S.ce = LDS ? candT + (long)cur * kVcK : cand + (long)S.e * kVcK;
  S.cp = candP + (long)cur * kVcK;
  S.own = LDS ? (unsigned)(cur - t0) : S.e;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    bool in = (unsigned)k <= S.nc;
    S.need[k] = in ? S.ne[k] : 0;
    S.x[k] = k == 0 ? S.own : (in ? (LDS ? S.ce[k - 1] - (unsigned)t0 : S.ce[k - 1]) : 0);
    S.p[k] = (k > 0 && in) ? S.cp[k - 1] : S.pi;
  }
}

// register "arrays" indexed at run time through selects (an indexed
// register array would live in scratch memory)
__device__ __forceinline__ unsigned sel4(const unsigned (&a)[4], unsigned k) {
  return k == 0 ? a[0] : (k == 1 ? a[1] : (k == 2 ? a[2] : a[3]));
}

template <bool LDS>
__device__ __forceinline__ unsigned vc_slot_x(const VcStep<LDS> &S, unsigned k, int t0) {
  if (k < 4) return sel4(S.x, k);
  return LDS ? S.ce[k - 1] - (unsigned)t0 : S.ce[k - 1];
}

template <bool LDS>
__device__ __forceinline__ unsigned vc_slot_need(const VcStep<LDS> &S, unsigned k) {
  return k < 4 ? sel4(S.need, k) : S.ne[k];
}

template <bool LDS>
__device__ __forceinline__ unsigned vc_slot_p(const VcStep<LDS> &S, unsigned k) {
  return k < 4 ? sel4(S.p, k) : S.cp[k - 1];
}

template <bool LDS>
__device__ void vc_sweep_group(int t0, int t1, unsigned *cnt, const unsigned long long *__restrict__ gk,
                               const unsigned long long *__restrict__ xk, const unsigned *__restrict__ posOf,
                               const unsigned *__restrict__ cand, const unsigned *__restrict__ candT,
                               const unsigned *__restrict__ candP, const unsigned *__restrict__ ncand,
                               const unsigned *__restrict__ need, float rad, float rad2, int spin, VaporState P,
                               unsigned *__restrict__ unfinished) {
  int cur = t0 + (int)threadIdx.x;
  long limit = 64L * (t1 - t0) + 4096;
  // this thread's steps are cur, cur + blockDim, ...
  VcStep<LDS> S;
  if (cur < t1) vc_step_load<LDS>(S, cur, t0, gk, xk, posOf, cand, candT, candP, ncand, need);
  unsigned blocker = 0;
  for (long it = 0; cur < t1 && it < limit; it++) {
    // readiness: every touched parcel's count at this step's position (the
    // counts are read without ordering; one acquire fence then orders the
    // parcel reads after them). A slot once seen ready stays ready.
    bool ready = true;
    for (unsigned k = blocker; k <= S.nc; k++) {
      unsigned want = vc_slot_need<LDS>(S, k);
      if (__hip_atomic_load(&cnt[vc_slot_x<LDS>(S, k, t0)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) !=
          want) {
        ready = false;
        blocker = k;
        break;
      }
    }
    if (!ready) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    Parcel I = parcel_load(P, S.pi, spin);
    Parcel Jn = S.nc ? parcel_load(P, S.p[1], spin) : I;
    for (unsigned k = 0; k < S.nc; k++) {
      // the next partner's state is fetched while this pair is resolved
      unsigned j = vc_slot_p<LDS>(S, k + 1);
      Parcel J = Jn;
      if (k + 1 < S.nc) Jn = parcel_load(P, vc_slot_p<LDS>(S, k + 2), spin);
      bool touched;
      bool go = vapor_pair(I, J, rad, rad2, spin, &touched);
      if (touched) parcel_store(P, j, J, spin);
      if (!go) break;
    }
    parcel_store(P, S.pi, I, spin);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    for (unsigned k = 0; k <= S.nc; k++)
      __hip_atomic_store(&cnt[vc_slot_x<LDS>(S, k, t0)], vc_slot_need<LDS>(S, k) + 1, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
    // the next step's metadata, after the counts are published (its loads
    // stay off the path of the steps waiting on this one)
    cur += (int)blockDim.x;
    if (cur < t1) vc_step_load<LDS>(S, cur, t0, gk, xk, posOf, cand, candT, candP, ncand, need);
    blocker = 0;
  }
  if (cur < t1) atomicAdd(unfinished, 1u);
}

// Large groups that fit in LDS run from a ready queue instead: each step's
// count of unfinished predecessors lives in LDS; a finished step releases
// its writes, decrements its successors and queues those that reach zero.
// A wavefront runs one step at a time: lane k fetches partner k (all
// partners at once), the pairs are resolved in order on values broadcast
// from the partners' lanes (every lane computes the same numbers), and the
// lanes of the touched partners write them back.
constexpr int kVcQ = 16000;  // steps of a group kept in LDS
constexpr unsigned short kVcEmpty = 0xffffu;
constexpr int kVcIdle = 1 << 22;  // polls without progress before a group is given up

__global__ __launch_bounds__(1024) void k_vc_sweepq(const unsigned *__restrict__ big, const unsigned *__restrict__ nBig,
                                                    const unsigned *__restrict__ stepP,
                                                    const unsigned *__restrict__ stepN,
                                                    const unsigned *__restrict__ candP,
                                                    const unsigned *__restrict__ succ,
                                                    const unsigned *__restrict__ depc, float rad, int spin,
                                                    VaporState P, unsigned *__restrict__ unfinished, int qMax) {
  __shared__ unsigned sdep[kVcQ / 2];  // two 16-bit counts per word
  __shared__ unsigned short sq[kVcQ];
  __shared__ unsigned shead, stail, sdone;
  const float rad2 = rad * rad;
  const unsigned lane = threadIdx.x & 63;
  unsigned nb = *nBig;
  for (unsigned gi = blockIdx.x; gi < nb; gi += gridDim.x) {
    int t0 = (int)big[2 * gi], t1 = (int)big[2 * gi + 1], m = t1 - t0;
    if (m > qMax) continue;  // left to the polling sweep (qMax <= kVcQ)
    if (threadIdx.x == 0) { shead = 0; stail = 0; sdone = 0; }
    for (int q = threadIdx.x; q < m; q += blockDim.x) sq[q] = kVcEmpty;
    for (int w = threadIdx.x; 2 * w < m; w += blockDim.x) {
      unsigned lo = depc[t0 + 2 * w];
      unsigned hi = 2 * w + 1 < m ? depc[t0 + 2 * w + 1] : 0u;
      sdep[w] = lo | (hi << 16);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < m; q += blockDim.x)
      if (((sdep[q >> 1] >> (16 * (q & 1))) & 0xffffu) == 0) sq[atomicAdd(&stail, 1u)] = (unsigned short)q;
    __syncthreads();
    unsigned lastDone = 0;
    int idle = 0, status = 0;  // 0 a step popped, 1 nothing ready, 2 all done, 3 given up
    while (true) {
      unsigned q = 0;
      if (lane == 0) {
        unsigned done = __hip_atomic_load(&sdone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (done != lastDone) { lastDone = done; idle = 0; }
        status = 1;
        if (done >= (unsigned)m) {
          status = 2;
        } else if (++idle > kVcIdle) {
          status = 3;
        } else {
          unsigned h = __hip_atomic_load(&shead, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          unsigned tl = __hip_atomic_load(&stail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (h < tl && atomicCAS(&shead, h, h + 1) == h) {
            // slot h is claimed by a push whose entry may still be on its way
            unsigned short v = kVcEmpty;
            for (int w = 0; w < kVcIdle && v == kVcEmpty; w++)
              v = __hip_atomic_load(&sq[h], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            status = v == kVcEmpty ? 3 : 0;
            q = v;
          }
        }
      }
      status = __builtin_amdgcn_readfirstlane(status);
      if (status >= 2) break;
      if (status == 1) {
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      q = __builtin_amdgcn_readfirstlane(q);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const long t = t0 + (long)q;
      // the step's lists are read whole (they lie inside the step's rows of
      // candP and succ) together with its parcel and partner count, so the
      // step's metadata costs one memory round trip; the count then masks them
      const unsigned jpAll = lane < (unsigned)kVcK ? candP[t * kVcK + lane] : 0u;
      const unsigned sAll = succ[t * (kVcK + 1) + lane];
      unsigned pi = __builtin_amdgcn_readfirstlane(stepP[t]);
      unsigned nc = __builtin_amdgcn_readfirstlane(stepN[t]);
      bool mine = lane < nc;
      unsigned jp = mine ? jpAll : 0u;
      unsigned sT = lane <= nc ? sAll : 0xffffffffu;
      Parcel I = parcel_load(P, pi, spin);
      Parcel J{};
      if (mine) J = parcel_load(P, jp, spin);
      // The pairs are resolved in partner order, each on the parcel as the
      // pairs before it left it. A pair that neither touches nor stops the
      // pass changes nothing, so every lane tests its partner against the
      // current parcel at once and only the first that touches (or stops)
      // is resolved in turn; the lanes after it are tested again.
      bool wrote = false;
      for (unsigned k = 0; k < nc;) {
        const bool act = lane >= k && lane < nc;
        const int tst = act ? vapor_pair_test(I, J, rad, rad2) : 0;
        const unsigned long long hit = __ballot(tst != 0);
        if (!hit) break;
        const unsigned f = (unsigned)__builtin_ctzll(hit);
        if (__builtin_amdgcn_readlane(tst, (int)f) == 2) break;
        Parcel K = parcel_bcast(J, f);
        bool touched;
        bool go = vapor_pair(I, K, rad, rad2, spin, &touched);
        if (touched && lane == f) {
          J = K;
          wrote = true;
        }
        if (!go) break;
        k = f + 1;
      }
      if (wrote) parcel_store(P, jp, J, spin);
      if (lane == 0) parcel_store(P, pi, I, spin);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (sT != 0xffffffffu) {
        unsigned ls = sT - (unsigned)t0, sh = 16 * (ls & 1);
        unsigned old = __hip_atomic_fetch_add(&sdep[ls >> 1], 0u - (1u << sh), __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
        if (((old >> sh) & 0xffffu) == 1) {
          unsigned at = atomicAdd(&stail, 1u);
          __hip_atomic_store(&sq[at], (unsigned short)ls, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      if (lane == 0) atomicAdd(&sdone, 1u);
    }
    if (lane == 0 && status == 3) atomicAdd(unfinished, 1u);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_vc_sweep(const unsigned *__restrict__ big, const unsigned *__restrict__ nBig,
                                                  const unsigned long long *__restrict__ gk,
                                                  const unsigned long long *__restrict__ xk,
                                                  const unsigned *__restrict__ posOf, const unsigned *__restrict__ cand,
                                                  const unsigned *__restrict__ candT,
                                                  const unsigned *__restrict__ candP,
                                                  const unsigned *__restrict__ ncand, const unsigned *__restrict__ need,
                                                  unsigned *__restrict__ done, float rad, int spin, VaporState P,
                                                  unsigned *__restrict__ unfinished, int qMax, int ldsMax) {
  __shared__ unsigned sdone[kVcLds];
  const float rad2 = rad * rad;
  unsigned nb = *nBig;
  for (unsigned gi = blockIdx.x; gi < nb; gi += gridDim.x) {
    int t0 = (int)big[2 * gi], t1 = (int)big[2 * gi + 1];
    if (t1 - t0 <= qMax) continue;  // run from the ready queue (k_vc_sweepq)
    if (t1 - t0 <= ldsMax) {        // counts in LDS (ldsMax <= kVcLds)
      for (int q = threadIdx.x; q < t1 - t0; q += blockDim.x) sdone[q] = 0;
      __syncthreads();
      vc_sweep_group<true>(t0, t1, sdone, gk, xk, posOf, cand, candT, candP, ncand, need, rad, rad2, spin, P,
                           unfinished);
    } else {
      vc_sweep_group<false>(t0, t1, done, gk, xk, posOf, cand, candT, candP, ncand, need, rad, rad2, spin, P,
                            unfinished);
    }
    __syncthreads();
  }
}

// each parcel's group position (step index), the particle of each step and
// its partner count
__global__ void k_vc_tpos(int n, const unsigned long long *__restrict__ gk, const unsigned long long *__restrict__ xk,
                          const unsigned *__restrict__ posOf, const unsigned *__restrict__ ncand,
                          unsigned *__restrict__ tpos, unsigned *__restrict__ stepP, unsigned *__restrict__ stepN) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  unsigned p = member(gk, xk, t), e = posOf[p];
  tpos[e] = (unsigned)t;
  stepP[t] = p;
  stepN[t] = ncand[e];
}

// the successor of every step in each list it belongs to (the next step
// touching that parcel; all ones: none) and each step's number of
// predecessors, by step
__global__ void k_vc_succ(int np, const unsigned long long *__restrict__ keys, const unsigned *__restrict__ chainStart,
                          const unsigned long long *__restrict__ xk, const unsigned *__restrict__ posOf,
                          const unsigned *__restrict__ tpos, unsigned *__restrict__ succ, unsigned *__restrict__ depc) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  unsigned long long k = keys[s];
  unsigned x = (unsigned)(k >> 32), rank = (unsigned)((k >> 6) & 0x3ffffffull), slot = (unsigned)(k & 63ull);
  unsigned e = posOf[(unsigned)(xk[rank] & 0xffffffffull)];
  unsigned nxt = 0xffffffffu;
  if (s + 1 < np && (unsigned)(keys[s + 1] >> 32) == x) {
    unsigned rb = (unsigned)((keys[s + 1] >> 6) & 0x3ffffffull);
    nxt = tpos[posOf[(unsigned)(xk[rb] & 0xffffffffull)]];
  }
  unsigned t = tpos[e];
  succ[(long)t * (kVcK + 1) + slot] = nxt;
  if ((unsigned)s != chainStart[x]) atomicAdd(&depc[t], 1u);
}

__global__ void k_vc_cand_maps(int n, const unsigned *__restrict__ list, const unsigned *__restrict__ tpos,
                               const unsigned *__restrict__ cand, const unsigned *__restrict__ ncand,
                               unsigned *__restrict__ candT, unsigned *__restrict__ candP) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  long t = tpos[e];
  for (unsigned k = 0; k < ncand[e]; k++) {
    unsigned o = cand[(long)e * kVcK + k];
    candT[t * kVcK + k] = tpos[o];
    candP[t * kVcK + k] = list[o];
  }
}

// _iceContacts (solver.js:3290-3324). Snow does not move in this pass, so
// every pair test is fixed in advance; only the stuck flags and velocities
// evolve. Stuck grains are never written: a pair of two stuck grains does
// nothing and a stuck grain only turns its free partner stuck. Free grains
// are therefore grouped by their free-free contacts only, and each free
// grain carries the sweep position of its first contact with a stuck grain
// (the pair key (earlier x rank, later x rank)). One thread runs a group's
// free-free pairs in the reference's order and applies a member's first
// stuck contact just before the first pair that comes after it (it changes
// nothing but that member, so applying it late is the same), or at the end.

// which members of the ice subset are free (pAir != 1)
__global__ void k_ct_ice_free(int n, const unsigned *__restrict__ list, const float *__restrict__ air,
                              unsigned *__restrict__ freeMask) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) freeMask[e] = air[list[e]] != 1.f;
}

// the reference's pair test in double: the later grain (in x order) lies
// less than rad ahead in x, and the squared distance is below rad^2
__device__ __forceinline__ bool ice_pair(const float *px, const float *py, const float *pz, unsigned a, unsigned b,
                                         double rad, double rad2) {
  double ddx = (double)px[b] - (double)px[a];
  if (ddx >= rad) return false;
  double ddy = (double)py[b] - (double)py[a], ddz = (double)pz[b] - (double)pz[a];
  return ddx * ddx + ddy * ddy + ddz * ddz < rad2;
}

// per free grain: the key of its first contact with a stuck grain (all ones
// when it has none)
__global__ void k_ct_ice_stuckkey(Geo g, int n, const unsigned *__restrict__ list, const float *__restrict__ px,
                                  const float *__restrict__ py, const float *__restrict__ pz,
                                  const unsigned *__restrict__ cellOf, const unsigned *__restrict__ start,
                                  const unsigned *__restrict__ binned, int span, const unsigned *__restrict__ freeMask,
                                  const unsigned *__restrict__ rankOf, double rad, double rad2,
                                  unsigned long long *__restrict__ skey) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || !freeMask[e]) return;
  unsigned p = list[e], re = rankOf[e];
  int c = cellOf[e];
  int ci = c % g.nx, cj = (c / g.nx) % g.ny, ck = c / (g.nx * g.ny);
  unsigned long long best = ~0ull;
  for (int dk = -span; dk <= span; dk++)
    for (int dj = -span; dj <= span; dj++)
      for (int di = -span; di <= span; di++) {
        int ii = ci + di, jj = cj + dj, kk = ck + dk;
        if (ii < 0 || jj < 0 || kk < 0 || ii >= g.nx || jj >= g.ny || kk >= g.nz) continue;
        int cc = (kk * g.ny + jj) * g.nx + ii;
        for (unsigned b = start[cc]; b < start[cc + 1]; b++) {
          unsigned o = binned[b];
          if (freeMask[o]) continue;
          unsigned q = list[o], ro = rankOf[o];
          bool hit = re < ro ? ice_pair(px, py, pz, p, q, rad, rad2) : ice_pair(px, py, pz, q, p, rad, rad2);
          if (!hit) continue;
          unsigned long long k = re < ro ? ((unsigned long long)re << 32 | ro) : ((unsigned long long)ro << 32 | re);
          if (k < best) best = k;
        }
      }
  skey[e] = best;
}

__device__ __forceinline__ void ice_stick(float *vx, float *vy, float *vz, float *air, unsigned m) {
  air[m] = 1.f;
  vx[m] = vy[m] = vz[m] = 0.f;
}

// one group of free grains per thread; stuck members (key all ones) skip
__global__ void k_ct_ice_groups(int n, const unsigned long long *__restrict__ gk,
                                const unsigned long long *__restrict__ xk, const unsigned *__restrict__ ends,
                                const unsigned *__restrict__ posOf,
                                const unsigned long long *__restrict__ skey, double rad, double rad2,
                                const float *__restrict__ px, const float *__restrict__ py,
                                const float *__restrict__ pz, float *__restrict__ vx, float *__restrict__ vy,
                                float *__restrict__ vz, float *__restrict__ air) {
  int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= n) return;
  if (gk[t0] == ~0ull) return;
  if (t0 > 0 && (gk[t0] >> 32) == (gk[t0 - 1] >> 32)) return;  // not the first member of its group
  int t1 = (int)ends[gk[t0] >> 32];
  for (int a = t0; a < t1; a++) {
    unsigned i = member(gk, xk, a);
    unsigned ri = (unsigned)(gk[a] & 0xffffffffull);
    unsigned long long ski = skey[posOf[i]];
    for (int b = a + 1; b < t1; b++) {
      unsigned j = member(gk, xk, b);
      if ((double)px[j] - (double)px[i] >= rad) break;
      if (!ice_pair(px, py, pz, i, j, rad, rad2)) continue;
      unsigned long long K = ((unsigned long long)ri << 32) | (unsigned)(gk[b] & 0xffffffffull);
      // first stuck contacts that come before this pair
      if (air[i] != 1.f && ski < K) ice_stick(vx, vy, vz, air, i);
      if (air[j] != 1.f && skey[posOf[j]] < K) ice_stick(vx, vy, vz, air, j);
      bool stuckI = air[i] == 1.f, stuckJ = air[j] == 1.f;
      if (stuckI && stuckJ) continue;
      if (stuckJ) {
        ice_stick(vx, vy, vz, air, i);
      } else if (stuckI) {
        ice_stick(vx, vy, vz, air, j);
      } else {
        float mx = (vx[i] + vx[j]) * 0.5f, my = (vy[i] + vy[j]) * 0.5f, mz = (vz[i] + vz[j]) * 0.5f;
        vx[i] = vx[j] = mx; vy[i] = vy[j] = my; vz[i] = vz[j] = mz;
      }
    }
  }
  // first stuck contacts after the group's last pair
  for (int a = t0; a < t1; a++) {
    unsigned i = member(gk, xk, a);
    if (air[i] != 1.f && skey[posOf[i]] != ~0ull) ice_stick(vx, vy, vz, air, i);
  }
}

// vapor x liquid: each parcel meets the liquid particles near it in index
// order and is pushed out of each it overlaps (the water never moves). One
// wavefront per parcel: the lanes share the cells of the search box and
// gather the liquid particles within three radii (the parcel moves by less
// than one radius per contact) into LDS, the list is sorted by particle
// index (rank of each entry), and the contacts run in that order. Once the
// parcel has moved two radii from the gather's centre the next contact may
// lie outside it: the lanes gather again around the parcel's position (the
// particles after the last one visited) and the walk goes on. Three radii
// hold about 25 particles at 8 per cell, but compressed water holds more (64
// was exceeded in a 128^3 / 2M run); beyond kVlMax the overflow counter is
// raised and one lane gathers the first kVlMax in cell order instead, so the
// result stays deterministic.
// Each pair is tested directly (x window, then distance). The reference
// finds a liquid particle's parcels by a binary search over the vapor list
// sorted by x before the vapor-vapor pushes (solver.js:3449-3451); a parcel
// pushed along x past its neighbours in that stale order can be skipped
// there and is met here.
constexpr int kVlMax = 1024;
constexpr int kVlWaves = 4;     // parcels per 256-thread workgroup
constexpr int kVlRounds = 64;   // gathers per parcel before the overflow counter is raised

__global__ __launch_bounds__(256) void k_ct_vapor_liquid(Geo g, int nV, const unsigned *__restrict__ vlist,
                                                         const unsigned *__restrict__ llist,
                                                         const unsigned *__restrict__ lstart,
                                                         const Slot *__restrict__ lslot, float rad,
                                                         float *__restrict__ px, float *__restrict__ py,
                                                         float *__restrict__ pz, float *__restrict__ vx,
                                                         float *__restrict__ vy, float *__restrict__ vz,
                                                         unsigned *__restrict__ overflow) {
  __shared__ unsigned scand[kVlWaves][kVlMax], ssort[kVlWaves][kVlMax];
  __shared__ unsigned sn[kVlWaves], safter[kVlWaves], sagain[kVlWaves];
  __shared__ float spos[kVlWaves][3];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e = blockIdx.x * kVlWaves + wv;
  if (e >= nV) return;  // the whole wavefront leaves; no workgroup barrier follows
  const unsigned j = vlist[e];
  const unsigned kNone = 0xffffffffu;
  const float reach = 3.f * rad, reach2 = reach * reach;
  // the parcel's state, held by lane 0 across gathers
  float jx = px[j], jy = py[j], jz = pz[j], ux = vx[j], uy = vy[j], uz = vz[j];
  if (lane == 0) { safter[wv] = kNone; spos[wv][0] = jx; spos[wv][1] = jy; spos[wv][2] = jz; }
  for (int round = 0; round < kVlRounds; round++) {
    if (lane == 0) sn[wv] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // gather around the parcel's position at the start of this round the
    // liquid particles after the last one already visited (a push moves the
    // parcel by less than one radius, so three radii hold every contact
    // until the parcel has moved two radii from this centre)
    const float x = spos[wv][0], y = spos[wv][1], z = spos[wv][2];
    const unsigned after = safter[wv];
    int ci = clampi((int)(x * g.inv_dx), 0, g.nx - 1), cj = clampi((int)(y * g.inv_dx), 0, g.ny - 1);
    int ck = clampi((int)(z * g.inv_dx), 0, g.nz - 1);
    int span = (int)ceilf(reach * g.inv_dx);
    const int ia = max(ci - span, 0), ib = min(ci + span, g.nx - 1);
    const int ja = max(cj - span, 0), jb = min(cj + span, g.ny - 1);
    const int ka = max(ck - span, 0), kb = min(ck + span, g.nz - 1);
    const int ni = ib - ia + 1, nj = jb - ja + 1, nk = kb - ka + 1;
    for (int cidx = lane; cidx < ni * nj * nk; cidx += 64) {
      int ii = ia + cidx % ni, jj = ja + (cidx / ni) % nj, kk = ka + cidx / (ni * nj);
      int cc = (kk * g.ny + jj) * g.nx + ii;
      for (unsigned b = lstart[cc]; b < lstart[cc + 1]; b++) {
        const Slot s = lslot[b];
        const unsigned pi = llist[s.o];
        if (after != kNone && pi <= after) continue;
        float ddx = s.x - x, ddy = s.y - y, ddz = s.z - z;
        if (ddx * ddx + ddy * ddy + ddz * ddz >= reach2) continue;
        unsigned at = atomicAdd(&sn[wv], 1u);
        if (at < (unsigned)kVlMax) scand[wv][at] = pi;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const unsigned found = sn[wv];
    const int nc = found < (unsigned)kVlMax ? (int)found : kVlMax;
    if (found > (unsigned)kVlMax && lane == 0) {
      atomicAdd(overflow, found - (unsigned)kVlMax);
      // the first kVlMax in cell and bin order (arrival order is not fixed)
      int m = 0;
      for (int cidx = 0; cidx < ni * nj * nk && m < kVlMax; cidx++) {
        int ii = ia + cidx % ni, jj = ja + (cidx / ni) % nj, kk = ka + cidx / (ni * nj);
        int cc = (kk * g.ny + jj) * g.nx + ii;
        for (unsigned b = lstart[cc]; b < lstart[cc + 1] && m < kVlMax; b++) {
          const Slot s = lslot[b];
          const unsigned pi = llist[s.o];
          if (after != kNone && pi <= after) continue;
          float ddx = s.x - x, ddy = s.y - y, ddz = s.z - z;
          if (ddx * ddx + ddy * ddy + ddz * ddz < reach2) scand[wv][m++] = pi;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // rank sort of the (distinct) particle indices
    for (int q = lane; q < nc; q += 64) {
      unsigned v = scand[wv][q];
      int rank = 0;
      for (int m = 0; m < nc; m++) rank += scand[wv][m] < v ? 1 : 0;
      ssort[wv][rank] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) {
      const float rad2 = rad * rad, drift = reach - rad, drift2 = drift * drift;
      unsigned again = 0, last = after;
      for (int k = 0; k < nc; k++) {
        // moved two radii from this round's centre: the next contacts may lie
        // outside the gather, so gather again around the current position
        float mx = jx - x, my = jy - y, mz = jz - z;
        if (mx * mx + my * my + mz * mz > drift2) { again = 1; break; }
        unsigned p = ssort[wv][k];
        last = p;
        float wx = px[p], wy = py[p], wz = pz[p];
        // the reference's x window: binary search to x >= wx - rad, stop at x - wx >= rad
        if (jx < wx - rad) continue;
        float ddx = jx - wx;
        if (ddx >= rad) continue;
        float ddy = jy - wy, ddz = jz - wz;
        float dd2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (dd2 >= rad2 || dd2 < 1e-12f) continue;
        float dd = sqrtf(dd2);
        float nX = ddx / dd, nY = ddy / dd, nZ = ddz / dd;
        float vR = (ux - vx[p]) * nX + (uy - vy[p]) * nY + (uz - vz[p]) * nZ;
        if (vR < 0.f) {
          float vB = vR * 1.4f;
          ux -= vB * nX; uy -= vB * nY; uz -= vB * nZ;
        }
        float ov = rad - dd;
        jx += nX * ov; jy += nY * ov; jz += nZ * ov;
        float kk2 = ov * 2.f;
        ux += nX * kk2; uy += nY * kk2; uz += nZ * kk2;
      }
      if (again && round + 1 == kVlRounds) atomicAdd(overflow, 1u);
      sagain[wv] = again;
      safter[wv] = last;
      spos[wv][0] = jx; spos[wv][1] = jy; spos[wv][2] = jz;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!sagain[wv]) break;
  }
  if (lane == 0) {
    px[j] = jx; py[j] = jy; pz[j] = jz;
    vx[j] = ux; vy[j] = uy; vz[j] = uz;
  }
}

__global__ void k_ct_census(Geo g, int nV, const unsigned *__restrict__ vlist, const uint8_t *__restrict__ pflag,
                            const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
                            int hasSun, float sx, float sy, float sz, unsigned *__restrict__ counts) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nV) return;
  unsigned p = vlist[e];
  if (pflag[p] != 2) return;
  atomicAdd(&counts[0], 1u);
  if (!hasSun || (px[p] - g.cx) * sx + (py[p] - g.cy) * sy + (pz[p] - g.cz) * sz < 0.f) atomicAdd(&counts[1], 1u);
}

__global__ void k_ct_iota(int n, unsigned *__restrict__ a) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = (unsigned)e;
}

__global__ void k_ct_zero(long n, unsigned *__restrict__ a) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = 0u;
}

}  // namespace

// scratch of the contact passes (allocated on first use, grown with nP)
struct Contacts {
  long cap = 0;
  unsigned *fv = nullptr, *fi = nullptr, *fl = nullptr, *ov = nullptr, *oi = nullptr, *ol = nullptr;
  unsigned *vlist = nullptr, *ilist = nullptr, *llist = nullptr;
  unsigned *cellOf = nullptr, *rankIn = nullptr, *binned = nullptr, *parent = nullptr, *posOf = nullptr,
           *rankOf = nullptr, *gend = nullptr;
  unsigned *cnt = nullptr, *start = nullptr, *lstart = nullptr, *lbinned = nullptr, *lcell = nullptr,
           *lrank = nullptr, *counts = nullptr;
  unsigned long long *xk = nullptr, *xk2 = nullptr, *gk = nullptr, *gk2 = nullptr, *slotKey = nullptr;
  Slot *slot = nullptr, *lslot = nullptr;
  void *tmp = nullptr;
  size_t tmpBytes = 0;
  // dependency-ordered sweep of large vapor groups, sized by the vapor count
  long vcap = 0;
  unsigned *cand = nullptr, *ncand = nullptr, *pairOff = nullptr, *need = nullptr, *done = nullptr,
           *serialOnly = nullptr, *big = nullptr, *vcCounters = nullptr, *chainStart = nullptr, *candT = nullptr,
           *candP = nullptr, *tpos = nullptr, *stepP = nullptr, *stepN = nullptr, *succ = nullptr, *depc = nullptr;
  unsigned *gbig = nullptr;  // per group root: 1 if the group is larger than serialMax
  unsigned long long *pairKeys = nullptr, *pairKeys2 = nullptr;
  void *vtmp = nullptr;
  size_t vtmpBytes = 0;
};

namespace {

template <typename T>
int calloc_dev(ws_ctx *c, T **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess) { c->err = std::string("hipMalloc (contacts): ") + hipGetErrorString(e); return WS_EHIP; }
  return WS_OK;
}

int contacts_ensure(ws_ctx *c) {
  Contacts *k = c->contacts;
  long n = c->capacity + 1;
  if (k && k->cap >= n) return WS_OK;
  if (!k) { k = new Contacts(); c->contacts = k; }
  contacts_release(c, false);
  long C = c->nC + 1;
  int r;
  if ((r = calloc_dev(c, &k->fv, n)) || (r = calloc_dev(c, &k->fi, n)) || (r = calloc_dev(c, &k->fl, n)) ||
      (r = calloc_dev(c, &k->ov, n)) || (r = calloc_dev(c, &k->oi, n)) || (r = calloc_dev(c, &k->ol, n)) ||
      (r = calloc_dev(c, &k->vlist, n)) || (r = calloc_dev(c, &k->ilist, n)) || (r = calloc_dev(c, &k->llist, n)) ||
      (r = calloc_dev(c, &k->cellOf, n)) || (r = calloc_dev(c, &k->rankIn, n)) || (r = calloc_dev(c, &k->binned, n)) ||
      (r = calloc_dev(c, &k->parent, n)) || (r = calloc_dev(c, &k->posOf, n)) || (r = calloc_dev(c, &k->rankOf, n)) ||
      (r = calloc_dev(c, &k->gend, n)) ||
      (r = calloc_dev(c, &k->cnt, C)) || (r = calloc_dev(c, &k->start, C)) || (r = calloc_dev(c, &k->lstart, C)) ||
      (r = calloc_dev(c, &k->lbinned, n)) || (r = calloc_dev(c, &k->lcell, n)) || (r = calloc_dev(c, &k->lrank, n)) ||
      (r = calloc_dev(c, &k->counts, 4)) || (r = calloc_dev(c, &k->xk, n)) || (r = calloc_dev(c, &k->xk2, n)) ||
      (r = calloc_dev(c, &k->gk, n)) || (r = calloc_dev(c, &k->gk2, n)) || (r = calloc_dev(c, &k->slotKey, n)) ||
      (r = calloc_dev(c, &k->slot, n)) || (r = calloc_dev(c, &k->lslot, n)))
    return r;
  size_t a = 0, b = 0, d = 0;
  WS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, a, k->fv, k->ov, (int)n, c->stream));
  WS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, k->cnt, k->start, (int)C, c->stream));
  WS_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, d, k->xk, k->xk2, (int)n, 0, 64, c->stream));
  k->tmpBytes = a > b ? a : b;
  if (d > k->tmpBytes) k->tmpBytes = d;
  if ((r = calloc_dev(c, (char **)&k->tmp, k->tmpBytes))) return r;
  k->cap = n;
  return WS_OK;
}

int bin_subset(ws_ctx *c, int n, const unsigned *list, unsigned *cellOf, unsigned *rankIn, unsigned *start,
               unsigned *binned, Slot *slot) {
  Contacts *k = c->contacts;
  long C = c->nC;
  hipLaunchKernelGGL(k_ct_zero, dim3(blocks_for(C + 1, 256)), dim3(256), 0, c->stream, C + 1, k->cnt);
  if (n > 0)
    hipLaunchKernelGGL(k_ct_count, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, c->geo, n, list, c->px, c->py,
                       c->pz, k->cnt, cellOf, rankIn);
  WS_LAUNCHED();
  WS_HIP(hipcub::DeviceScan::ExclusiveSum(k->tmp, k->tmpBytes, k->cnt, start, (int)C + 1, c->stream));
  if (n > 0)
    hipLaunchKernelGGL(k_ct_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, list, c->px, c->py, c->pz,
                       cellOf, rankIn, start, binned, slot);
  WS_LAUNCHED();
  return WS_OK;
}

// neighbour cells to visit for a search radius
int cell_span(ws_ctx *c, double r) {
  int s = (int)ceil(r / c->dx);
  return s < 1 ? 1 : s;
}

// sorted x order of a subset, its connected groups (pairs closer than thr)
// and the group keys sorted by (group, x rank); leaves the sorted x keys in
// xk and the sorted group keys in gk
int group_subset(ws_ctx *c, int n, const unsigned *list, float thr, const unsigned *mask) {
  Contacts *k = c->contacts;
  int nb = blocks_for(n, 256);
  hipLaunchKernelGGL(k_ct_keys, dim3(nb), dim3(256), 0, c->stream, n, list, c->px, k->xk2);
  WS_LAUNCHED();
  WS_HIP(hipcub::DeviceRadixSort::SortKeys(k->tmp, k->tmpBytes, k->xk2, k->xk, n, 0, 64, c->stream));
  hipLaunchKernelGGL(k_ct_pos, dim3(nb), dim3(256), 0, c->stream, n, list, k->posOf);
  hipLaunchKernelGGL(k_ct_rank, dim3(nb), dim3(256), 0, c->stream, n, k->xk, k->posOf, k->rankOf);
  WS_LAUNCHED();
  int r;
  if ((r = bin_subset(c, n, list, k->cellOf, k->rankIn, k->start, k->binned, k->slot))) return r;
  hipLaunchKernelGGL(k_ct_iota, dim3(nb), dim3(256), 0, c->stream, n, k->parent);
  hipLaunchKernelGGL(k_ct_union, dim3((n + 3) / 4), dim3(256), 0, c->stream, c->geo, n, list, c->px, c->py, c->pz,
                     k->cellOf,
                     k->start, k->slot, thr * thr, cell_span(c, thr), mask, k->parent);
  hipLaunchKernelGGL(k_ct_groupkeys, dim3(nb), dim3(256), 0, c->stream, n, k->parent, k->rankOf, mask, k->cellOf,
                     k->rankIn, k->start, k->gk2, k->slotKey);
  WS_LAUNCHED();
  WS_HIP(hipcub::DeviceRadixSort::SortKeys(k->tmp, k->tmpBytes, k->gk2, k->gk, n, 0, 64, c->stream));
  hipLaunchKernelGGL(k_ct_group_ends, dim3(nb), dim3(256), 0, c->stream, n, k->gk, k->gend);
  WS_LAUNCHED();
  return WS_OK;
}

// scratch of the dependency-ordered vapor sweep for nV parcels (grow-only)
int vapor_ensure(ws_ctx *c, long nV) {
  Contacts *k = c->contacts;
  if (k->vcap >= nV) return WS_OK;
  long n = nV + nV / 4 + 1024;
  void *old[] = {k->cand, k->ncand, k->pairOff, k->need, k->done, k->serialOnly, k->big, k->vcCounters,
                 k->chainStart, k->pairKeys, k->pairKeys2, k->vtmp, k->candT, k->candP, k->tpos, k->stepP,
                 k->stepN, k->succ, k->depc, k->gbig};
  for (void *p : old)
    if (p) (void)hipFree(p);
  k->vcap = 0;
  long np = n * (kVcK + 1);
  int r;
  if ((r = calloc_dev(c, &k->cand, n * kVcK)) || (r = calloc_dev(c, &k->ncand, n + 1)) ||
      (r = calloc_dev(c, &k->pairOff, n + 1)) || (r = calloc_dev(c, &k->need, np)) ||
      (r = calloc_dev(c, &k->done, n)) || (r = calloc_dev(c, &k->serialOnly, n)) ||
      (r = calloc_dev(c, &k->big, 2 * n)) || (r = calloc_dev(c, &k->vcCounters, 4)) ||
      (r = calloc_dev(c, &k->chainStart, n)) || (r = calloc_dev(c, &k->pairKeys, np)) ||
      (r = calloc_dev(c, &k->pairKeys2, np)) || (r = calloc_dev(c, &k->candT, n * kVcK)) ||
      (r = calloc_dev(c, &k->candP, n * kVcK)) || (r = calloc_dev(c, &k->tpos, n)) ||
      (r = calloc_dev(c, &k->stepP, n)) || (r = calloc_dev(c, &k->stepN, n)) || (r = calloc_dev(c, &k->succ, np)) || (r = calloc_dev(c, &k->depc, n)) ||
      (r = calloc_dev(c, &k->gbig, n)))
    return r;
  size_t a = 0, b = 0;
  WS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, a, k->ncand, k->pairOff, (int)(n + 1), c->stream));
  WS_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b, k->pairKeys2, k->pairKeys, (int)np, 0, 64, c->stream));
  k->vtmpBytes = a > b ? a : b;
  if ((r = calloc_dev(c, (char **)&k->vtmp, k->vtmpBytes))) return r;
  k->vcap = n;
  return WS_OK;
}

int subset(ws_ctx *c, unsigned *flag, unsigned *off, unsigned *list, int *count) {
  Contacts *k = c->contacts;
  int n = c->nP;
  WS_HIP(hipcub::DeviceScan::ExclusiveSum(k->tmp, k->tmpBytes, flag, off, n + 1, c->stream));
  hipLaunchKernelGGL(k_ct_compact, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, flag, off, list);
  WS_LAUNCHED();
  unsigned h = 0;
  WS_HIP(hipMemcpyAsync(&h, off + n, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  *count = (int)h;
  return WS_OK;
}

int flags(ws_ctx *c) {
  Contacts *k = c->contacts;
  double band = c->oceanR - c->dx * 2;
  hipLaunchKernelGGL(k_ct_flags, dim3(blocks_for(c->nP + 1, 256)), dim3(256), 0, c->stream, c->geo, c->nP, c->px,
                     c->py, c->pz, c->pflag, (float)(band * band), k->fv, k->fi, k->fl);
  WS_LAUNCHED();
  return WS_OK;
}

}  // namespace

void contacts_release(ws_ctx *c, bool all) {
  Contacts *k = c->contacts;
  if (!k) return;
  void *ptrs[] = {k->fv, k->fi, k->fl, k->ov, k->oi, k->ol, k->vlist, k->ilist, k->llist, k->cellOf, k->rankIn,
                  k->binned, k->parent, k->posOf, k->rankOf, k->gend, k->cnt, k->start, k->lstart, k->lbinned, k->lcell,
                  k->lrank, k->counts, k->xk, k->xk2, k->gk, k->gk2, k->tmp, k->cand, k->ncand, k->pairOff,
                  k->need, k->done, k->serialOnly, k->big, k->vcCounters, k->chainStart, k->pairKeys,
                  k->pairKeys2, k->vtmp, k->candT, k->candP, k->tpos, k->stepP, k->stepN, k->succ,
                  k->depc, k->slotKey, k->slot, k->lslot};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (all) {
    delete k;
    c->contacts = nullptr;
  } else {
    *k = Contacts();
  }
}

// _vaporCollisions, solver.js:3345-3479
int frame_vapor_collisions(ws_ctx *c) {
  if (c->nP < 1) return WS_OK;
  int r;
  if ((r = contacts_ensure(c)) || (r = flags(c))) return r;
  Contacts *k = c->contacts;
  int nV = 0;
  if ((r = subset(c, k->fv, k->ov, k->vlist, &nV))) return r;
  hipLaunchKernelGGL(k_ct_zero, dim3(1), dim3(64), 0, c->stream, 4L, k->counts);
  if (nV > 0) {
    double sx = c->sunX - c->cx, sy = c->sunY - c->cy, sz = c->sunZ - c->cz;
    hipLaunchKernelGGL(k_ct_census, dim3(blocks_for(nV, 256)), dim3(256), 0, c->stream, c->geo, nV, k->vlist, c->pflag,
                       c->px, c->py, c->pz, c->hasSun, (float)sx, (float)sy, (float)sz, k->counts);
  }
  WS_LAUNCHED();
  if (nV >= 2) {
    float rad = (float)(c->spacing * 0.6);
    if ((r = group_subset(c, nV, k->vlist, 2.f * rad, nullptr))) return r;
    if ((r = vapor_ensure(c, nV))) return r;
    VaporState P{c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pWx, c->pWy, c->pWz};
    int nb = blocks_for(nV, 256);
    int serialMax = c->vaporSerialMax;
    // groups too large for one thread's serial sweep
    hipLaunchKernelGGL(k_ct_zero, dim3(blocks_for(nV, 256)), dim3(256), 0, c->stream, (long)nV, k->serialOnly);
    hipLaunchKernelGGL(k_ct_zero, dim3(1), dim3(64), 0, c->stream, 4L, k->vcCounters);
    hipLaunchKernelGGL(k_vc_big, dim3(nb), dim3(256), 0, c->stream, nV, k->gk, k->gend, serialMax, k->serialOnly, k->big,
                       k->vcCounters);
    WS_LAUNCHED();
    unsigned nBig = 0;
    WS_HIP(hipMemcpyAsync(&nBig, k->vcCounters, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    WS_HIP(hipStreamSynchronize(c->stream));
    if (nBig > 0) {
      // partners, the steps touching each parcel in sweep order, positions
      hipLaunchKernelGGL(k_vc_mark_big, dim3(nb), dim3(256), 0, c->stream, nV, k->gk, k->gend, serialMax, k->gbig);
      hipLaunchKernelGGL(k_vc_cand, dim3((nV + 3) / 4), dim3(256), 0, c->stream, c->geo, nV, k->vlist, k->gbig, c->px,
                         c->py, c->pz,
                         k->cellOf, k->start, k->slot, k->slotKey, cell_span(c, 2.0 * rad), k->gk2,
                         4.f * rad * rad, k->cand, k->ncand, k->serialOnly);
      hipLaunchKernelGGL(k_vc_pair_count, dim3(blocks_for(nV + 1, 256)), dim3(256), 0, c->stream, nV, k->ncand,
                         k->pairOff);
      WS_LAUNCHED();
      WS_HIP(hipcub::DeviceScan::ExclusiveSum(k->vtmp, k->vtmpBytes, k->pairOff, k->pairOff, nV + 1, c->stream));
      unsigned np = 0;
      WS_HIP(hipMemcpyAsync(&np, k->pairOff + nV, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
      hipLaunchKernelGGL(k_vc_pair_emit, dim3(nb), dim3(256), 0, c->stream, nV, k->rankOf, k->cand, k->ncand,
                         k->pairOff, k->pairKeys2);
      WS_LAUNCHED();
      WS_HIP(hipStreamSynchronize(c->stream));
      WS_HIP(hipcub::DeviceRadixSort::SortKeys(k->vtmp, k->vtmpBytes, k->pairKeys2, k->pairKeys, (int)np, 0, 64,
                                               c->stream));
      hipLaunchKernelGGL(k_vc_chain_heads, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, (int)np, k->pairKeys,
                         k->chainStart);
      hipLaunchKernelGGL(k_vc_need, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, (int)np, k->pairKeys,
                         k->chainStart, k->xk, k->posOf, k->need);
      hipLaunchKernelGGL(k_ct_zero, dim3(nb), dim3(256), 0, c->stream, (long)nV, k->done);
      hipLaunchKernelGGL(k_vc_tpos, dim3(nb), dim3(256), 0, c->stream, nV, k->gk, k->xk, k->posOf, k->ncand, k->tpos,
                         k->stepP, k->stepN);
      hipLaunchKernelGGL(k_ct_zero, dim3(nb), dim3(256), 0, c->stream, (long)nV, k->depc);
      hipLaunchKernelGGL(k_vc_succ, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, (int)np, k->pairKeys,
                         k->chainStart, k->xk, k->posOf, k->tpos, k->succ, k->depc);
      hipLaunchKernelGGL(k_vc_cand_maps, dim3(nb), dim3(256), 0, c->stream, nV, k->vlist, k->tpos, k->cand, k->ncand,
                         k->candT, k->candP);
      // the large groups again, now without those marked for the serial sweep
      hipLaunchKernelGGL(k_ct_zero, dim3(1), dim3(64), 0, c->stream, 4L, k->vcCounters);
      hipLaunchKernelGGL(k_vc_big, dim3(nb), dim3(256), 0, c->stream, nV, k->gk, k->gend, serialMax, k->serialOnly, k->big,
                         k->vcCounters);
      WS_LAUNCHED();
    }
    hipLaunchKernelGGL(k_ct_vapor_groups, dim3(nb), dim3(256), 0, c->stream, nV, k->gk, k->xk, k->gend, rad, c->spinOn, P,
                       serialMax, k->serialOnly);
    if (nBig > 0) {
      int grid = (int)(nBig < 2048 ? nBig : 2048);
      hipLaunchKernelGGL(k_vc_sweepq, dim3(grid), dim3(1024), 0, c->stream, k->big, k->vcCounters, k->stepP, k->stepN,
                         k->candP, k->succ, k->depc, rad, c->spinOn, P, k->vcCounters + 1,
                         std::min(std::max(c->vaporQueueMax, 0), kVcQ));
      hipLaunchKernelGGL(k_vc_sweep, dim3(grid), dim3(256), 0, c->stream, k->big, k->vcCounters, k->gk, k->xk,
                         k->posOf, k->cand, k->candT, k->candP, k->ncand, k->need, k->done, rad, c->spinOn, P,
                         k->vcCounters + 1, std::min(std::max(c->vaporQueueMax, 0), kVcQ),
                         std::min(std::max(c->vaporLdsMax, 0), kVcLds));
    }
    WS_LAUNCHED();
    c->vaporBigGroups = (long)nBig;
    if (nBig > 0) {
      unsigned vc[4];
      WS_HIP(hipMemcpyAsync(vc, k->vcCounters, sizeof vc, hipMemcpyDeviceToHost, c->stream));
      WS_HIP(hipStreamSynchronize(c->stream));
      unsigned unfinished = vc[1];
      c->vaporGroupMax = (long)vc[2];
      c->vaporBigParcels = (long)vc[3];
      if (unfinished) {
        c->err = "vapor contact sweep: a group did not finish within its step bound";
        return WS_ESTATE;
      }
    }
    // liquid partners, binned by cell (they do not move in this pass)
    int nL = 0;
    if ((r = subset(c, k->fl, k->ol, k->llist, &nL))) return r;
    if ((r = bin_subset(c, nL, k->llist, k->lcell, k->lrank, k->lstart, k->lbinned, k->lslot))) return r;
    hipLaunchKernelGGL(k_ct_vapor_liquid, dim3((nV + kVlWaves - 1) / kVlWaves), dim3(256), 0, c->stream, c->geo, nV,
                       k->vlist,
                       k->llist, k->lstart, k->lslot, rad, c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz,
                       k->counts + 2);
    WS_LAUNCHED();
  }
  unsigned h[4];
  WS_HIP(hipMemcpyAsync(h, k->counts, sizeof h, hipMemcpyDeviceToHost, c->stream));
  WS_HIP(hipStreamSynchronize(c->stream));
  c->vaporCount = (int)h[0];
  c->nightVaporCount = (int)h[1];
  c->contactOverflow += h[2];
  return WS_OK;
}

// _iceContacts, solver.js:3290-3324
int frame_ice_contacts(ws_ctx *c) {
  if (c->nP < 1) return WS_OK;
  int r;
  if ((r = contacts_ensure(c)) || (r = flags(c))) return r;
  Contacts *k = c->contacts;
  int nI = 0;
  if ((r = subset(c, k->fi, k->oi, k->ilist, &nI))) return r;
  if (nI < 2) return WS_OK;
  double rad = c->spacing * 0.7, rad2 = rad * rad;
  int nb = blocks_for(nI, 256);
  // free grains, grouped by free-free contacts from a radius a hair above
  // the contact radius, so that no pair the double test accepts is split
  hipLaunchKernelGGL(k_ct_ice_free, dim3(nb), dim3(256), 0, c->stream, nI, k->ilist, c->pAir, k->fi);
  WS_LAUNCHED();
  if ((r = group_subset(c, nI, k->ilist, (float)rad * (1.f + 1e-5f), k->fi))) return r;
  hipLaunchKernelGGL(k_ct_ice_stuckkey, dim3(nb), dim3(256), 0, c->stream, c->geo, nI, k->ilist, c->px, c->py, c->pz,
                     k->cellOf, k->start, k->binned, cell_span(c, rad), k->fi, k->rankOf, rad, rad2, k->gk2);
  hipLaunchKernelGGL(k_ct_ice_groups, dim3(nb), dim3(256), 0, c->stream, nI, k->gk, k->xk, k->gend, k->posOf, k->gk2, rad,
                     rad2, c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pAir);
  WS_LAUNCHED();
  return WS_OK;
}

}  // namespace ws
