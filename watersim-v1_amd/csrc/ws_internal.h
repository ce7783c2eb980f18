// Internal definitions of the watersim HIP library (gfx950): the context,
// device-side scalar block, launch helpers and the device functions shared by
// the grid, particle and surface kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/watersim.h"

namespace ws {

enum : uint8_t { AIR = 0, FLUID = 1, SOLID = 2 };

// P2G and the density splat accumulate in 64-bit fixed point with 50
// fractional bits: exact and order independent. A face sums at most a few
// hundred contributions of |w v| <= 3.6, far below the 2^13 headroom.
constexpr double kFixScale = 1125899906842624.0;  // 2^50
constexpr float kFixScaleF = 1125899906842624.0f;
constexpr float kFixInvF = 1.0f / 1125899906842624.0f;

// Tile shapes (cells) for P2G and the extrapolation wavefront.
// P2G bins (cells per tile along i, j, k), chosen by the context's p2gBin
// (the autotuner times them; every bin gives the same sums)
constexpr int P2G_BINS[4][3] = {{8, 8, 8}, {4, 4, 4}, {16, 8, 8}, {8, 8, 4}};
constexpr int P2G_NBINS = 4;
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
// extrapolation tile: 64 faces along i (one lane each), 8 x 8 rows in j and
// k; one wavefront per k-plane of the tile
constexpr int EX_TI = 64, EX_TJ = 8, EX_TK = 8;
constexpr int EX_THREADS = 64 * EX_TK;
// the row-pipelined extrapolation: k-planes per workgroup (one wavefront each)
constexpr int EXR_TK = 8;
constexpr int EXR_THREADS = 64 * EXR_TK;
constexpr int SOR_TX = 32, SOR_TY = 8, SOR_TZ = 8;
constexpr int SOR_TB_X = 32, SOR_TB_Y = 16, SOR_TB_Z = 16;  // temporal-blocking tile
constexpr int SOR_TB_THREADS = 512;
// The four colour phases of a two-sweep temporal-blocking launch: phase ph
// relaxes one colour (red first) of the tile plus a ring of 3 - ph cells.
// Their cells are listed per tile, at most tb_items(ph) each, at offset
// tb_off(ph) of the tile's TB_CAP entries.
__host__ __device__ constexpr int tb_ring(int ph) { return 3 - ph; }
__host__ __device__ constexpr int tb_items(int ph) {
  return ((SOR_TB_X + 2 * tb_ring(ph) + 1) / 2) * (SOR_TB_Y + 2 * tb_ring(ph)) * (SOR_TB_Z + 2 * tb_ring(ph));
}
__host__ __device__ constexpr int tb_off(int ph) { return ph <= 0 ? 0 : tb_off(ph - 1) + tb_items(ph - 1); }
constexpr int TB_CAP = tb_off(4);
constexpr int TB_H = 4;  // q halo of the tile in LDS

// Evaporation chooser: candidates per window, threads of its one workgroup,
// and the ejection counts per window tested at once (each ejection moves
// the stream by six draws)
constexpr int EV_WIN = 4096, EV_THREADS = 1024, EV_HYP = 16;

// Scalars that live on the device so that substeps can run back to back
// without host round trips (dt law, gravity ramp, reductions).
struct DevScalars {
  float dt;            // current substep length
  float gs;            // gravity scale of this substep (ramp * gravity), 0 = off
  double gRamp;        // gravity ramp timer (solver.js:2657-2658)
  double simTime;
  unsigned umax_bits;  // max |face velocity| after the clamp, float bits
  unsigned airborne;   // airborneCount of the last G2P
  unsigned df_refined; // demotion predicates decided by the float-float path
  unsigned nan_seen;
  unsigned nRed, nBlack;  // fluid cells of each colour (pressure lists)
  double rr, bb;       // residual^2, rhs^2 (FP64 norms)
  double ke;
  double vol;          // marching-tets volume (FP64)
  unsigned long long mt_count;
  float dtLast;
  float pad;
  double dtD;          // the substep length as the reference holds it (double)
  // the on-device solve (k_sor_solve): sweeps and residual of the last
  // solve, running totals of sweeps and residual evaluations
  unsigned sorIters, sorChecksLast;
  double sorRes;
  unsigned long long sorSweeps, sorChecks;
  // XCD-resident sweeps (k_sor_xcd): launches that declined (workgroups
  // not on one XCD, or a range larger than its LDS; q untouched), and a
  // sticky flag set when one of their waits gave up, and the declines
  // whose cause was a range or halo larger than the workgroup holds
  unsigned xcdDeclined, xcdTimedOut, xcdTooBig;
  // the row-pipelined extrapolation: 1 a poll gave up, 2 a component's jobs
  // were not all done (sticky until the host reads and clears it)
  unsigned exFailed;
  unsigned scanFailed;  // a single-launch scan gave up waiting for an earlier tile
};

// Geometry and per-substep constants passed by value to kernels.
struct Geo {
  int nx, ny, nz;
  int uN, vN, wN, nC;
  float dx, inv_dx;
  float dx_lo;              // double dx minus its float rounding
  float cx, cy, cz;         // domain centre (float)
  float cx_lo, cy_lo, cz_lo;  // double centre minus its float rounding
};

// words between two workgroups' progress counters in the neighbour-
// synchronised solve (one 128-byte line each)
constexpr int kFlowStride = 32;

// n red-black sweeps of the fluid lists in one launch on the workgroups of
// one XCD (k_sor_xcd); progress counters and placement reports share the
// flow buffer (word 0 and word 1 of each workgroup's line)
constexpr int kXcdMaxG = 64;
struct XcdSweep {
  Geo g;
  DevScalars *ds;
  const unsigned *code;
  const float *lrhs, *lpw;
  const unsigned long long *off;  // the lists' scan keys (nC + 1 entries)
  float *q;
  long half;
  float om1;
  int G, sweeps;
  unsigned epoch;  // every tag of this launch is at least epoch; earlier launches' are below it
  unsigned *flow;  // per workgroup: arrival counter (word 0) and placement report (word 1)
  unsigned long long *xq;  // per cell: a boundary cell's value and half-sweep tag
};

// One pressure solve run to its end by one workgroup (k_sor_solve): the
// fluid lists of a context, its q, and the stopping rule of the host loop
struct SorJob {
  Geo g;
  DevScalars *ds;
  const unsigned *code;
  const float *lrhs, *lpw;
  const uint8_t *ldg;
  float *q;
  long half;
  float om1;
  int every, maxIters, adaptive, fixedIters, measure;
  double tol;        // <= 0: exactly fixedIters sweeps
  double *chunk;     // residual chunk sums, 2 x ceil(nC / 2048)
  unsigned *arrive;  // G > 1: barrier arrivals, then a timed-out flag
  int G;             // workgroups sharing this solve
  // batched solve, G > 1: G + 1 words, non-null when the G workgroups were
  // laid out on one XCD (each writes its XCD id to word w; word G records
  // whether they met there and took the L2-local barriers)
  unsigned *place = nullptr;
  // the neighbour-synchronised solve (k_sor_flow): the lists' scan keys
  // (red count | black count << 32 before each cell, nC + 1 entries) and one
  // progress counter per workgroup, kFlowStride words apart
  const unsigned long long *off;
  unsigned *flow;
};

// one launch of the row-pipelined extrapolation (k_extrap_rows): per
// component a ticket counter and a count of finished jobs, the exchange
// arrays (per component, slab and chunk: the top plane's rows, 64 words each,
// and the last face of each plane's rows), their dimensions (the most slabs
// and chunks of any component) and the launch's tag
struct ExRun {
  unsigned *ticket, *done;
  unsigned long long *xk, *xi;
  int S, C;
  unsigned epoch;
};

// one tile diagonal d = I + J + K of the extrapolation wavefront: its tiles
// are those of component 0, then 1, then 2 (cnt of each), each in the order
// K, then J ascending (I = d - J - K)
struct ExDiag {
  int d;
  int cnt[3];
  int ti[3], tj[3], tk[3];  // tiles per axis of each component
};

// a planet's face fields for the batched extrapolation wavefront
struct ExPlanet {
  Geo g;
  float *vel;
  uint8_t *valid;
};

struct Net;       // the learned warm start (ws_learn.h)
struct Contacts;  // scratch of the frame-end contact passes (ws_contacts.hip)
struct Recorder;  // a recording of stage launches (ws_stage.h)

struct Ctx {
  // geometry
  int nx, ny, nz, capacity;
  int device = 0;  // the GPU current when the context was created
  double dx, W, H, D, cx, cy, cz, domainR;
  long nC, uN, vN, wN, nD;
  Geo geo;
  // tunables (solver.js:56-75) and world constants
  double cfl = 3.0, pic = 0.15, gravity = 9.81, viscosity = 5e-4, omega = 1.55, maxSpeed = 3.2;
  double vorticity = 0, coreR = 0, oceanR = 0, atmosphereH = 0.35, spacing = 0, particleVolume = 0;
  double gRamp = 1e9, dtLast = 0, iso = 1.8, substeps = 1;
  double rainT = 0.22, cloudT = 0.34, rainLiftAge = 0.4;
  int pressureIters = 20, maxPressureIters = 4000, nP = 0, tn = 0;
  double pressureTol = 0, dv = 0;
  int residualEvery = 8;
  int measureResidual = 0, ceilReflect = 0, hasSun = 0, adaptiveChecks = 1;
  double sunX = 0, sunY = 0, sunZ = 0, dtFrameBase = 0;
  bool hasVapor = false;
  // frame tick: thermal, evaporation, currents, phase (solver.js:137-193, 2670-2700)
  double heatK = 0.6, Tamb = 0.32, shadeCool = 0.035, sunActivity = 0, evapIntensity = 0.05;
  double snowT = 0.10, meltT = 0.105, cloudP = 0.35, currents = 0, Rlo = 0, Rhi = 0;
  double simTime = 0, heatAcc = 0;
  int spinOn = 0, evapCount = 0, haveSolid = 0;
  uint8_t *tsolid = nullptr, *pDepth = nullptr;
  float *pLight = nullptr, *pWx = nullptr, *pWy = nullptr, *pWz = nullptr, *pPh = nullptr;
  unsigned *evBuf = nullptr;    // 6 x capacity: cand, off, thresh, list, thrC, evList | evState
  uint32_t *jmp = nullptr;      // 2 x EV_WIN jump-ahead coefficients of the LCG (A, then C)
  float *currField = nullptr;   // 3 x lattice
  long currN = 0;
  // live timers (HIP events on the context stream), enabled by "timing"
  int timing = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // timed spans of the solver's own work (SOR batches, whole pressure solves,
  // warm-start inferences): event pairs left queued on the stream and read
  // only when a total is asked for (or the queue is long), so timing adds no
  // host synchronization to the run
  struct Span {
    hipEvent_t a, b;
    int kind;
    long launches;
  };
  std::vector<Span> spans;
  std::vector<hipEvent_t> evFree;
  double sorMs = 0, pressureMs = 0;
  long sorLaunches = 0, pressureSolves = 0, sweepsTotal = 0, residualsTotal = 0, ticksTotal = 0;
  long sorCellSweeps = 0;  // cells updated by the timed SOR launches
  // results of the last calls
  double umax = 0, residual = 0;
  int itersDone = 0, airborne = 0, nFluid = 0;
  bool gRampDirty = true;
  bool haveStatic = false, haveTerrain = false;
  std::string err;
  // learned pressure warm start: the network, whether the pressure solve
  // starts from its prediction, and whether solves record training samples
  Net *net = nullptr;
  int warmStart = 0, recordSamples = 0;
  int netTiled = 1;  // convolutions staged through LDS (0: the untiled kernel)
  int netCoarse = 0;  // the network runs on the grid coarsened 2x per axis
  // set by the batched driver: the forward pass and the correction run for
  // all planets at once after the planets' solve set-up
  int netDeferred = 0;
  double netLossCoarse = 0;  // weight of the 4x4x4-averaged error term of the training loss
  // balls (moving solids, solver.js:115, 918-938, 1005-1026, 2294-2564):
  // nine doubles per ball on the device (x y z vx vy vz r rho submerged),
  // the solid cell velocities (svx | svy | svz, 3 x nC, allocated with the
  // first ball), the waterline histogram (24 buckets + count per ball) and
  // the waterline levels of the last update
  int nBalls = 0, hadSolidVelocity = 0;
  double *balls = nullptr;
  float *sv = nullptr, *ballSurf = nullptr;
  unsigned *ballHist = nullptr;
  // frame-end contact passes: scratch, the vapor census, candidate overflows
  Contacts *contacts = nullptr;
  int vaporCount = 0, nightVaporCount = 0;
  int vaporSerialMax = 48;  // larger vapor groups run the dependency-ordered sweep
  // of those, groups up to vaporQueueMax steps (at most 16000) run from the
  // LDS ready queue, the rest by the polling sweep, with its counts in LDS up
  // to vaporLdsMax steps (at most 16384)
  int vaporQueueMax = 16000, vaporLdsMax = 16384;
  long vaporBigGroups = 0, vaporGroupMax = 0, vaporBigParcels = 0;  // of the last vapor pass
  long contactOverflow = 0;
  // both solves per pressure stage, timed separately (measurement)
  int pressureCompare = 0;
  double cmpMsPlain = 0, cmpMsLearned = 0;
  long cmpSweepsPlain = 0, cmpSweepsLearned = 0, cmpSolves = 0;
  double netMs = 0;  // timed network inferences inside pressure solves
  long netCalls = 0;

  hipStream_t stream = nullptr;
  // while set, stage launches are recorded here instead of launched
  Recorder *rec = nullptr;
  // static world
  uint8_t *staticTypes = nullptr;
  float *gu = nullptr, *gv = nullptr, *gw = nullptr;
  float *terrainR = nullptr;
  float *thr = nullptr;  // exact floor(x/dx) thresholds, max(n)+2 entries
  // cells
  uint8_t *cellType = nullptr, *nbm = nullptr, *dg = nullptr;
  float *q = nullptr, *q2 = nullptr, *rhs = nullptr, *pw = nullptr;
  uint8_t *sorActive = nullptr;  // per SOR tile: holds a fluid cell
  uint8_t *tbActive = nullptr;   // per temporal-blocking tile: holds a fluid cell
  unsigned *tbPk = nullptr;      // per tile and phase: listed cells (LDS index, mask, diagonal)
  float *tbRhs = nullptr;        // ... and their rhs
  int *tbCnt = nullptr;          // ... and their counts (4 per tile)
  int tbTiles[3] = {0, 0, 0};
  // fluid cells per colour, in cell order (red from 0, black from sorHalf):
  // code = cell | neighbour mask << 26, with rhs, omega/diag and diag per entry
  unsigned long long *sorOff = nullptr;
  unsigned *sorCode = nullptr;
  float *sorRhs = nullptr, *sorPw = nullptr;
  uint8_t *sorDg = nullptr;
  long sorHalf = 0, nRed = 0, nBlack = 0;
  long tbLaunchesPending = 0;
  int sorTiles[3] = {0, 0, 0};
  // 2: one launch per colour over the fluid-cell lists (the fastest
  // measured at 128^3), 3: two sweeps per launch over LDS tiles (temporal
  // blocking), 1: one LDS-tiled sweep per launch, 0: one launch per colour
  // over the whole grid. All give the same bits for the same sweeps.
  // 4: the whole solve in one workgroup (k_sor_solve, sorThreads threads),
  // residual checks on the device, no host round trips
  int sorMode = 2;
  int sorThreads = 1024;
  int p2gBin = 0;  // index into P2G_BINS
  double *sorChunk = nullptr;  // residual chunk sums of the on-device solve
  int sorGroups = 1;           // workgroups of one context's on-device solve
  unsigned *sorArrive = nullptr;  // their barrier counter and timed-out flag
  // 5: the whole solve on flowGroups workgroups that wait only for their
  // neighbours between half-sweeps (k_sor_flow), on the first flowXcds XCDs
  int flowGroups = 256, flowXcds = 8;
  unsigned *sorFlow = nullptr;  // progress counters, kFlowStride words apart
  // 6: n sweeps per launch on xcdGroups workgroups of one XCD
  // (k_sor_xcd; 0 = one per compute unit of an XCD). A launch that
  // declines leaves q untouched; the host sees it at its next read of the
  // device scalars and runs the sweeps with k_sor_list instead. A range too
  // large for its workgroup raises the cells per thread of later launches;
  // any other decline, or ranges too large at the most cells per thread,
  // keeps the context on sorMode 2 from then on (xcdFallbacks counts each
  // batch run by the list kernel instead)
  int xcdGroups = 0;
  int xcdUnits = 0;             // cells per thread and colour (1..3); 0: chosen from the list lengths
  unsigned xcdEpoch = 1;        // the next launch's counter and tag base
  unsigned long long *xcdX = nullptr;  // the tagged exchange array (nC words, allocated on first use)
  unsigned *xcdFlow = nullptr;  // arrival counters and placement reports (only k_sor_xcd writes them)
  unsigned xcdDeclinedSeen = 0, xcdTooBigSeen = 0;
  int xcdPending = 0;           // sweeps of the last launch not yet confirmed
  int xcdPendingUnits = 0;      // its cells per thread
  int xcdMinUnits = 1;          // raised when a launch's ranges did not fit
  long xcdFallbacks = 0;
  unsigned long long sorSweepsSeen = 0, sorChecksSeen = 0;  // device totals folded into the host counters
  bool devStatsPending = false;
  unsigned *cellCount = nullptr, *cellStart = nullptr;
  float *vort = nullptr;  // 10*C scratch for vorticity confinement
  // faces (u | v | w contiguous in one allocation each)
  float *vel = nullptr, *velO = nullptr, *wgt = nullptr, *scr = nullptr, *scr2 = nullptr;
  uint8_t *valid = nullptr;
  float *dens = nullptr;
  // particles
  float *px = nullptr, *py = nullptr, *pz = nullptr, *pvx = nullptr, *pvy = nullptr, *pvz = nullptr;
  float *pT = nullptr, *pAir = nullptr, *advU = nullptr, *advV = nullptr, *advW = nullptr;
  uint8_t *pflag = nullptr, *pcool = nullptr;
  int *pcell = nullptr;
  unsigned *prank = nullptr, *sortedIdx = nullptr;
  unsigned *vneed = nullptr;  // vapor draw flags and stream offsets, 4 x capacity
  float *vscr = nullptr;      // vapor phase scratch, 4 x capacity
  uint32_t *rng = nullptr;    // _cfS, _bfS, _evS stream states
  // reductions
  DevScalars *ds = nullptr;     // device
  DevScalars *hs = nullptr;     // pinned host mirror
  // pinned host word: the sequence number of the last scalars published
  // into hs by k_publish_scalars (hsPub: the last one requested)
  unsigned *hsSeq = nullptr;
  unsigned hsPub = 0;
  double *partials = nullptr;   // residual block partials
  int nPartials = 0;
  void *scanTmp = nullptr;
  size_t scanTmpBytes = 0;
  // the single-launch scans (ws_scan.hip): status words (two per tile of the
  // smallest tile size) and the launch epoch
  unsigned long long *scanStatus = nullptr;
  unsigned *scanCtl = nullptr;
  long scanTiles = 0;
  // extrapolation tile schedule
  int4 *exTiles = nullptr;
  // the row-pipelined extrapolation (exRows 1; 0, the default: one launch
  // per tile diagonal). Switched off: its first GPU run stalled with no
  // cause found; see the note in ws_extrap.hip. Final-flag words, exchange
  // arrays (cleared once, tags are launch epochs from 1), tickets and job
  // counts
  int exRows = 0;
  unsigned long long *exFin = nullptr, *exXk = nullptr, *exXi = nullptr;
  unsigned *exCnt = nullptr;
  int exS = 0, exC = 0;
  unsigned exEpoch = 0;
  std::vector<int4> exTilesHost;  // the same schedule on the host (batches merge the planets')
  std::vector<int> exDiagOff;  // tiles of diagonal d: [off[d], off[d+1])
  std::vector<ExDiag> exDiag;  // the same diagonals, for tiles found from the block index
  // marching tets scratch
  float *smpBuf = nullptr;  // sampleVel points and results, 6 floats per point
  long smpCap = 0;
  unsigned *mtCount = nullptr, *mtOff = nullptr;
  float *mtPos = nullptr, *mtNrm = nullptr, *mtDens = nullptr;
  long mtCells = 0, mtCap = 0, mtDensN = 0;
  // pinned host memory of the render readback
  uint8_t *renderHost = nullptr;
  size_t renderCap = 0;
  // metaball field (its own corner lattice of mbM cells a side) and its
  // fixed-point accumulator
  float *mbField = nullptr;
  unsigned long long *mbAcc = nullptr;
  long mbCap = 0;
  int mbM = 0;

  float *u() const { return vel; }
  float *v() const { return vel + uN; }
  float *w() const { return vel + uN + vN; }
};

#define WS_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      c->err = std::string(#call) + ": " + hipGetErrorString(e_);         \
      return WS_EHIP;                                                     \
    }                                                                     \
  } while (0)

#define WS_LAUNCHED()                                                     \
  do {                                                                    \
    hipError_t e_ = hipGetLastError();                                    \
    if (e_ != hipSuccess) {                                               \
      c->err = std::string("kernel launch: ") + hipGetErrorString(e_);    \
      return WS_EHIP;                                                     \
    }                                                                     \
  } while (0)

inline int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

// ------------------------------------------------------------------ device
// Exact floor(x / dx) of a float coordinate as the reference computes it in
// double: thr[i] is the smallest float x with floor(double(x)/dx) >= i.
__device__ __forceinline__ int cell_floor(float x, float inv_dx, const float *__restrict__ thr, int n) {
  int c = (int)floorf(x * inv_dx);
  c = c < 0 ? 0 : (c > n ? n : c);
  if (c < n && x >= thr[c + 1]) c++;
  else if (c > 0 && x < thr[c]) c--;
  return c;  // in [0, n]; callers clamp to their range
}

__device__ __forceinline__ int clampi(int v, int a, int b) { return v < a ? a : (v > b ? b : v); }

// Trilinear MAC stencil of one face component (solver.js:737-748).
struct Stencil {
  int idx, sj, sk;
  float s0, s1, t0, t1, r0, r1;
};

// Cell-local coordinate along one axis: index i0 = floor(x/dx - off) and
// fraction t = x/dx - off - i0. The fraction is formed as
// (x - (i0+off)*dx)/dx with an FMA and the low part of dx, so it keeps
// float precision relative to the cell instead of to the coordinate
// (x/dx - i0 in float would lose ~ulp(n) ~ 1e-5 at 128 cells).
__device__ __forceinline__ float axis_frac(const Geo &g, float x, float off, int i0) {
  float base = (float)i0 + off;  // exact: i0 < 2^23, off in {0, 0.5}
  float r = __builtin_fmaf(-base, g.dx, x) - base * g.dx_lo;
  return r * g.inv_dx;
}

__device__ __forceinline__ void face_stencil(const Geo &g, int comp, float x, float y, float z, Stencil &st) {
  float oa = comp == 0 ? 0.f : 0.5f, ob = comp == 1 ? 0.f : 0.5f, oc = comp == 2 ? 0.f : 0.5f;
  int i0 = (int)floorf(x * g.inv_dx - oa), j0 = (int)floorf(y * g.inv_dx - ob), k0 = (int)floorf(z * g.inv_dx - oc);
  float tx, ty, tz;
  if (i0 < 0) { i0 = 0; tx = 0.f; } else if (i0 > g.nx - 1) { i0 = g.nx - 1; tx = 1.f; } else tx = axis_frac(g, x, oa, i0);
  if (j0 < 0) { j0 = 0; ty = 0.f; } else if (j0 > g.ny - 2) { j0 = g.ny - 2; ty = 1.f; } else ty = axis_frac(g, y, ob, j0);
  if (k0 < 0) { k0 = 0; tz = 0.f; } else if (k0 > g.nz - 2) { k0 = g.nz - 2; tz = 1.f; } else tz = axis_frac(g, z, oc, k0);
  if (comp == 0) { st.sj = g.nx + 1; st.sk = g.ny * (g.nx + 1); st.idx = (k0 * g.ny + j0) * (g.nx + 1) + i0; }
  else if (comp == 1) { st.sj = g.nx; st.sk = (g.ny + 1) * g.nx; st.idx = (k0 * (g.ny + 1) + j0) * g.nx + i0; }
  else { st.sj = g.nx; st.sk = g.ny * g.nx; st.idx = (k0 * g.ny + j0) * g.nx + i0; }
  st.s0 = 1.f - tx; st.s1 = tx; st.t0 = 1.f - ty; st.t1 = ty; st.r0 = 1.f - tz; st.r1 = tz;
}

__device__ __forceinline__ float stencil_eval(const Stencil &st, const float *__restrict__ A) {
  int i = st.idx, sj = st.sj, sk = st.sk;
  float v00 = A[i] * st.s0 + A[i + 1] * st.s1;
  float v10 = A[i + sj] * st.s0 + A[i + sj + 1] * st.s1;
  float v01 = A[i + sk] * st.s0 + A[i + sk + 1] * st.s1;
  float v11 = A[i + sk + sj] * st.s0 + A[i + sk + sj + 1] * st.s1;
  return (v00 * st.t0 + v10 * st.t1) * st.r0 + (v01 * st.t0 + v11 * st.t1) * st.r1;
}

// Logical block index with consecutive chunks of the grid on one XCD: the
// hardware deals workgroups round-robin over the 8 XCDs, so block b runs on
// XCD b%8; remapping gives each XCD (and its L2) a contiguous slab.
// The sum of the fluid neighbours' q in the order -i, +i, -j, +j, -k, +k
// (mask bits 0..5), every load issued unconditionally so that all six are in
// flight together: an absent neighbour loads the cell itself and adds +0.
// The sum starts at +0 and so never holds -0, and adding +0 to it leaves its
// bits unchanged: the result is that of skipping the absent neighbours.
__device__ __forceinline__ float nb_sum(const float *__restrict__ q, int c, unsigned m, int nx, int nxy) {
  const float a = q[(m & 1u) ? c - 1 : c], b = q[(m & 2u) ? c + 1 : c];
  const float cc = q[(m & 4u) ? c - nx : c], d = q[(m & 8u) ? c + nx : c];
  const float e = q[(m & 16u) ? c - nxy : c], f = q[(m & 32u) ? c + nxy : c];
  float sum = 0.f;
  sum += (m & 1u) ? a : 0.f;
  sum += (m & 2u) ? b : 0.f;
  sum += (m & 4u) ? cc : 0.f;
  sum += (m & 8u) ? d : 0.f;
  sum += (m & 16u) ? e : 0.f;
  sum += (m & 32u) ? f : 0.f;
  return sum;
}

// The relaxation of solver.js:1402, q(1 - omega) + omega/diag (rhs + sum),
// with every rounding explicit so that all SOR kernels produce the same
// bits for the same inputs. The operators are written here under the
// pragma (HIP's __fmul_rn / __fadd_rn are plain operators defined elsewhere,
// which the compiler may contract into an FMA differently in each kernel).
__device__ __forceinline__ float sor_relax(float q, float sum, float rhs, float pw, float om1) {
#pragma clang fp contract(off)
  float keep = q * om1;
  float src = rhs + sum;
  float upd = pw * src;
  return keep + upd;
}

__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int X = 8;
  int per = nb / X, rem = nb % X, xcd = b % X, slot = b / X;
  // XCDs [0, rem) hold per+1 blocks, the others per
  return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

}  // namespace ws
