// Batched planets: many contexts advanced together by device-dt-law
// substeps, every stage as one launch over all planets on the batch stream:
// - the planets' own stages (binning, P2G, boundary conditions, viscosity,
//   the system's assembly, projection, confinement, G2P and advection): each
//   planet's substep is recorded once per run (its launches, with their
//   blocks and arguments, instead of launching them), the recordings are
//   merged in order and each merged step replays as one k_planets launch
//   that gives every planet its own blocks;
// - the learned warm start's forward pass and correction (one launch per
//   step of the network over every planet's grid);
// - the pressure solve (one launch of k_sor_solve_batch: every planet's
//   whole solve, a few workgroups each, the residual checks on the device);
// - the extrapolation wavefront (one launch per tile diagonal over every
//   planet's tiles of that diagonal).
// With stageBatch off, each planet's own stages are queued on its own stream
// instead (fork and join around the batched parts). A substep costs no host
// round trip. With a graph, the batched substep is captured once per run and
// replayed for every substep.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ws_frame.h"
#include "ws_learn.h"

using namespace ws;

struct ws_batch {
  std::vector<ws_ctx *> ctx;
  int useGraph = 0, threads = 1024;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, solved = nullptr;
  std::vector<hipEvent_t> pre, post;
  SorJob *jobs = nullptr;  // device copy of the planets' solve descriptions
  std::vector<SorJob> hjobs;
  // workgroups of the joint solve: G per planet (about one per
  // cellsPerGroup cells, at most maxGroups in all, so that every workgroup
  // is resident while the launch runs alone on the GPU), block -> (planet,
  // workgroup), and per planet a barrier counter and a timed-out flag
  double cellsPerGroup = 110000;
  int maxGroups = 256, nBlocks = 0;
  std::vector<int> G;
  int2 *blk = nullptr;
  unsigned *arrive = nullptr;
  // xcdLocal: lay each planet's workgroups out on one XCD where they fit
  // (blocks b with b % 8 equal), so that its barriers stay in that XCD's L2;
  // place holds the planets' placement words (G + 1 each, at placeOff)
  int xcdLocal = 1;
  unsigned *place = nullptr;
  // netBatch: the warm start's forward passes run as one launch per step
  // over all planets (else each planet runs its own on its stream; with
  // stageBatch on they are batched either way)
  int netBatch = 1;
  NetBatch *net = nullptr;
  // stageBatch: the planets' own stages as merged launches over all planets
  // (else on the planets' streams); a stage of the merged sequence is one
  // k_planets launch of the planets whose next recorded launch it is
  struct Stage {
    const StageOps *ops;
    int threads, nj;
    long blocks;
    size_t args, starts;  // offsets into stageMem
  };
  int stageBatch = 1;
  bool staged = false;
  std::vector<Stage> stages[3];  // before the solve, before and after the extrapolation
  unsigned char *stageMem = nullptr;
  std::vector<int> placeOff;
  // the planets' extrapolation schedules merged by tile diagonal: one launch
  // per diagonal covers every planet's tiles of that diagonal
  ExPlanet *exPlanets = nullptr;
  int4 *exTiles = nullptr;
  int *exPlanetOf = nullptr;
  std::vector<int> exDiagOff;
  std::vector<hipEvent_t> mid;
  hipEvent_t extrapolated = nullptr;
  // measurement: with timing on, substeps run eagerly with HIP events around
  // the joint solve launch and a synchronization after each
  int timing = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  double solveMs = 0, cellSweeps = 0, sweeps = 0;
  long solveLaunches = 0;
  std::vector<DevScalars> hs;  // the planets' device scalars after a timed substep
  std::string err;
};

namespace {

int fail(ws_batch *b, const std::string &m, int code) {
  b->err = m;
  return code;
}

#define BH(x)                                                                               \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail(b, std::string(#x) + ": " + hipGetErrorString(e_), WS_EHIP); \
  } while (0)

int planet_fail(ws_batch *b, int i, int r) {
  return fail(b, "planet " + std::to_string(i) + ": " + b->ctx[i]->err, r);
}

int check_planets(ws_batch *b) {
  for (size_t i = 0; i < b->ctx.size(); i++) {
    ws_ctx *c = b->ctx[i];
    if (c->timing || c->recordSamples || c->pressureCompare)
      return fail(b, "planet " + std::to_string(i) + ": timing, recordSamples and pressureCompare must be off",
                  WS_ESTATE);
    if (c->nC >= (1L << 26)) return fail(b, "planet " + std::to_string(i) + ": 2^26 cells or more", WS_EINVAL);
  }
  return WS_OK;
}

__global__ void k_batch_arm(int n, unsigned *__restrict__ arrive) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) arrive[2 * p] = 0u;
}

void launch_solves(ws_batch *b) {
  int n = (int)b->ctx.size();
  hipLaunchKernelGGL(k_batch_arm, dim3((n + 255) / 256), dim3(256), 0, b->stream, n, b->arrive);
  if (b->threads == 256)
    hipLaunchKernelGGL(k_sor_solve_batch<256>, dim3(b->nBlocks), dim3(256), 0, b->stream, b->jobs, b->blk);
  else if (b->threads == 512)
    hipLaunchKernelGGL(k_sor_solve_batch<512>, dim3(b->nBlocks), dim3(512), 0, b->stream, b->jobs, b->blk);
  else
    hipLaunchKernelGGL(k_sor_solve_batch<1024>, dim3(b->nBlocks), dim3(1024), 0, b->stream, b->jobs, b->blk);
}

// the blocks of the joint solve with each planet's workgroups on one XCD:
// workgroups are dealt to the 8 XCDs in turn (block b to XCD b % 8), so slot
// s of XCD x is block 8 s + x. Planets go to the XCD with the most free slots
// that holds them, the largest first; a planet that fits no XCD whole takes
// free slots wherever they are (its workgroups then find, on the device,
// that they do not share an XCD). Returns the table (x < 0: an empty slot)
// and marks which planets were laid out whole on one XCD.
std::vector<int2> xcd_layout(const std::vector<int> &G, int slots, std::vector<char> &whole) {
  const int n = (int)G.size();
  std::vector<int> order(n), fill(8, 0), xcd(n, -1);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return G[a] > G[b]; });
  whole.assign(n, 0);
  for (int i : order) {
    int best = -1;
    for (int x = 0; x < 8; x++)
      if (fill[x] + G[i] <= slots && (best < 0 || fill[x] < fill[best])) best = x;
    if (best >= 0) {
      xcd[i] = best;
      fill[best] += G[i];
      whole[i] = 1;
    }
  }
  int rows = 0;
  for (int x = 0; x < 8; x++) rows = std::max(rows, fill[x]);
  // the planets that fit nowhere whole, spread over the free slots
  int spill = 0;
  for (int i = 0; i < n; i++)
    if (!whole[i]) spill += G[i];
  while (8 * rows - std::accumulate(fill.begin(), fill.end(), 0) < spill) rows++;
  std::vector<int2> hb(8 * (size_t)rows, make_int2(-1, -1));
  std::vector<int> next(8, 0);
  for (int i : order)
    if (whole[i])
      for (int w = 0; w < G[i]; w++) hb[8 * (size_t)next[xcd[i]]++ + xcd[i]] = make_int2(i, w);
  size_t free_slot = 0;
  for (int i : order)
    if (!whole[i])
      for (int w = 0; w < G[i]; w++) {
        while (hb[free_slot].x >= 0) free_slot++;
        hb[free_slot] = make_int2(i, w);
      }
  return hb;
}

// workgroups per planet from its cell count, the block table, the jobs
int plan_solves(ws_batch *b) {
  int n = (int)b->ctx.size();
  if (n > b->maxGroups) return fail(b, "more planets than workgroups allowed for the joint solve", WS_EINVAL);
  b->G.assign(n, 1);
  int total = 0;
  for (int i = 0; i < n; i++) {
    double g = std::ceil((double)b->ctx[i]->nC / b->cellsPerGroup);
    b->G[i] = g < 1 ? 1 : (g > 32 ? 32 : (int)g);
    total += b->G[i];
  }
  while (total > b->maxGroups) {  // take from the planets with the most
    int m = 0;
    for (int i = 1; i < n; i++)
      if (b->G[i] > b->G[m]) m = i;
    b->G[m]--;
    total--;
  }
  std::vector<int2> hb;
  std::vector<char> whole(n, 0);
  if (b->xcdLocal && b->maxGroups >= 8) {
    hb = xcd_layout(b->G, b->maxGroups / 8, whole);
  } else {
    for (int i = 0; i < n; i++)
      for (int w = 0; w < b->G[i]; w++) hb.push_back(make_int2(i, w));
  }
  if (!b->blk) BH(hipMalloc((void **)&b->blk, sizeof(int2) * 256));
  if (hb.size() > 256) return fail(b, "the joint solve's layout exceeds 256 blocks", WS_EINVAL);
  b->nBlocks = (int)hb.size();
  BH(hipMemcpy(b->blk, hb.data(), sizeof(int2) * hb.size(), hipMemcpyHostToDevice));
  if (!b->place) BH(hipMalloc((void **)&b->place, sizeof(unsigned) * (256 + 256)));
  BH(hipMemset(b->place, 0, sizeof(unsigned) * (256 + 256)));
  b->placeOff.assign(n, -1);
  for (int i = 0, at = 0; i < n; i++) {
    b->hjobs[i] = batch_sor_job(b->ctx[i]);
    b->hjobs[i].G = b->G[i];
    b->hjobs[i].arrive = b->arrive + 2 * i;
    if (whole[i] && b->G[i] > 1) {
      b->placeOff[i] = at;
      b->hjobs[i].place = b->place + at;
      at += b->G[i] + 1;
    }
  }
  BH(hipMemcpy(b->jobs, b->hjobs.data(), sizeof(SorJob) * n, hipMemcpyHostToDevice));
  return WS_OK;
}

void launch_stages(ws_batch *b, int part) {
  for (const ws_batch::Stage &st : b->stages[part])
    st.ops->planets(b->stageMem + st.args, (const long *)(b->stageMem + st.starts), st.nj, st.blocks, st.threads,
                    b->stream);
}

void launch_extrapolation(ws_batch *b) {
  for (size_t d = 0; d + 1 < b->exDiagOff.size(); d++) {
    int nt = b->exDiagOff[d + 1] - b->exDiagOff[d];
    if (nt > 0)
      hipLaunchKernelGGL(k_extrap_diag_batch, dim3(nt), dim3(EX_THREADS), 0, b->stream, b->exPlanets, b->exTiles,
                         b->exPlanetOf, b->exDiagOff[d]);
  }
}

// one batched substep with every stage on the batch stream
int enqueue_staged(ws_batch *b) {
  launch_stages(b, 0);
  net_batch_forward(b->net, b->stream);
  BH(hipGetLastError());
  if (b->timing) BH(hipEventRecord(b->t0, b->stream));
  launch_solves(b);
  BH(hipGetLastError());
  if (b->timing) BH(hipEventRecord(b->t1, b->stream));
  launch_stages(b, 1);
  launch_extrapolation(b);
  launch_stages(b, 2);
  BH(hipGetLastError());
  return WS_OK;
}

// one batched substep
int enqueue_substep(ws_batch *b) {
  if (b->staged) return enqueue_staged(b);
  int n = (int)b->ctx.size(), r;
  BH(hipEventRecord(b->fork, b->stream));
  for (int i = 0; i < n; i++) {
    ws_ctx *c = b->ctx[i];
    BH(hipStreamWaitEvent(c->stream, b->fork, 0));
    if ((r = batch_substep_pre(c))) return planet_fail(b, i, r);
    BH(hipEventRecord(b->pre[i], c->stream));
    BH(hipStreamWaitEvent(b->stream, b->pre[i], 0));
  }
  net_batch_forward(b->net, b->stream);
  BH(hipGetLastError());
  if (b->timing) BH(hipEventRecord(b->t0, b->stream));
  launch_solves(b);
  BH(hipGetLastError());
  if (b->timing) BH(hipEventRecord(b->t1, b->stream));
  BH(hipEventRecord(b->solved, b->stream));
  for (int i = 0; i < n; i++) {
    ws_ctx *c = b->ctx[i];
    BH(hipStreamWaitEvent(c->stream, b->solved, 0));
    if ((r = batch_substep_post1(c))) return planet_fail(b, i, r);
    BH(hipEventRecord(b->mid[i], c->stream));
    BH(hipStreamWaitEvent(b->stream, b->mid[i], 0));
  }
  launch_extrapolation(b);
  BH(hipGetLastError());
  BH(hipEventRecord(b->extrapolated, b->stream));
  for (int i = 0; i < n; i++) {
    ws_ctx *c = b->ctx[i];
    BH(hipStreamWaitEvent(c->stream, b->extrapolated, 0));
    if ((r = batch_substep_post2(c))) return planet_fail(b, i, r);
    BH(hipEventRecord(b->post[i], c->stream));
    BH(hipStreamWaitEvent(b->stream, b->post[i], 0));
  }
  return WS_OK;
}

// the merged extrapolation schedule (the planets' grids and face buffers do
// not change over a batch's life)
int plan_extrapolation(ws_batch *b) {
  int n = (int)b->ctx.size();
  std::vector<ExPlanet> pl(n);
  size_t nd = 0;
  for (int i = 0; i < n; i++) {
    pl[i] = batch_ex_planet(b->ctx[i]);
    if (b->ctx[i]->exDiagOff.size() > nd) nd = b->ctx[i]->exDiagOff.size();
  }
  std::vector<int4> tiles;
  std::vector<int> owner;
  b->exDiagOff.assign(1, 0);
  for (size_t d = 0; d + 1 < nd; d++) {
    for (int i = 0; i < n; i++) {
      const ws_ctx *c = b->ctx[i];
      if (d + 1 >= c->exDiagOff.size()) continue;
      for (int t = c->exDiagOff[d]; t < c->exDiagOff[d + 1]; t++) {
        tiles.push_back(c->exTilesHost[t]);
        owner.push_back(i);
      }
    }
    b->exDiagOff.push_back((int)tiles.size());
  }
  BH(hipMalloc((void **)&b->exPlanets, sizeof(ExPlanet) * n));
  BH(hipMalloc((void **)&b->exTiles, sizeof(int4) * (tiles.size() + 1)));
  BH(hipMalloc((void **)&b->exPlanetOf, sizeof(int) * (owner.size() + 1)));
  BH(hipMemcpy(b->exPlanets, pl.data(), sizeof(ExPlanet) * n, hipMemcpyHostToDevice));
  if (!tiles.empty()) {
    BH(hipMemcpy(b->exTiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
    BH(hipMemcpy(b->exPlanetOf, owner.data(), sizeof(int) * owner.size(), hipMemcpyHostToDevice));
  }
  return WS_OK;
}

// after a timed substep: the solve's duration and the cells it relaxed
int account(ws_batch *b) {
  BH(hipStreamSynchronize(b->stream));
  float ms = 0;
  BH(hipEventElapsedTime(&ms, b->t0, b->t1));
  b->solveMs += ms;
  b->solveLaunches++;
  for (size_t i = 0; i < b->ctx.size(); i++) {
    BH(hipMemcpy(&b->hs[i], b->ctx[i]->ds, sizeof(DevScalars), hipMemcpyDeviceToHost));
    b->cellSweeps += (double)(b->hs[i].nRed + b->hs[i].nBlack) * b->hs[i].sorIters;
    b->sweeps += b->hs[i].sorIters;
  }
  return WS_OK;
}

int finish(ws_batch *b) {
  BH(hipStreamSynchronize(b->stream));
  for (ws_ctx *c : b->ctx) batch_solved(c);
  // a planet whose solve barrier timed out (its workgroups were not all
  // resident): the flag is cleared and the run reported as failed
  int n = (int)b->ctx.size();
  std::vector<unsigned> st(2 * n);
  BH(hipMemcpy(st.data(), b->arrive, sizeof(unsigned) * 2 * n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++)
    if (st[2 * i + 1]) {
      BH(hipMemset(b->arrive, 0, sizeof(unsigned) * 2 * n));
      return fail(b, "planet " + std::to_string(i) + ": the joint pressure solve's barrier timed out", WS_ESTATE);
    }
  return WS_OK;
}

// The planets' substeps recorded (pre: through the solve's set-up; post1:
// projection to clamp; post2: G2P to the substep's end) and merged into one
// sequence of stages per part. Each planet keeps its own order; a merged
// stage takes every planet whose next launch is of that kernel and block
// size, chosen among the kernels at the planets' heads: one that no other
// planet still has further on (else it would split a later stage), the one
// most planets are at. Buffers a recording creates are cleared on the
// planets' streams, which are drained before the batch stream runs.
int plan_stages(ws_batch *b) {
  for (auto &v : b->stages) v.clear();
  if (b->stageMem) (void)hipFree(b->stageMem);
  b->stageMem = nullptr;
  b->staged = false;
  if (!b->stageBatch) return WS_OK;
  const int n = (int)b->ctx.size();
  // planets with vapor parcels or balls take their own streams: those
  // passes are not recorded
  for (ws_ctx *c : b->ctx)
    if (c->hasVapor || c->nBalls || c->hadSolidVelocity) return WS_OK;
  std::vector<Recorder> rec((size_t)3 * n);
  for (int i = 0; i < n; i++) {
    ws_ctx *c = b->ctx[i];
    c->rec = &rec[3 * i];
    int r = batch_substep_pre(c);
    if (!r) {
      c->rec = &rec[3 * i + 1];
      r = batch_substep_post1(c);
    }
    if (!r) {
      c->rec = &rec[3 * i + 2];
      r = batch_substep_post2(c);
    }
    c->rec = nullptr;
    if (r) return planet_fail(b, i, r);
  }
  std::vector<unsigned char> mem;
  auto align = [&](size_t a) { mem.resize((mem.size() + a - 1) / a * a); };
  for (int part = 0; part < 3; part++) {
    std::vector<size_t> at(n, 0);
    auto seq = [&](int i) -> const std::vector<Recorded> & { return rec[3 * i + part].seq; };
    auto same = [](const Recorded &x, const Recorded &y) { return x.ops == y.ops && x.threads == y.threads; };
    for (;;) {
      int pick = -1, pickCount = 0;
      bool pickFree = false;
      for (int i = 0; i < n; i++) {
        if (at[i] >= seq(i).size()) continue;
        const Recorded &h = seq(i)[at[i]];
        int count = 0;
        bool free = true;
        for (int k = 0; k < n; k++) {
          if (at[k] >= seq(k).size()) continue;
          if (same(seq(k)[at[k]], h)) {
            count++;
            continue;
          }
          for (size_t q = at[k] + 1; q < seq(k).size() && free; q++)
            if (same(seq(k)[q], h)) free = false;
        }
        if (pick < 0 || (free && !pickFree) || (free == pickFree && count > pickCount)) {
          pick = i;
          pickCount = count;
          pickFree = free;
        }
      }
      if (pick < 0) break;
      const Recorded head = seq(pick)[at[pick]];
      std::vector<int> who;
      for (int i = 0; i < n; i++)
        if (at[i] < seq(i).size() && same(seq(i)[at[i]], head)) who.push_back(i);
      ws_batch::Stage st{head.ops, head.threads, (int)who.size(), 0, 0, 0};
      align(256);
      st.args = mem.size();
      for (int i : who) {
        const Recorded &x = seq(i)[at[i]];
        if (x.args.size() != head.ops->argSize)
          return fail(b, std::string("stage ") + head.ops->name + ": argument size mismatch", WS_EINVAL);
        mem.insert(mem.end(), x.args.begin(), x.args.end());
      }
      align(8);
      st.starts = mem.size();
      long total = 0;
      auto put = [&](long v) {
        const unsigned char *p = (const unsigned char *)&v;
        mem.insert(mem.end(), p, p + sizeof v);
      };
      for (int i : who) {
        put(total);
        total += seq(i)[at[i]].blocks;
        at[i]++;
      }
      put(total);
      st.blocks = total;
      if (total > 0) b->stages[part].push_back(st);
    }
  }
  BH(hipMalloc((void **)&b->stageMem, mem.size() ? mem.size() : 1));
  BH(hipMemcpy(b->stageMem, mem.data(), mem.size(), hipMemcpyHostToDevice));
  for (ws_ctx *c : b->ctx) BH(hipStreamSynchronize(c->stream));
  b->staged = true;
  return WS_OK;
}

// the substeps of a run: eagerly, timed, or captured once and replayed
int run_substeps(ws_batch *b, int substeps) {
  int r;
  // the solves' descriptions as the contexts stand now (tolerances, caps,
  // omega); every planet runs the one-workgroup solve
  for (size_t i = 0; i < b->ctx.size(); i++) {
    b->ctx[i]->sorMode = 4;
    b->ctx[i]->sorThreads = b->threads;
  }
  if ((r = plan_solves(b))) return r;
  int done = 0;
  if (b->timing) {
    for (; done < substeps; done++)
      if ((r = enqueue_substep(b)) || (r = account(b))) return r;
    return finish(b);
  }
  if (!b->useGraph) {
    for (; done < substeps; done++)
      if ((r = enqueue_substep(b))) return r;
    return finish(b);
  }
  // the first substep of every run goes eagerly: buffers the contexts create
  // on first use (network activations, after a change of settings) are
  // allocated outside the capture
  if (substeps > 0) {
    if ((r = enqueue_substep(b)) || (r = finish(b))) return r;
    done = 1;
  }
  if (done < substeps) {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    BH(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    r = enqueue_substep(b);
    hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (r || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (r) return r;
      return fail(b, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec), WS_EHIP);
    }
    hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(g);
      return fail(b, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei), WS_EHIP);
    }
    hipError_t el = hipSuccess;
    for (; done < substeps && el == hipSuccess; done++) el = hipGraphLaunch(x, b->stream);
    hipError_t es = hipStreamSynchronize(b->stream);
    (void)hipGraphExecDestroy(x);
    (void)hipGraphDestroy(g);
    if (el != hipSuccess) return fail(b, std::string("hipGraphLaunch: ") + hipGetErrorString(el), WS_EHIP);
    if (es != hipSuccess) return fail(b, std::string("hipStreamSynchronize: ") + hipGetErrorString(es), WS_EHIP);
  }
  return finish(b);
}
}  // namespace

extern "C" {

int ws_batch_create(ws_batch **out, ws_ctx *const *ctx, int n, int use_graph, int threads) {
  if (!out || !ctx || n < 1 || (threads != 256 && threads != 512 && threads != 1024)) return WS_EINVAL;
  *out = nullptr;
  for (int i = 0; i < n; i++)
    if (!ctx[i]) return WS_EINVAL;
  ws_batch *b = new ws_batch();
  b->ctx.assign(ctx, ctx + n);
  b->useGraph = use_graph != 0;
  b->threads = threads;
  b->pre.assign(n, nullptr);
  b->post.assign(n, nullptr);
  b->mid.assign(n, nullptr);
  b->hjobs.resize(n);
  b->hs.resize(n);
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->solved, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->t0);
  if (e == hipSuccess) e = hipEventCreate(&b->t1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->extrapolated, hipEventDisableTiming);
  for (int i = 0; i < n && e == hipSuccess; i++) {
    e = hipEventCreateWithFlags(&b->pre[i], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->post[i], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->mid[i], hipEventDisableTiming);
  }
  // one resident 1024-thread workgroup per compute unit
  int dev = 0, cus = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess && cus > 0 && cus < b->maxGroups) b->maxGroups = cus;
  if (e == hipSuccess) e = hipMalloc((void **)&b->jobs, sizeof(SorJob) * n);
  if (e == hipSuccess) e = hipMalloc((void **)&b->arrive, sizeof(unsigned) * 2 * n);
  if (e == hipSuccess) e = hipMemset(b->arrive, 0, sizeof(unsigned) * 2 * n);
  if (e != hipSuccess || plan_extrapolation(b) != WS_OK) {
    ws_batch_destroy(b);
    return WS_EHIP;
  }
  *out = b;
  return WS_OK;
}

int ws_batch_run(ws_batch *b, int substeps) {
  if (!b || substeps < 0) return WS_EINVAL;
  int r;
  if ((r = check_planets(b))) return r;
  // the warm starts' forward passes: batched, or each planet's own
  net_batch_free(b->net);
  b->net = nullptr;
  // (merged stages need the batched pass: a planet's own pass is not recorded)
  if (b->netBatch || b->stageBatch) {
    int bad = -1;
    if ((r = net_batch_plan(b->ctx.data(), (int)b->ctx.size(), &b->net, &bad)))
      return bad >= 0 ? planet_fail(b, bad, r) : fail(b, "warm-start network: batched pass", r);
  }
  for (ws_ctx *c : b->ctx) c->netDeferred = b->net && c->warmStart;
  // every planet runs the one-workgroup solve; the recordings are of that
  for (ws_ctx *c : b->ctx) {
    c->sorMode = 4;
    c->sorThreads = b->threads;
  }
  r = plan_stages(b);
  if (!r) r = run_substeps(b, substeps);
  for (ws_ctx *c : b->ctx) c->netDeferred = 0;
  return r;
}


int ws_batch_set(ws_batch *b, const char *name, double v) {
  if (!b || !name) return WS_EINVAL;
  if (!strcmp(name, "timing")) {
    b->timing = v != 0;
    b->solveMs = b->cellSweeps = b->sweeps = 0;
    b->solveLaunches = 0;
    return WS_OK;
  }
  if (!strcmp(name, "graph")) {
    b->useGraph = v != 0;
    return WS_OK;
  }
  if (!strcmp(name, "threads")) {
    if (v != 256 && v != 512 && v != 1024) return fail(b, "threads must be 256, 512 or 1024", WS_EINVAL);
    b->threads = (int)v;
    return WS_OK;
  }
  if (!strcmp(name, "xcdLocal")) {
    b->xcdLocal = v != 0;
    return WS_OK;
  }
  if (!strcmp(name, "netBatch")) {
    b->netBatch = v != 0;
    return WS_OK;
  }
  if (!strcmp(name, "stageBatch")) {
    b->stageBatch = v != 0;
    return WS_OK;
  }
  if (!strcmp(name, "cellsPerGroup")) {
    if (!(v >= 1)) return fail(b, "cellsPerGroup must be at least 1", WS_EINVAL);
    b->cellsPerGroup = v;
    return WS_OK;
  }
  if (!strcmp(name, "maxGroups")) {
    if (!(v >= 1 && v <= 256)) return fail(b, "maxGroups must be 1..256", WS_EINVAL);
    int dev = 0, cus = 0;
    BH(hipGetDevice(&dev));
    BH(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (v > cus) return fail(b, "maxGroups exceeds the compute units", WS_EINVAL);
    b->maxGroups = (int)v;
    return WS_OK;
  }
  return fail(b, std::string("unknown batch setting ") + name, WS_EINVAL);
}

int ws_batch_get(ws_batch *b, const char *name, double *v) {
  if (!b || !name || !v) return WS_EINVAL;
  if (!strcmp(name, "solveMs")) *v = b->solveMs;
  else if (!strcmp(name, "solveLaunches")) *v = (double)b->solveLaunches;
  else if (!strcmp(name, "cellSweeps")) *v = b->cellSweeps;
  else if (!strcmp(name, "sweeps")) *v = b->sweeps;
  else if (!strcmp(name, "threads")) *v = b->threads;
  else if (!strcmp(name, "graph")) *v = b->useGraph;
  else if (!strcmp(name, "timing")) *v = b->timing;
  else if (!strcmp(name, "cellsPerGroup")) *v = b->cellsPerGroup;
  else if (!strcmp(name, "maxGroups")) *v = b->maxGroups;
  else if (!strcmp(name, "groups")) *v = b->nBlocks;
  else if (!strcmp(name, "xcdLocal")) *v = b->xcdLocal;
  else if (!strcmp(name, "netBatch")) *v = b->netBatch;
  else if (!strcmp(name, "stageBatch")) *v = b->stageBatch;
  else if (!strcmp(name, "stageLaunches")) {
    // launches of the planets' own stages per batched substep
    *v = b->staged ? (double)(b->stages[0].size() + b->stages[1].size() + b->stages[2].size()) : 0.0;
  }
  else if (!strcmp(name, "xcdLocalPlanets")) {
    // planets whose last joint solve ran with its workgroups on one XCD
    int k = 0;
    BH(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < b->placeOff.size(); i++) {
      if (b->placeOff[i] < 0) continue;
      unsigned u = 0;
      BH(hipMemcpy(&u, b->place + b->placeOff[i] + b->G[i], sizeof(unsigned), hipMemcpyDeviceToHost));
      k += u != 0;
    }
    *v = k;
  }
  else return fail(b, std::string("unknown batch value ") + name, WS_EINVAL);
  return WS_OK;
}

const char *ws_batch_last_error(const ws_batch *b) { return b ? b->err.c_str() : "null batch"; }

void ws_batch_destroy(ws_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (hipEvent_t e : b->pre)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->post)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->mid)
    if (e) (void)hipEventDestroy(e);
  if (b->extrapolated) (void)hipEventDestroy(b->extrapolated);
  if (b->exPlanets) (void)hipFree(b->exPlanets);
  if (b->exTiles) (void)hipFree(b->exTiles);
  if (b->exPlanetOf) (void)hipFree(b->exPlanetOf);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->solved) (void)hipEventDestroy(b->solved);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->jobs) (void)hipFree(b->jobs);
  if (b->blk) (void)hipFree(b->blk);
  if (b->arrive) (void)hipFree(b->arrive);
  if (b->place) (void)hipFree(b->place);
  net_batch_free(b->net);
  if (b->stageMem) (void)hipFree(b->stageMem);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"
