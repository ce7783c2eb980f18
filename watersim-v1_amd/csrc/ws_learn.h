// The learned pressure warm start: network description, device buffers and
// the host entry points the pressure solve and the C ABI use.
#pragma once
#include <vector>

#include "ws_frame.h"

namespace ws {

typedef unsigned short bf16_t;

struct NetDims {
  int Z, Y, X;
};

// one 3x3x3 (or the 1x1x1 head) layer of the U-Net; the order and sizes are
// those of the flat parameter vector
struct NetLayer {
  int cin;   // real input channels
  int cinp;  // channels of the input tensor (a multiple of 8)
  int cout;
  int taps;  // 27, or 1 for the head
  long woff, boff;  // offsets of W[cout][taps][cin] and b[cout] in the flat vector
  int K, Kd;        // packed K of the forward product and of the input-gradient product
  bf16_t *wf = nullptr, *wd = nullptr;
};

constexpr int kNetLayers = 11;

struct Net {
  NetLayer L[kNetLayers];
  long nparam = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32 master, gradient, Adam moments
  long adamStep = 0;
  bool packed = false;
  // activations and gradients for grids of up to capV voxels
  long capV = 0;
  bf16_t *x0 = nullptr, *a0 = nullptr, *s1 = nullptr, *p1 = nullptr, *a2 = nullptr, *s2 = nullptr, *p2 = nullptr;
  bf16_t *a4 = nullptr, *a5 = nullptr, *u1 = nullptr, *a6 = nullptr, *a7 = nullptr, *u2 = nullptr, *a8 = nullptr;
  bf16_t *a9 = nullptr, *gz = nullptr;
  float *out = nullptr, *gout = nullptr, *tgt = nullptr;
  uint8_t *mask = nullptr;
  float *gA = nullptr, *gB = nullptr, *gC = nullptr, *gD = nullptr, *gE = nullptr, *gF = nullptr, *gG = nullptr;
  float *part = nullptr;
  long partN = 0;
  double *sums = nullptr;  // loss partials and totals
  // the solver side: initial residual, previous pressure, scales
  float *r0 = nullptr, *qprev = nullptr, *scales = nullptr;
  double *rpart = nullptr;
  // recorded (input, target) samples of the context's grid
  int bankCap = 0, bankCount = 0, bankNext = 0;
  int bankCoarse = 0;  // the samples are of the coarsened grid
  bf16_t *bankX = nullptr;
  float *bankT = nullptr;
  std::vector<void *> owned;
};

long net_param_count();
int net_ensure(ws_ctx *c);
int net_forward(ws_ctx *c, NetDims d);
int net_backward(ws_ctx *c, NetDims d, int accumulate);
int net_pack(ws_ctx *c);
void net_free(ws_ctx *c);
// the warm start inside the pressure solve: q += s * net(r0, types, q)
int net_warm_start(ws_ctx *c);
// many planets' forward passes and corrections at once, one launch per
// step: the planets' solves call net_warm_inputs in place of net_warm_start,
// then net_batch_forward runs on a stream ordered after all of them. A plan
// holds the planets' buffers (valid while their grids and networks stand);
// *bad names the planet of an error
struct NetBatch;
int net_warm_inputs(ws_ctx *c);
int net_batch_plan(ws_ctx *const *cs, int n, NetBatch **nb, int *bad);
void net_batch_forward(const NetBatch *nb, hipStream_t s);
void net_batch_free(NetBatch *nb);
// recording: inputs before the solve, the target after it
int net_record_inputs(ws_ctx *c);
int net_record_target(ws_ctx *c);

}  // namespace ws
