// Stage kernels written once for one planet or for many planets in one
// launch.
//
// A stage kernel's work lives in a body  void body(long bx, long nbx, P...)
// in which bx and nbx stand for blockIdx.x and gridDim.x. Its own kernel
// calls the body with the launch's block index and grid; k_planets runs the
// body for many planets at once: planet j owns the launch's blocks
// [starts[j], starts[j + 1]) and its block k there is block k of the launch
// it would have had alone, with its own arguments. The results are those of
// the planets' separate launches.
//
// stage_launch launches a stage for one planet, or, while the planet's
// context records (Ctx::rec), appends the launch (its blocks, block size
// and argument bytes) to the recording; a batch of planets whose recordings
// agree step by step replays them as one k_planets launch per step.
#pragma once
#include <cstring>
#include <vector>

#include "ws_internal.h"

namespace ws {

// the arguments of a kernel as one trivially copyable block
template <class... T>
struct Args;
template <>
struct Args<> {};
template <class H, class... T>
struct Args<H, T...> {
  H h;
  Args<T...> t;
};
template <class... T>
struct MakeArgs;
template <>
struct MakeArgs<> {
  static Args<> make() { return Args<>{}; }
};
template <class H, class... T>
struct MakeArgs<H, T...> {
  static Args<H, T...> make(H h, T... t) { return Args<H, T...>{h, MakeArgs<T...>::make(t...)}; }
};

template <class K>
struct ArgsOf;
template <class... P>
struct ArgsOf<void (*)(P...)> {
  using type = Args<P...>;
};

// the planet that owns block b: the last j with starts[j] <= b (starts
// ascending), as a wave-uniform value
__device__ __forceinline__ int planet_of_block(const long *__restrict__ starts, int nj, long b) {
  const int lane = threadIdx.x & 63;
  int k = 0;
  for (int base = 0; base < nj; base += 64) {
    const int j = base + lane;
    k += __popcll(__ballot(j < nj && starts[j] <= b));
  }
  return __builtin_amdgcn_readfirstlane(k - 1);
}

template <auto Body, class... Done>
__device__ __forceinline__ void call_body(long bx, long nbx, const Args<> &, Done... d) {
  Body(bx, nbx, d...);
}
template <auto Body, class H, class... T, class... Done>
__device__ __forceinline__ void call_body(long bx, long nbx, const Args<H, T...> &a, Done... d) {
  call_body<Body>(bx, nbx, a.t, d..., a.h);
}

template <auto Body, class A>
__global__ __launch_bounds__(256) void k_planets(const A *__restrict__ args, const long *__restrict__ starts, int nj) {
  const int j = planet_of_block(starts, nj, blockIdx.x);
  const long s0 = starts[j];
  call_body<Body>((long)blockIdx.x - s0, starts[j + 1] - s0, args[j]);
}

// a stage's launch over many planets: args holds nj argument blocks of
// argSize bytes, starts nj + 1 block offsets
struct StageOps {
  const char *name;
  size_t argSize;
  void (*planets)(const void *args, const long *starts, int nj, long blocks, int threads, hipStream_t s);
};

// defines the StageOps `ops` of kernel `kernel` (an expression naming it)
// with body `body`; ops is declared extern beside the kernel. (Not const: a
// const object would be placed in device memory as well, where the host
// launcher it points to does not exist.)
#define WS_PLANETS_OPS(ops, kernel, body)                                                                   \
  static void ops##_launch(const void *a, const long *s, int nj, long blocks, int threads, hipStream_t st) { \
    using A_ = typename ArgsOf<decltype(&kernel)>::type;                                                    \
    hipLaunchKernelGGL((k_planets<body, A_>), dim3(blocks), dim3(threads), 0, st, (const A_ *)a, s, nj);    \
  }                                                                                                         \
  StageOps ops = {#kernel, sizeof(typename ArgsOf<decltype(&kernel)>::type), ops##_launch};

struct Recorded {
  const StageOps *ops;
  long blocks;
  int threads;
  std::vector<unsigned char> args;
};
struct Recorder {
  std::vector<Recorded> seq;
};

// launch a stage kernel on the context's stream, or record the launch
template <class... P, class... A>
inline void stage_launch(Ctx *c, const StageOps &ops, void (*kernel)(P...), long blocks, int threads, A... a) {
  if (c->rec) {
    const Args<P...> t = MakeArgs<P...>::make((P)a...);
    Recorded r{&ops, blocks, threads, std::vector<unsigned char>(sizeof t)};
    std::memcpy(r.args.data(), &t, sizeof t);
    c->rec->seq.push_back(std::move(r));
    return;
  }
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, c->stream, (P)a...);
}

}  // namespace ws
