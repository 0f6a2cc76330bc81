// A CPU stand-in for the parts of <hip/hip_runtime.h> that the engine's
// blocked kernels use, for running those kernels on the host in tests
// (tests/cpp/ts_emu.cpp; put this directory first on the include path and
// build with g++).  Every workgroup runs its threads as ucontext coroutines
// that switch only at __syncthreads, one workgroup after another, so a kernel
// that reads LDS written by another thread without a barrier in between, or
// whose threads reach different numbers of barriers, gives wrong results or
// aborts deterministically.  __shared__ variables become statics shared by the
// threads of the running workgroup; wave-level builtins that assume a uniform
// value (readfirstlane) return it, and the kernels' empty register-renaming
// asm statements compile to nothing.
#pragma once
#include <ucontext.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...) unused
#define asm
#define volatile(...)
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
typedef struct ihipStream_t* hipStream_t;
typedef int hipError_t;
enum { hipSuccess = 0 };
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipFree(void*) { return 0; }

namespace emu {
struct Th {
  ucontext_t ctx;
  dim3 tid;
  bool done = false, at_bar = false;
  int flag = 0;
};
inline std::vector<Th>& ths() { static std::vector<Th> v; return v; }
inline std::vector<std::vector<char>>& stacks() { static std::vector<std::vector<char>> v; return v; }
inline ucontext_t& sched() { static ucontext_t c; return c; }
inline int& cur() { static int c = 0; return c; }
inline int& bar_or() { static int v = 0; return v; }
inline long long& barriers() { static long long v = 0; return v; }
inline dim3& block_idx() { static dim3 b; return b; }
inline std::function<void()>& body() { static std::function<void()> f; return f; }
inline void tramp() {
  body()();
  ths()[cur()].done = true;
  swapcontext(&ths()[cur()].ctx, &sched());
}
inline void sync() {
  ths()[cur()].at_bar = true;
  swapcontext(&ths()[cur()].ctx, &sched());
}
inline int sync_or(int v) {
  ths()[cur()].flag = v != 0;
  sync();
  return bar_or();
}
inline void launch(dim3 grid, dim3 block, std::function<void()> f) {
  const int n = block.x * block.y * block.z;
  const size_t ss = 1 << 19;
  if ((int)stacks().size() < n) stacks().resize(n, std::vector<char>(ss));
  body() = std::move(f);
  for (unsigned b = 0; b < grid.x; ++b) {
    block_idx() = dim3(b);
    ths().assign(n, Th{});
    for (int i = 0; i < n; ++i) {
      Th& t = ths()[i];
      t.tid = dim3(i % block.x, (i / block.x) % block.y, i / (block.x * block.y));
      getcontext(&t.ctx);
      t.ctx.uc_stack.ss_sp = stacks()[i].data();
      t.ctx.uc_stack.ss_size = ss;
      t.ctx.uc_link = nullptr;
      makecontext(&t.ctx, tramp, 0);
    }
    for (;;) {
      int ndone = 0, nbar = 0, orv = 0;
      for (int i = 0; i < n; ++i) {
        if (ths()[i].done) continue;
        cur() = i;
        swapcontext(&sched(), &ths()[i].ctx);
      }
      for (int i = 0; i < n; ++i) {
        ndone += ths()[i].done;
        nbar += ths()[i].at_bar;
        orv |= ths()[i].flag;
      }
      if (ndone == n) break;
      if (ndone != 0 || nbar != n) {
        std::fprintf(stderr, "emu: barrier divergence in block %u (%d done, %d at barrier of %d)\n",
                     b, ndone, nbar, n);
        std::abort();
      }
      ++barriers();
      bar_or() = orv;
      for (int i = 0; i < n; ++i) ths()[i].at_bar = false, ths()[i].flag = 0;
    }
  }
}
}  // namespace emu

#define threadIdx (emu::ths()[emu::cur()].tid)
#define blockIdx (emu::block_idx())
inline void __syncthreads() { emu::sync(); }
inline int __syncthreads_or(int v) { return emu::sync_or(v); }
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) emu::launch(G, B, [&] { K(__VA_ARGS__); })
