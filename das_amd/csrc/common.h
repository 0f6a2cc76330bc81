// Shared definitions for the tsw_rtm HIP engine (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace tsw {

// Status codes returned by every int-returning entry point.
enum Status : int {
  kOk = 0,
  kErrArg = -1,      // invalid argument (shape, index, null pointer)
  kErrHip = -2,      // HIP runtime error
  kErrMem = -3,      // device allocation failed or budget exceeded
  kErrState = -4,    // call not valid in the object's current state
  kErrInjected = -5,  // failure injected on purpose (tsw_debug_fail_next)
  kErrDiverged = -6   // non-finite image or illumination: the wavefield diverged
};

void set_error(const char* fmt, ...);
int fail_injected();  // consumes a pending injected failure (0 if none)

#define TSW_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::tsw::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                       __FILE__, __LINE__);                                        \
      return ::tsw::kErrHip;                                                       \
    }                                                                              \
  } while (0)

#define TSW_ARG(cond, ...)                \
  do {                                    \
    if (!(cond)) {                        \
      ::tsw::set_error(__VA_ARGS__);      \
      return ::tsw::kErrArg;              \
    }                                     \
  } while (0)

#define TSW_TRY(expr)               \
  do {                              \
    int rc_ = (expr);               \
    if (rc_ != 0) return rc_;       \
  } while (0)

// 4th-order central second-derivative weights and 4th-order staggered
// first-derivative weights, as float literals (wave2d.c:37-39, :117-118).
#define TSW_C0 (-2.5f)
#define TSW_C1 (4.0f / 3.0f)
#define TSW_C2 (-1.0f / 12.0f)
#define TSW_S1 (9.0f / 8.0f)
#define TSW_S2 (-1.0f / 24.0f)

// Global-address-space view of a device pointer.  Loads through a generic
// pointer (one read from a descriptor in memory) become flat loads, which
// also count against the LDS wait counter, so every LDS wait would wait for
// them too; global loads do not.
using gfloat = __attribute__((address_space(1))) float;
__device__ __forceinline__ const gfloat* as_global(const float* p) { return (const gfloat*)p; }
__device__ __forceinline__ gfloat* as_global(float* p) { return (gfloat*)p; }
using gint = __attribute__((address_space(1))) int;
__device__ __forceinline__ const gint* as_global(const int* p) { return (const gint*)p; }

// A pointer that is uniform over the wave, passed through readfirstlane so
// that the compiler keeps it in SGPRs as an opaque base: every access is then
// that scalar base plus one 32-bit vector byte offset (`at`), which global
// memory instructions take directly.  Without it the compiler folds the lane
// offset into 64-bit per-lane pointers and adds row offsets to those.
template <typename P>
__device__ __forceinline__ P* uniform_ptr(P* p) {
  const unsigned long long a = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  return (P*)((unsigned long long)hi << 32 | lo);
}
__device__ __forceinline__ const gfloat* at(const gfloat* base, unsigned boff) {
  using C = const __attribute__((address_space(1))) char;
  return (const gfloat*)((C*)base + boff);
}
__device__ __forceinline__ gfloat* at(gfloat* base, unsigned boff) {
  using C = __attribute__((address_space(1))) char;
  return (gfloat*)((C*)base + boff);
}

// Element of a row-major float field at byte offset `soff` + `voff` from
// `p`: `soff` uniform over the wave (a row's offset), `voff` the lane's (its
// column's).  32-bit offsets: fields stay under 4 GiB.
__device__ __forceinline__ float ld_at(const float* p, unsigned voff, unsigned soff) {
  using C = const __attribute__((address_space(1))) char;
  return *(const gfloat*)((C*)p + soff + voff);
}
__device__ __forceinline__ void st_at(float v, float* p, unsigned voff, unsigned soff) {
  using C = __attribute__((address_space(1))) char;
  *(gfloat*)((C*)p + soff + voff) = v;
}
// nontemporal forms for data streamed through HBM once
__device__ __forceinline__ float ld_at_nt(const float* p, unsigned voff, unsigned soff) {
  using C = const __attribute__((address_space(1))) char;
  return __builtin_nontemporal_load((const gfloat*)((C*)p + soff + voff));
}
__device__ __forceinline__ void st_at_nt(float v, float* p, unsigned voff, unsigned soff) {
  using C = __attribute__((address_space(1))) char;
  __builtin_nontemporal_store(v, (gfloat*)((C*)p + soff + voff));
}

// `x` as a value the compiler cannot relate to earlier ones.  A load in a
// conditional block then computes its 64-bit address in that block, where
// instruction selection sees the scalar base plus the zero-extended 32-bit
// offset and emits the SGPR-base form; with the extended offset inherited
// from a dominating block it falls back to 64-bit vector addresses.
__device__ __forceinline__ unsigned fresh(unsigned x) {
  asm volatile("" : "+v"(x));
  return x;
}

// `x` (uniform) as a value the compiler cannot relate to earlier ones: per
// step of an unrolled loop this keeps it from hoisting one copy of every
// step- and stage-dependent uniform expression out of the loop, which would
// need more scalar registers than there are.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Integer depth of cell (ix, iz) into the Cerjan strip (wave2d.c:247-256).
__host__ __device__ inline int sponge_depth(int ix, int iz, int nx, int nz, int n, int fs) {
  int d = 0;
  int t = n - ix;
  d = t > d ? t : d;
  t = ix - (nx - 1 - n);
  d = t > d ? t : d;
  t = iz - (nz - 1 - n);
  d = t > d ? t : d;
  if (!fs) {
    t = n - iz;
    d = t > d ? t : d;
  }
  return d;
}

}  // namespace tsw
