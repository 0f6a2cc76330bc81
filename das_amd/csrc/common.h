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
