// Stacking and post-processing kernels (tsw_rtm/rtm_engine.py:142-158).
#pragma once
#include "common.h"

namespace tsw {

// 5-point Laplacian of one pixel from its four neighbours (symmetric boundary
// resolved by the caller) and centre, weights wz = FP32(1/dz^2),
// wx = FP32(1/dx^2), wc = FP32(-2/dx^2 - 2/dz^2) (rtm_engine.py:31-35).
// Terms are summed in FP32 in the order scipy.signal.convolve2d visits the
// 3x3 window (last kernel row first, last column first), which makes the
// result bit-identical to the reference's scipy call.
__host__ __device__ inline float lap_combine(float up, float dn, float lf, float rt, float c,
                                             float wz, float wx, float wc) {
  float s = dn * wz;
  s = s + rt * wx;
  s = s + c * wc;
  s = s + lf * wx;
  s = s + up * wz;
  return s;
}

void launch_stack_add(float* acc, const float* x, long long n, hipStream_t st);
void launch_stack_window(float* acc, int nx, const float* x, int wnx, int wnz, int x0, int z0,
                         hipStream_t st);
// Zeroes a (and b if non-null) where sponge < 0.99 (rtm_engine.py:147-149).
void launch_mute(float* a, float* b, const float* sponge, long long n, hipStream_t st);
// *count (zeroed by the caller) += number of NaN / +-inf values of x.
void launch_count_nonfinite(const float* x, long long n, unsigned* count, hipStream_t st);
// *out (zeroed by the caller) := bit pattern of max(x, 0).
void launch_max_nonneg(const float* x, long long n, unsigned* out, hipStream_t st);
void launch_normalize(const float* stacked, const float* illum, float* out, long long n,
                      const unsigned* maxbits, hipStream_t st);
// Laplacian of img into out; cells with mute_sponge < 0.99 are zeroed when
// mute_sponge is non-null (rtm_engine.py:156-158).
void launch_laplacian(const float* img, float* out, int nx, int nz, double dx, double dz,
                      const float* mute_sponge, hipStream_t st);

}  // namespace tsw
