// Multi-shot stacking and image post-processing on the device
// (tsw_rtm/rtm_engine.py:142-158): FP32 stack, sponge mute, illumination
// normalisation and the 5-point Laplacian with symmetric boundary
// (scipy.signal.convolve2d(img, K, mode='same', boundary='symm')).
#include "post.h"

namespace tsw {

static inline unsigned blocks_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

__global__ void k_stack_add(float* acc, const float* x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc[i] = acc[i] + x[i];
}

// Adds a (wnz, wnx) window image into a larger (nz, nx) image at (z0, x0).
__global__ void k_stack_window(float* acc, int nx, const float* x, int wnx, int wnz, int x0,
                               int z0) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iz = blockIdx.y;
  if (ix >= wnx || iz >= wnz) return;
  float* a = acc + (long long)(z0 + iz) * nx + x0 + ix;
  *a = *a + x[(long long)iz * wnx + ix];
}

__global__ void k_mute(float* a, float* b, const float* sponge, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (sponge[i] < 0.99f) {
      a[i] = 0.0f;
      if (b) b[i] = 0.0f;
    }
  }
}

// Max of non-negative values via integer max on the IEEE bit pattern
// (order-independent, exact).  out must start at 0.
__global__ void k_max_nonneg(const float* x, long long n, unsigned* out) {
  unsigned m = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    const unsigned u = v > 0.0f ? __float_as_uint(v) : 0u;
    m = u > m ? u : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// *count += number of values that are NaN or +-inf (exponent bits all set).
__global__ void k_count_nonfinite(const float* x, long long n, unsigned* count) {
  unsigned c = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    c += (__float_as_uint(x[i]) & 0x7f800000u) == 0x7f800000u ? 1u : 0u;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

// normalized = stacked / (sqrt(illum) + eps), eps = 1e-4*max if max > 0 else 1e-12
// (numpy scalar rules: every operand rounded to FP32).
__global__ void k_normalize(const float* stacked, const float* illum, float* out, long long n,
                            const unsigned* maxbits) {
  const float mx = __uint_as_float(*maxbits);
  const float eps = mx > 0.0f ? 1e-4f * mx : 1e-12f;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = stacked[i] / (sqrtf(illum[i]) + eps);
}

// out = wz*up + wx*left ... with symmetric (edge-duplicating) boundary.
__global__ void k_laplacian(const float* img, float* out, int nx, int nz, float wz, float wx,
                            float wc, const float* mute_sponge) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iz = blockIdx.y;
  if (ix >= nx || iz >= nz) return;
  const int zu = iz > 0 ? iz - 1 : 0, zd = iz < nz - 1 ? iz + 1 : nz - 1;
  const int xl = ix > 0 ? ix - 1 : 0, xr = ix < nx - 1 ? ix + 1 : nx - 1;
  const long long c = (long long)iz * nx + ix;
  float v = lap_combine(img[(long long)zu * nx + ix], img[(long long)zd * nx + ix],
                        img[(long long)iz * nx + xl], img[(long long)iz * nx + xr], img[c], wz,
                        wx, wc);
  if (mute_sponge && mute_sponge[c] < 0.99f) v = 0.0f;
  out[c] = v;
}

void launch_stack_add(float* acc, const float* x, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_stack_add, dim3(blocks_for(n)), dim3(256), 0, st, acc, x, n);
}

void launch_stack_window(float* acc, int nx, const float* x, int wnx, int wnz, int x0, int z0,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_stack_window, dim3((wnx + 255) / 256, wnz), dim3(256), 0, st, acc, nx, x,
                     wnx, wnz, x0, z0);
}

void launch_mute(float* a, float* b, const float* sponge, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_mute, dim3(blocks_for(n)), dim3(256), 0, st, a, b, sponge, n);
}

void launch_count_nonfinite(const float* x, long long n, unsigned* count, hipStream_t st) {
  hipLaunchKernelGGL(k_count_nonfinite, dim3(blocks_for(n)), dim3(256), 0, st, x, n, count);
}

void launch_max_nonneg(const float* x, long long n, unsigned* out, hipStream_t st) {
  long long b = (n + 255) / 256;
  if (b > 1024) b = 1024;
  hipLaunchKernelGGL(k_max_nonneg, dim3((unsigned)(b < 1 ? 1 : b)), dim3(256), 0, st, x, n, out);
}

void launch_normalize(const float* stacked, const float* illum, float* out, long long n,
                      const unsigned* maxbits, hipStream_t st) {
  hipLaunchKernelGGL(k_normalize, dim3(blocks_for(n)), dim3(256), 0, st, stacked, illum, out, n,
                     maxbits);
}

void launch_laplacian(const float* img, float* out, int nx, int nz, double dx, double dz,
                      const float* mute_sponge, hipStream_t st) {
  const float wz = (float)(1.0 / (dz * dz));
  const float wx = (float)(1.0 / (dx * dx));
  const float wc = (float)(-2.0 / (dx * dx) - 2.0 / (dz * dz));
  hipLaunchKernelGGL(k_laplacian, dim3((nx + 255) / 256, nz), dim3(256), 0, st, img, out, nx, nz,
                     wz, wx, wc, mute_sponge);
}

}  // namespace tsw
