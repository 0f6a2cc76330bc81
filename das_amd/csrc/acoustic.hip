// Acoustic 4th-order leapfrog step for gfx950 with the time loop's per-step
// side work fused in (tsw_rtm/c_engine/wave2d.c:19-82 plus
// tsw_rtm/wave_solver.py:157-183 / :210-242).
//
// FP32 association follows wave2d.c expression by expression and the library
// is compiled with -ffp-contract=off, so a step is bit-identical to the
// reference compiled for x86-64.
#include "acoustic.h"

namespace tsw {

__global__ __launch_bounds__(kAcuBX* kAcuBZ) void k_acoustic_step(AcuStep s) {
  const int ix = blockIdx.x * kAcuBX + threadIdx.x;
  const int iz = blockIdx.y * kAcuBZ + threadIdx.y;
  if (ix >= s.nx || iz >= s.nz) return;
  acu_cell(s, ix, iz);
}

// Injection only (no step, no recording), into s.nxt.
__global__ void k_acoustic_inject(AcuStep s) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= s.nx) return;
  if (s.roff != nullptr && s.rdata_it >= 0 && s.riz >= 0 && s.riz < s.nz) {
    const int a = s.roff[ix], b = s.roff[ix + 1];
    if (a < b) {
      const int c = s.riz * s.nx + ix;
      float v = s.nxt[c];
      const float* row = s.rdata + (long long)s.rdata_it * s.nrec;
      for (int r = a; r < b; ++r) v = v + row[s.ridx[r]] * s.inj;
      s.nxt[c] = v;
    }
  }
  if (ix == 0 && s.src_it >= 0) s.nxt[s.src_cell] = s.nxt[s.src_cell] + s.src[s.src_it] * s.inj;
}

__global__ void k_fill(float* p, long long n, float v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_coef(const float* vel, float* coef, long long n, float dt2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = vel[i];
    coef[i] = (v * v) * dt2;
  }
}

__global__ void k_scale(float* p, long long n, float w) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = p[i] * w;
}

__global__ void k_correlate(const float* a, const float* b, float* img, long long n, float w) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    img[i] = img[i] + (a[i] * b[i]) * w;
}

// table[d] = expf(-alpha*d*d) for d = 0..n (table[0] unused: depth 0 -> 1).
__global__ void k_sponge(float* sp, int nx, int nz, int n, int fs, const float* table) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iz = blockIdx.y;
  if (ix >= nx) return;
  const int d = sponge_depth(ix, iz, nx, nz, n, fs);
  sp[iz * nx + ix] = d > 0 ? table[d] : 1.0f;
}

static inline unsigned grid1d(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

void launch_acoustic_step(const AcuStep& s, hipStream_t st) {
  dim3 blk(kAcuBX, kAcuBZ);
  dim3 grd((s.nx + kAcuBX - 1) / kAcuBX, (s.nz + kAcuBZ - 1) / kAcuBZ);
  hipLaunchKernelGGL(k_acoustic_step, grd, blk, 0, st, s);
}

void launch_acoustic_inject(const AcuStep& s, hipStream_t st) {
  hipLaunchKernelGGL(k_acoustic_inject, dim3((s.nx + 255) / 256), dim3(256), 0, st, s);
}

void launch_fill(float* p, long long n, float v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(grid1d(n)), dim3(256), 0, st, p, n, v);
}

void launch_coef(const float* vel, float* coef, long long n, float dt2, hipStream_t st) {
  hipLaunchKernelGGL(k_coef, dim3(grid1d(n)), dim3(256), 0, st, vel, coef, n, dt2);
}

void launch_scale(float* p, long long n, float w, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(grid1d(n)), dim3(256), 0, st, p, n, w);
}

void launch_correlate(const float* a, const float* b, float* img, long long n, float w,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_correlate, dim3(grid1d(n)), dim3(256), 0, st, a, b, img, n, w);
}

void launch_sponge(float* sp, int nx, int nz, int n, int fs, const float* table, hipStream_t st) {
  hipLaunchKernelGGL(k_sponge, dim3((nx + 255) / 256, nz), dim3(256), 0, st, sp, nx, nz, n, fs,
                     table);
}

}  // namespace tsw
