// Model-side builders on the device: the lateral velocity flood of
// tsw_rtm/velocity_flood.py:24-94.  A flood model is a column gather of the
// base model (every column copies the base column the reference's slice and
// loop assignments leave there, computed on the host from those same index
// operations), optionally scaled by (1 + perturbation) in the base model's
// precision and stored as float32.
#include <cstring>
#include <vector>

#include "engine.h"
#include "tswrtm.h"

namespace tsw {

template <typename T>
__global__ void k_flood(const T* base, int nx, int nz, const int* colmap, int scale, T factor,
                        float* out) {
  const long long n = (long long)nx * nz;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(e / nx), ix = (int)(e - (long long)iz * nx);
    const int c = colmap[ix];
    T v = c < 0 ? T(0) : base[(long long)iz * nx + c];
    if (scale) v = v * factor;
    out[e] = (float)v;
  }
}

template <typename T>
static int flood(const T* base, int nx, int nz, const int* colmap, int scale, T factor,
                 float* out) {
  for (int i = 0; i < nx; ++i)
    TSW_ARG(colmap[i] >= -1 && colmap[i] < nx, "column map entry %d out of range", colmap[i]);
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const size_t n = (size_t)nx * nz;
  DevBuf b, m, o;
  TSW_TRY(b.ensure(n * sizeof(T)));
  TSW_TRY(m.ensure((size_t)nx * sizeof(int)));
  TSW_TRY(o.ensure(n * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(b.p, base, n * sizeof(T), hipMemcpyHostToDevice, st));
  TSW_HIP(hipMemcpyAsync(m.p, colmap, (size_t)nx * sizeof(int), hipMemcpyHostToDevice, st));
  const unsigned g = (unsigned)std::min<long long>(65536, ((long long)n + 255) / 256);
  hipLaunchKernelGGL(k_flood<T>, dim3(g), dim3(256), 0, st, (const T*)b.p, nx, nz, m.i(), scale,
                     factor, o.f());
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(out, o.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

}  // namespace tsw

using namespace tsw;

extern "C" {

int tsw_flood_columns(const void* base, int precision, int nx, int nz, const int* colmap,
                      int scale, double factor, float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(base && colmap && out && nx >= 1 && nz >= 1, "bad arguments");
  TSW_ARG(precision == 4 || precision == 8, "precision must be 4 or 8");
  if (precision == 4)
    return flood<float>((const float*)base, nx, nz, colmap, scale, (float)factor, out);
  return flood<double>((const double*)base, nx, nz, colmap, scale, factor, out);
}

}  // extern "C"
