// Elastic P-SV velocity-stress step for gfx950 (Levander staggered grid,
// tsw_rtm/c_engine/wave2d.c:95-212).  FP32 association follows the reference;
// built with -ffp-contract=off for bit-identical results.
#include <cstring>
#include <mutex>
#include <vector>

#include "elastic.h"
#include "engine.h"

namespace tsw {

constexpr int kElaBX = 64;
constexpr int kElaBZ = 4;

__device__ __forceinline__ float row_or_zero(const float* f, int i) { return i < 0 ? 0.0f : f[i]; }

__global__ __launch_bounds__(kElaBX* kElaBZ) void k_elastic_velocity(ElaStep s) {
  const int ix = blockIdx.x * kElaBX + threadIdx.x;
  const int iz = blockIdx.y * kElaBZ + threadIdx.y;
  const int z0 = s.fs ? 1 : 2;
  if (ix < 2 || ix >= s.nx - 2 || iz < z0 || iz >= s.nz - 2) return;
  const int nx = s.nx;
  const int i = iz * nx + ix;
  const float a = TSW_S1, b = TSW_S2;
  const float* sxx = s.sxx;
  const float* szz = s.szz;
  const float* sxz = s.sxz;
  const float dsxx = (a * (sxx[i + 1] - sxx[i]) + b * (sxx[i + 2] - sxx[i - 1])) * s.rdx;
  const float dsxz_z =
      (a * (sxz[i] - sxz[i - nx]) + b * (sxz[i + nx] - row_or_zero(sxz, i - 2 * nx))) * s.rdz;
  float vx = s.vx[i] + (s.dt * s.bx[i]) * (dsxx + dsxz_z);
  const float dsxz_x = (a * (sxz[i] - sxz[i - 1]) + b * (sxz[i + 1] - sxz[i - 2])) * s.rdx;
  const float dszz = (a * (szz[i + nx] - szz[i]) + b * (szz[i + 2 * nx] - szz[i - nx])) * s.rdz;
  float vz = s.vz[i] + (s.dt * s.bz[i]) * (dsxz_x + dszz);
  if (s.sponge != nullptr) {
    const float g = s.sponge[i];
    vx = vx * g;
    vz = vz * g;
  }
  s.vx[i] = vx;
  s.vz[i] = vz;
}

__global__ __launch_bounds__(kElaBX* kElaBZ) void k_elastic_stress(ElaStep s) {
  const int ix = blockIdx.x * kElaBX + threadIdx.x;
  const int iz = blockIdx.y * kElaBZ + threadIdx.y;
  const int nx = s.nx;
  if (ix >= nx || iz >= s.nz) return;
  const float a = TSW_S1, b = TSW_S2;
  const float* vx = s.vx;
  const float* vz = s.vz;
  if (iz == 0) {
    if (!s.fs) return;
    // free surface (wave2d.c:198-211): assignment without dt, as the reference
    s.szz[ix] = 0.0f;
    s.sxz[ix] = 0.0f;
    if (ix >= 2 && ix < nx - 2) {
      const float l = s.lam[ix], m = s.mu[ix];
      const float d = (a * (vx[ix] - vx[ix - 1]) + b * (vx[ix + 1] - vx[ix - 2])) * s.rdx;
      s.sxx[ix] = (((4.0f * m) * (l + m)) / ((l + 2.0f * m) + 1e-8f)) * d;
    }
    return;
  }
  const int z0 = s.fs ? 1 : 2;
  if (ix < 2 || ix >= nx - 2 || iz < z0 || iz >= s.nz - 2) return;
  const int i = iz * nx + ix;
  const float dvx_x = (a * (vx[i] - vx[i - 1]) + b * (vx[i + 1] - vx[i - 2])) * s.rdx;
  const float dvz_z =
      (a * (vz[i] - vz[i - nx]) + b * (vz[i + nx] - row_or_zero(vz, i - 2 * nx))) * s.rdz;
  const float l = s.lam[i], m = s.mu[i];
  const float l2m = l + 2.0f * m;
  float sxx = s.sxx[i] + s.dt * (l2m * dvx_x + l * dvz_z);
  float szz = s.szz[i] + s.dt * (l * dvx_x + l2m * dvz_z);
  const float dvz_x = (a * (vz[i + 1] - vz[i]) + b * (vz[i + 2] - vz[i - 1])) * s.rdx;
  const float dvx_z = (a * (vx[i + nx] - vx[i]) + b * (vx[i + 2 * nx] - vx[i - nx])) * s.rdz;
  float sxz = s.sxz[i] + (s.dt * m) * (dvz_x + dvx_z);
  if (s.sponge != nullptr) {
    const float g = s.sponge[i];
    sxx = sxx * g;
    szz = szz * g;
    sxz = sxz * g;
  }
  s.sxx[i] = sxx;
  s.szz[i] = szz;
  s.sxz[i] = sxz;
}

void launch_elastic_velocity(const ElaStep& s, hipStream_t st) {
  dim3 blk(kElaBX, kElaBZ);
  dim3 grd((s.nx + kElaBX - 1) / kElaBX, (s.nz + kElaBZ - 1) / kElaBZ);
  hipLaunchKernelGGL(k_elastic_velocity, grd, blk, 0, st, s);
}

void launch_elastic_stress(const ElaStep& s, hipStream_t st) {
  dim3 blk(kElaBX, kElaBZ);
  dim3 grd((s.nx + kElaBX - 1) / kElaBX, (s.nz + kElaBZ - 1) / kElaBZ);
  hipLaunchKernelGGL(k_elastic_stress, grd, blk, 0, st, s);
}

std::recursive_mutex& api_mutex();

}  // namespace tsw

using namespace tsw;

namespace {
DevBuf g_ela[10];

int ref_elastic(float* vx, float* vz, float* sxx, float* szz, float* sxz, const float* lam,
                const float* mu, const float* bx, const float* bz, const float* sponge, int nx,
                int nz, float dx, float dz, float dt, int fs) {
  TSW_ARG(vx && vz && sxx && szz && sxz && lam && mu && bx && bz, "null argument");
  if (nx <= 0 || nz <= 0) return kOk;
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const size_t nb = (size_t)nx * nz * sizeof(float);
  float* state[5] = {vx, vz, sxx, szz, sxz};
  const float* par[5] = {lam, mu, bx, bz, sponge};
  for (int j = 0; j < 5; ++j) {
    TSW_TRY(g_ela[j].ensure(nb));
    TSW_HIP(hipMemcpyAsync(g_ela[j].p, state[j], nb, hipMemcpyHostToDevice, st));
  }
  for (int j = 0; j < 5; ++j) {
    if (!par[j]) continue;
    TSW_TRY(g_ela[5 + j].ensure(nb));
    TSW_HIP(hipMemcpyAsync(g_ela[5 + j].p, par[j], nb, hipMemcpyHostToDevice, st));
  }
  ElaStep s;
  std::memset(&s, 0, sizeof s);
  s.nx = nx;
  s.nz = nz;
  s.fs = fs ? 1 : 0;
  s.rdx = 1.0f / dx;
  s.rdz = 1.0f / dz;
  s.dt = dt;
  s.vx = g_ela[0].f();
  s.vz = g_ela[1].f();
  s.sxx = g_ela[2].f();
  s.szz = g_ela[3].f();
  s.sxz = g_ela[4].f();
  s.lam = g_ela[5].f();
  s.mu = g_ela[6].f();
  s.bx = g_ela[7].f();
  s.bz = g_ela[8].f();
  s.sponge = sponge ? g_ela[9].f() : nullptr;
  launch_elastic_velocity(s, st);
  launch_elastic_stress(s, st);
  TSW_HIP(hipGetLastError());
  for (int j = 0; j < 5; ++j)
    TSW_HIP(hipMemcpyAsync(state[j], g_ela[j].p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}
}  // namespace

extern "C" void elastic_2d_step(float* vx, float* vz, float* sxx, float* szz, float* sxz,
                                const float* lambda, const float* mu, const float* buo_x,
                                const float* buo_z, const float* sponge, int nx, int nz, float dx,
                                float dz, float dt, int free_surface_top) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  tsw::set_error("%s", "");
  (void)ref_elastic(vx, vz, sxx, szz, sxz, lambda, mu, buo_x, buo_z, sponge, nx, nz, dx, dz, dt,
                    free_surface_top);
}
