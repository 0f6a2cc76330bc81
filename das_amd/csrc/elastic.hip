// Elastic P-SV velocity-stress step for gfx950 (Levander staggered grid,
// tsw_rtm/c_engine/wave2d.c:95-212).  FP32 association follows the reference;
// built with -ffp-contract=off for bit-identical results.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "elastic.h"
#include "engine.h"
#include "tswrtm.h"

namespace tsw {

constexpr int kElaBX = 64;
constexpr int kElaBZ = 4;

__global__ __launch_bounds__(kElaBX* kElaBZ) void k_elastic_velocity(ElaStep s) {
  const int ix = blockIdx.x * kElaBX + threadIdx.x;
  const int iz = blockIdx.y * kElaBZ + threadIdx.y;
  if (ix >= s.nx || iz >= s.nz) return;
  ela_vel_cell(s, ix, iz);
}

__global__ __launch_bounds__(kElaBX* kElaBZ) void k_elastic_stress(ElaStep s) {
  const int ix = blockIdx.x * kElaBX + threadIdx.x;
  const int iz = blockIdx.y * kElaBZ + threadIdx.y;
  if (ix >= s.nx || iz >= s.nz) return;
  ela_str_cell(s, ix, iz);
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
  s.src_it = -1;
  s.rdata_it = -1;
  s.rec_it = -1;
  s.src_cell = -1;
  s.riz = -1;
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

// ------------------------------------------------------------ driver loops
namespace tsw {
namespace {

ElaStep ela_base(Grid* g, int slot) {
  ElaStep s;
  std::memset(&s, 0, sizeof s);
  s.nx = g->nx;
  s.nz = g->nz;
  s.fs = g->fs;
  s.rdx = 1.0f / g->dxf;
  s.rdz = 1.0f / g->dzf;
  s.dt = g->dtf;
  s.vx = g->ef[0].f();
  s.vz = g->ef[1].f();
  s.sxx = g->ef[2].f();
  s.szz = g->ef[3].f();
  s.sxz = g->ef[4].f();
  s.lam = g->epar[slot][0].f();
  s.mu = g->epar[slot][1].f();
  s.bx = g->epar[slot][2].f();
  s.bz = g->epar[slot][3].f();
  s.sponge = g->sponge.f();
  s.src_cell = -1;
  s.src_it = -1;
  s.rdata_it = -1;
  s.rec_it = -1;
  s.riz = -1;
  s.roff = g->roff.i();
  s.ridx = g->ridx.i();
  return s;
}

int ela_upload(Grid* g, int slot, const float* lam, const float* mu, const float* bx,
               const float* bz, hipStream_t st) {
  const size_t nb = (size_t)g->cells() * sizeof(float);
  const float* p[4] = {lam, mu, bx, bz};
  for (int j = 0; j < 4; ++j) {
    TSW_ARG(p[j], "null elastic model array");
    TSW_TRY(g->epar[slot][j].ensure(nb));
    TSW_HIP(hipMemcpyAsync(g->epar[slot][j].p, p[j], nb, hipMemcpyHostToDevice, st));
  }
  for (auto& f : g->ef) TSW_TRY(f.ensure(nb));
  return kOk;
}

int ela_recv(Grid* g, const int* rix, int nrec, hipStream_t st) {
  RecvMap m;
  TSW_TRY(m.build(rix, nrec, g->nx));
  TSW_TRY(g->roff.ensure(m.off.size() * sizeof(int)));
  TSW_TRY(g->ridx.ensure(m.idx.size() * sizeof(int)));
  TSW_HIP(hipMemcpyAsync(g->roff.p, m.off.data(), m.off.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  TSW_HIP(hipMemcpyAsync(g->ridx.p, m.idx.data(), m.idx.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

// Forward elastic loop: injection, side work, step, strain-rate recording.
int ela_forward(Grid* g, int slot, const float* d_src, int nt, int six, int siz, int nrec,
                int riz, int k, float* d_rec, float* d_snaps, float* d_illum, hipStream_t st) {
  const long long n = g->cells();
  for (auto& f : g->ef) launch_fill(f.f(), n, 0.0f, st);
  ElaStep s = ela_base(g, slot);
  s.src = d_src;
  s.src_cell = siz * g->nx + six;
  s.riz = riz;
  s.nrec = nrec;
  s.rec = d_rec;
  for (int it = 0; it < nt; ++it) {
    s.src_it = it;
    s.snap = nullptr;
    s.illum = nullptr;
    if (d_snaps && it % k == 0) {
      s.snap = d_snaps + (long long)(it / k) * n;
      s.illum = d_illum;
    }
    s.rec_it = -1;
    launch_elastic_velocity(s, st);
    s.rec_it = d_rec ? it : -1;
    launch_elastic_stress(s, st);
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}

int ela_backward(Grid* g, int slot, const float* d_data, int nt, int nrec, int riz,
                 const float* d_snaps, long long nsnap, float* d_image, int k, hipStream_t st) {
  const long long n = g->cells();
  for (auto& f : g->ef) launch_fill(f.f(), n, 0.0f, st);
  ElaStep s = ela_base(g, slot);
  s.riz = riz;
  s.nrec = nrec;
  s.rdata = d_data;
  s.w = (float)(g->dt * k);
  long long cursor = (nt - 1) / k;
  for (int it = nt - 1; it >= 0; --it) {
    s.rdata_it = it;
    s.image = nullptr;
    s.isnap = nullptr;
    if (it % k == 0 && cursor >= 0 && cursor < nsnap) {
      s.image = d_image;
      s.isnap = d_snaps + cursor * n;
      --cursor;
    }
    launch_elastic_velocity(s, st);
    launch_elastic_stress(s, st);
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}

int ela_shot_checks(Grid* g, int nt, int nrec, int riz, int k) {
  TSW_ARG(nt >= 1, "nt must be >= 1");
  TSW_ARG(nrec >= 0, "n_rec must be >= 0");
  TSW_ARG(riz >= 0 && riz < g->nz, "receiver_iz %d outside [0, %d)", riz, g->nz);
  TSW_ARG(k >= 1, "snapshot_interval must be >= 1");
  return kOk;
}

}  // namespace
}  // namespace tsw

extern "C" {

int tsw_forward_elastic(tsw_grid_t h, const float* lam, const float* mu, const float* bx,
                        const float* bz, const float* src, int nt, int six, int siz,
                        const int* rix, int nrec, int riz, int k, float* rec_out,
                        float* snaps_out, float* illum_out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_TRY(fail_injected());
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && src, "null argument");
  TSW_TRY(ela_shot_checks(g, nt, nrec, riz, k));
  TSW_ARG(six >= 0 && six < g->nx && siz >= 0 && siz < g->nz, "source outside grid");
  TSW_ARG(nrec == 0 || (rix && rec_out), "receivers need rix and rec_out");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = g->cells();
  TSW_TRY(ela_upload(g, 0, lam, mu, bx, bz, st));
  TSW_TRY(ela_recv(g, rix, nrec, st));
  TSW_TRY(g->src.ensure((size_t)nt * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(g->src.p, src, (size_t)nt * sizeof(float), hipMemcpyHostToDevice, st));
  TSW_TRY(g->rec.ensure((size_t)nt * (nrec > 0 ? nrec : 1) * sizeof(float)));
  const long long nsnap = (nt - 1) / k + 1;
  const bool side = snaps_out || illum_out;
  if (side) {
    TSW_TRY(g->snaps.ensure((size_t)nsnap * n * sizeof(float)));
    TSW_TRY(g->illum.ensure((size_t)n * sizeof(float)));
    launch_fill(g->illum.f(), n, 0.0f, st);
  }
  TSW_TRY(ela_forward(g, 0, g->src.f(), nt, six, siz, nrec, riz, k, nrec ? g->rec.f() : nullptr,
                      side ? g->snaps.f() : nullptr, side ? g->illum.f() : nullptr, st));
  if (nrec)
    TSW_HIP(hipMemcpyAsync(rec_out, g->rec.p, (size_t)nt * nrec * sizeof(float),
                           hipMemcpyDeviceToHost, st));
  if (snaps_out)
    TSW_HIP(hipMemcpyAsync(snaps_out, g->snaps.p, (size_t)nsnap * n * sizeof(float),
                           hipMemcpyDeviceToHost, st));
  if (illum_out)
    TSW_HIP(hipMemcpyAsync(illum_out, g->illum.p, (size_t)n * sizeof(float),
                           hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_backward_correlate_elastic(tsw_grid_t h, const float* lam, const float* mu,
                                   const float* bx, const float* bz, const float* data, int nt,
                                   const int* rix, int nrec, int riz, const float* snaps,
                                   long long nsnap, float* image_inout, int k) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_TRY(fail_injected());
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && image_inout, "null argument");
  TSW_TRY(ela_shot_checks(g, nt, nrec, riz, k));
  TSW_ARG(nrec == 0 || (rix && data), "receivers need rix and data");
  TSW_ARG(nsnap >= 0 && (nsnap == 0 || snaps), "snapshots missing");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = g->cells();
  TSW_TRY(ela_upload(g, 1, lam, mu, bx, bz, st));
  TSW_TRY(ela_recv(g, rix, nrec, st));
  TSW_TRY(g->data.ensure((size_t)nt * (nrec > 0 ? nrec : 1) * sizeof(float)));
  if (nrec)
    TSW_HIP(hipMemcpyAsync(g->data.p, data, (size_t)nt * nrec * sizeof(float),
                           hipMemcpyHostToDevice, st));
  long long used = std::min<long long>(nsnap, (nt - 1) / k + 1);
  if ((nt - 1) / k >= nsnap) used = 0;
  TSW_TRY(g->isnap.ensure((size_t)(used > 0 ? used : 1) * n * sizeof(float)));
  if (used)
    TSW_HIP(hipMemcpyAsync(g->isnap.p, snaps, (size_t)used * n * sizeof(float),
                           hipMemcpyHostToDevice, st));
  TSW_TRY(g->image.ensure((size_t)n * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(g->image.p, image_inout, (size_t)n * sizeof(float),
                         hipMemcpyHostToDevice, st));
  TSW_TRY(ela_backward(g, 1, g->data.f(), nt, nrec, riz, g->isnap.f(), used, g->image.f(), k, st));
  TSW_HIP(hipMemcpyAsync(image_inout, g->image.p, (size_t)n * sizeof(float),
                         hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_migrate_shot_elastic(tsw_grid_t h, const float* const* model_fwd,
                             const float* const* model_bwd, const float* src, int six, int siz,
                             const float* data, int nt, const int* rix, int nrec, int riz, int k,
                             float* image_out, float* illum_out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_TRY(fail_injected());
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && model_fwd && model_bwd && src && image_out && illum_out, "null argument");
  TSW_TRY(ela_shot_checks(g, nt, nrec, riz, k));
  TSW_ARG(six >= 0 && six < g->nx && siz >= 0 && siz < g->nz, "source outside grid");
  TSW_ARG(nrec == 0 || (rix && data), "receivers need rix and data");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = g->cells();
  TSW_TRY(ela_upload(g, 0, model_fwd[0], model_fwd[1], model_fwd[2], model_fwd[3], st));
  TSW_TRY(ela_upload(g, 1, model_bwd[0], model_bwd[1], model_bwd[2], model_bwd[3], st));
  TSW_TRY(ela_recv(g, rix, nrec, st));
  TSW_TRY(g->src.ensure((size_t)nt * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(g->src.p, src, (size_t)nt * sizeof(float), hipMemcpyHostToDevice, st));
  TSW_TRY(g->data.ensure((size_t)nt * (nrec > 0 ? nrec : 1) * sizeof(float)));
  if (nrec)
    TSW_HIP(hipMemcpyAsync(g->data.p, data, (size_t)nt * nrec * sizeof(float),
                           hipMemcpyHostToDevice, st));
  const long long nsnap = (nt - 1) / k + 1;
  TSW_TRY(g->snaps.ensure((size_t)nsnap * n * sizeof(float)));
  TSW_TRY(g->illum.ensure((size_t)n * sizeof(float)));
  TSW_TRY(g->image.ensure((size_t)n * sizeof(float)));
  launch_fill(g->illum.f(), n, 0.0f, st);
  launch_fill(g->image.f(), n, 0.0f, st);
  TSW_TRY(ela_forward(g, 0, g->src.f(), nt, six, siz, nrec, riz, k, nullptr, g->snaps.f(),
                      g->illum.f(), st));
  launch_scale(g->illum.f(), n, (float)(g->dt * k), st);
  TSW_TRY(ela_backward(g, 1, g->data.f(), nt, nrec, riz, g->snaps.f(), nsnap, g->image.f(), k,
                       st));
  TSW_HIP(hipMemcpyAsync(image_out, g->image.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                         st));
  TSW_HIP(hipMemcpyAsync(illum_out, g->illum.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                         st));
  TSW_HIP(hipStreamSynchronize(st));
  const float* arr[2] = {g->image.f(), g->illum.f()};
  const long long len[2] = {n, n};
  return check_finite(arr, len, 2, 2, 0, "elastic shot", st);
}

}  // extern "C"
