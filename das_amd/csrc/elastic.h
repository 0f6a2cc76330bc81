// Elastic P-SV velocity-stress kernels: descriptor and launch helpers.
#pragma once
#include "common.h"

namespace tsw {

// One staggered-grid step (tsw_rtm/c_engine/wave2d.c:95-212) in two passes:
// velocities from stresses, then stresses from velocities plus the free
// surface on row 0.  Rows above the grid read as 0.
struct ElaStep {
  int nx, nz, fs;
  float rdx, rdz, dt;                    // 1/dx, 1/dz, dt in FP32
  float *vx, *vz, *sxx, *szz, *sxz;      // state, updated in place
  const float *lam, *mu, *bx, *bz;       // Lame parameters, buoyancies
  const float* sponge;                   // per-cell sponge; nullptr: see sponge_table
  const float* sponge_table;             // used when sponge is null: table[depth]
  int n_sponge;                          //   (nullptr: no absorbing strip)
  // Driver side work (all disabled when the *_it fields are < 0 and the
  // pointers null).  Velocity pass, before a cell's update: vertical-force
  // injection vz += (dt*bz)*sample of the source (src_cell, src[src_it]) or of
  // the receivers on row riz (rdata[rdata_it*nrec + r], receiver order), then
  // snapshot/illumination or imaging of that post-injection vz.  Stress pass:
  // vertical strain rate of the updated vz recorded at the receivers.
  int src_cell, src_it;
  const float* src;
  int riz, nrec;
  const int* roff;
  const int* ridx;
  const float* rdata;
  int rdata_it;
  float* rec;
  int rec_it;
  float* snap;
  float* illum;
  const float* isnap;
  float* image;
  float w;
};

__device__ __forceinline__ float ela_row_or_zero(const float* f, int i) {
  return i < 0 ? 0.0f : f[i];
}

__device__ __forceinline__ float ela_sponge(const ElaStep& s, int ix, int iz, int i) {
  if (s.sponge != nullptr) return as_global(s.sponge)[i];
  if (s.sponge_table == nullptr) return 1.0f;
  const int d = sponge_depth(ix, iz, s.nx, s.nz, s.n_sponge, s.fs);
  return d > 0 ? as_global(s.sponge_table)[d] : 1.0f;
}

// Velocity pass at one cell, with the driver's injection and side work on the
// cell's vz before its update (wave2d.c:123-155).  Every read of the update is
// issued first (the side work writes only the snapshot, illumination and
// image arrays), through global-address-space pointers.
__device__ __forceinline__ void ela_vel_cell(const ElaStep& s, int ix, int iz) {
  const int nx = s.nx;
  const int i = iz * nx + ix;
  const int z0 = s.fs ? 1 : 2;
  const bool upd = !(ix < 2 || ix >= nx - 2 || iz < z0 || iz >= s.nz - 2);
  const gfloat* sxx = as_global((const float*)s.sxx);
  const gfloat* szz = as_global((const float*)s.szz);
  const gfloat* sxz = as_global((const float*)s.sxz);
  const float vz0 = as_global((const float*)s.vz)[i];
  const float bzi = as_global(s.bz)[i];
  float vxi = 0.0f, bxi = 0.0f, g = 1.0f;
  float sxx_p1 = 0.0f, sxx_0 = 0.0f, sxx_p2 = 0.0f, sxx_m1 = 0.0f;
  float sxz_0 = 0.0f, sxz_mz = 0.0f, sxz_pz = 0.0f, sxz_m2z = 0.0f;
  float sxz_m1 = 0.0f, sxz_p1 = 0.0f, sxz_m2 = 0.0f;
  float szz_pz = 0.0f, szz_0 = 0.0f, szz_p2z = 0.0f, szz_mz = 0.0f;
  if (upd) {
    vxi = as_global((const float*)s.vx)[i];
    bxi = as_global(s.bx)[i];
    sxx_p1 = sxx[i + 1];
    sxx_0 = sxx[i];
    sxx_p2 = sxx[i + 2];
    sxx_m1 = sxx[i - 1];
    sxz_0 = sxz[i];
    sxz_mz = sxz[i - nx];
    sxz_pz = sxz[i + nx];
    sxz_m2z = i - 2 * nx < 0 ? 0.0f : sxz[i - 2 * nx];  // rows above the grid read 0
    sxz_m1 = sxz[i - 1];
    sxz_p1 = sxz[i + 1];
    sxz_m2 = sxz[i - 2];
    szz_pz = szz[i + nx];
    szz_0 = szz[i];
    szz_p2z = szz[i + 2 * nx];
    szz_mz = szz[i - nx];
    g = ela_sponge(s, ix, iz, i);
  }
  float vzc = vz0;
  bool injected = false;
  if (s.src_it >= 0 && i == s.src_cell) {
    vzc = vzc + (s.dt * bzi) * as_global(s.src)[s.src_it];
    injected = true;
  }
  if (s.rdata_it >= 0 && iz == s.riz) {
    const int r0 = as_global(s.roff)[ix], r1 = as_global(s.roff)[ix + 1];
    const gfloat* row = as_global(s.rdata) + (long long)s.rdata_it * s.nrec;
    for (int r = r0; r < r1; ++r) {
      vzc = vzc + (s.dt * bzi) * row[as_global(s.ridx)[r]];
      injected = true;
    }
  }
  if (s.snap != nullptr) {
    as_global(s.snap)[i] = vzc;
    gfloat* il = as_global(s.illum);
    il[i] = il[i] + vzc * vzc;
  }
  if (s.image != nullptr) {
    gfloat* im = as_global(s.image);
    im[i] = im[i] + (as_global(s.isnap)[i] * vzc) * s.w;
  }
  if (!upd) {
    if (injected) as_global(s.vz)[i] = vzc;
    return;
  }
  const float a = TSW_S1, b = TSW_S2;
  const float dsxx = (a * (sxx_p1 - sxx_0) + b * (sxx_p2 - sxx_m1)) * s.rdx;
  const float dsxz_z = (a * (sxz_0 - sxz_mz) + b * (sxz_pz - sxz_m2z)) * s.rdz;
  float vx = vxi + (s.dt * bxi) * (dsxx + dsxz_z);
  const float dsxz_x = (a * (sxz_0 - sxz_m1) + b * (sxz_p1 - sxz_m2)) * s.rdx;
  const float dszz = (a * (szz_pz - szz_0) + b * (szz_p2z - szz_mz)) * s.rdz;
  float vz = vzc + (s.dt * bzi) * (dsxz_x + dszz);
  vx = vx * g;
  vz = vz * g;
  as_global(s.vx)[i] = vx;
  as_global(s.vz)[i] = vz;
}

// Vertical strain rate (4th-order staggered D-z of vz) at (iz, ix); rows
// outside the grid read 0.
__device__ __forceinline__ float ela_strain_rate(const ElaStep& s, int iz, int ix) {
  const int nx = s.nx, nz = s.nz;
  auto at = [&](int z) { return (z < 0 || z >= nz) ? 0.0f : s.vz[z * nx + ix]; };
  return (TSW_S1 * (at(iz) - at(iz - 1)) + TSW_S2 * (at(iz + 1) - at(iz - 2))) * s.rdz;
}

// Stress pass at one cell (wave2d.c:158-195), free surface on row 0
// (wave2d.c:198-211) and receiver recording; reads through
// global-address-space pointers.
__device__ __forceinline__ void ela_str_cell(const ElaStep& s, int ix, int iz) {
  const int nx = s.nx;
  const float a = TSW_S1, b = TSW_S2;
  const gfloat* vx = as_global((const float*)s.vx);
  const gfloat* vz = as_global((const float*)s.vz);
  if (s.rec_it >= 0 && iz == s.riz) {
    const int r0 = as_global(s.roff)[ix], r1 = as_global(s.roff)[ix + 1];
    if (r0 < r1) {
      const float e = ela_strain_rate(s, iz, ix);
      gfloat* row = as_global(s.rec) + (long long)s.rec_it * s.nrec;
      for (int r = r0; r < r1; ++r) row[as_global(s.ridx)[r]] = e;
    }
  }
  if (iz == 0) {
    if (!s.fs) return;
    // assignment without dt, as the reference does
    as_global(s.szz)[ix] = 0.0f;
    as_global(s.sxz)[ix] = 0.0f;
    if (ix >= 2 && ix < nx - 2) {
      const float l = as_global(s.lam)[ix], m = as_global(s.mu)[ix];
      const float d = (a * (vx[ix] - vx[ix - 1]) + b * (vx[ix + 1] - vx[ix - 2])) * s.rdx;
      as_global(s.sxx)[ix] = (((4.0f * m) * (l + m)) / ((l + 2.0f * m) + 1e-8f)) * d;
    }
    return;
  }
  const int z0 = s.fs ? 1 : 2;
  if (ix < 2 || ix >= nx - 2 || iz < z0 || iz >= s.nz - 2) return;
  const int i = iz * nx + ix;
  gfloat* sxxp = as_global(s.sxx);
  gfloat* szzp = as_global(s.szz);
  gfloat* sxzp = as_global(s.sxz);
  const float dvx_x = (a * (vx[i] - vx[i - 1]) + b * (vx[i + 1] - vx[i - 2])) * s.rdx;
  // rows above the grid read 0
  const float vz_m2z = i - 2 * nx < 0 ? 0.0f : vz[i - 2 * nx];
  const float dvz_z = (a * (vz[i] - vz[i - nx]) + b * (vz[i + nx] - vz_m2z)) * s.rdz;
  const float l = as_global(s.lam)[i], m = as_global(s.mu)[i];
  const float l2m = l + 2.0f * m;
  float sxx = sxxp[i] + s.dt * (l2m * dvx_x + l * dvz_z);
  float szz = szzp[i] + s.dt * (l * dvx_x + l2m * dvz_z);
  const float dvz_x = (a * (vz[i + 1] - vz[i]) + b * (vz[i + 2] - vz[i - 1])) * s.rdx;
  const float dvx_z = (a * (vx[i + nx] - vx[i]) + b * (vx[i + 2 * nx] - vx[i - nx])) * s.rdz;
  float sxz = sxzp[i] + (s.dt * m) * (dvz_x + dvx_z);
  const float g = ela_sponge(s, ix, iz, i);
  sxx = sxx * g;
  szz = szz * g;
  sxz = sxz * g;
  sxxp[i] = sxx;
  szzp[i] = szz;
  sxzp[i] = sxz;
}

void launch_elastic_velocity(const ElaStep& s, hipStream_t st);
void launch_elastic_stress(const ElaStep& s, hipStream_t st);

}  // namespace tsw
