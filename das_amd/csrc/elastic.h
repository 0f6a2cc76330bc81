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
  if (s.sponge != nullptr) return s.sponge[i];
  if (s.sponge_table == nullptr) return 1.0f;
  const int d = sponge_depth(ix, iz, s.nx, s.nz, s.n_sponge, s.fs);
  return d > 0 ? s.sponge_table[d] : 1.0f;
}

// Velocity pass at one cell, with the driver's injection and side work on the
// cell's vz before its update (wave2d.c:123-155).
__device__ __forceinline__ void ela_vel_cell(const ElaStep& s, int ix, int iz) {
  const int nx = s.nx;
  const int i = iz * nx + ix;
  float vzc = s.vz[i];
  bool injected = false;
  if (s.src_it >= 0 && i == s.src_cell) {
    vzc = vzc + (s.dt * s.bz[i]) * s.src[s.src_it];
    injected = true;
  }
  if (s.rdata_it >= 0 && iz == s.riz) {
    const int r0 = s.roff[ix], r1 = s.roff[ix + 1];
    const float* row = s.rdata + (long long)s.rdata_it * s.nrec;
    for (int r = r0; r < r1; ++r) {
      vzc = vzc + (s.dt * s.bz[i]) * row[s.ridx[r]];
      injected = true;
    }
  }
  if (s.snap != nullptr) {
    s.snap[i] = vzc;
    s.illum[i] = s.illum[i] + vzc * vzc;
  }
  if (s.image != nullptr) s.image[i] = s.image[i] + (s.isnap[i] * vzc) * s.w;
  const int z0 = s.fs ? 1 : 2;
  if (ix < 2 || ix >= nx - 2 || iz < z0 || iz >= s.nz - 2) {
    if (injected) s.vz[i] = vzc;
    return;
  }
  const float a = TSW_S1, b = TSW_S2;
  const float* sxx = s.sxx;
  const float* szz = s.szz;
  const float* sxz = s.sxz;
  const float dsxx = (a * (sxx[i + 1] - sxx[i]) + b * (sxx[i + 2] - sxx[i - 1])) * s.rdx;
  const float dsxz_z =
      (a * (sxz[i] - sxz[i - nx]) + b * (sxz[i + nx] - ela_row_or_zero(sxz, i - 2 * nx))) * s.rdz;
  float vx = s.vx[i] + (s.dt * s.bx[i]) * (dsxx + dsxz_z);
  const float dsxz_x = (a * (sxz[i] - sxz[i - 1]) + b * (sxz[i + 1] - sxz[i - 2])) * s.rdx;
  const float dszz = (a * (szz[i + nx] - szz[i]) + b * (szz[i + 2 * nx] - szz[i - nx])) * s.rdz;
  float vz = vzc + (s.dt * s.bz[i]) * (dsxz_x + dszz);
  const float g = ela_sponge(s, ix, iz, i);
  vx = vx * g;
  vz = vz * g;
  s.vx[i] = vx;
  s.vz[i] = vz;
}

// Vertical strain rate (4th-order staggered D-z of vz) at (iz, ix); rows
// outside the grid read 0.
__device__ __forceinline__ float ela_strain_rate(const ElaStep& s, int iz, int ix) {
  const int nx = s.nx, nz = s.nz;
  auto at = [&](int z) { return (z < 0 || z >= nz) ? 0.0f : s.vz[z * nx + ix]; };
  return (TSW_S1 * (at(iz) - at(iz - 1)) + TSW_S2 * (at(iz + 1) - at(iz - 2))) * s.rdz;
}

// Stress pass at one cell (wave2d.c:158-195), free surface on row 0
// (wave2d.c:198-211) and receiver recording.
__device__ __forceinline__ void ela_str_cell(const ElaStep& s, int ix, int iz) {
  const int nx = s.nx;
  const float a = TSW_S1, b = TSW_S2;
  const float* vx = s.vx;
  const float* vz = s.vz;
  if (s.rec_it >= 0 && iz == s.riz) {
    const int r0 = s.roff[ix], r1 = s.roff[ix + 1];
    if (r0 < r1) {
      const float e = ela_strain_rate(s, iz, ix);
      float* row = s.rec + (long long)s.rec_it * s.nrec;
      for (int r = r0; r < r1; ++r) row[s.ridx[r]] = e;
    }
  }
  if (iz == 0) {
    if (!s.fs) return;
    // assignment without dt, as the reference does
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
      (a * (vz[i] - vz[i - nx]) + b * (vz[i + nx] - ela_row_or_zero(vz, i - 2 * nx))) * s.rdz;
  const float l = s.lam[i], m = s.mu[i];
  const float l2m = l + 2.0f * m;
  float sxx = s.sxx[i] + s.dt * (l2m * dvx_x + l * dvz_z);
  float szz = s.szz[i] + s.dt * (l * dvx_x + l2m * dvz_z);
  const float dvz_x = (a * (vz[i + 1] - vz[i]) + b * (vz[i + 2] - vz[i - 1])) * s.rdx;
  const float dvx_z = (a * (vx[i + nx] - vx[i]) + b * (vx[i + 2 * nx] - vx[i - nx])) * s.rdz;
  float sxz = s.sxz[i] + (s.dt * m) * (dvz_x + dvx_z);
  const float g = ela_sponge(s, ix, iz, i);
  sxx = sxx * g;
  szz = szz * g;
  sxz = sxz * g;
  s.sxx[i] = sxx;
  s.szz[i] = szz;
  s.sxz[i] = sxz;
}

void launch_elastic_velocity(const ElaStep& s, hipStream_t st);
void launch_elastic_stress(const ElaStep& s, hipStream_t st);

}  // namespace tsw
