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
// cell's vz before its update (wave2d.c:123-155), in two halves: every read of
// the update (ela_vel_load), then the side work, the update and the writes
// (ela_vel_finish).  The side work writes only the snapshot, illumination and
// image arrays, so a thread can issue the reads of several cells before
// finishing any.  Reads go through global-address-space pointers.
struct ElaVelIn {
  bool upd;  // interior: the cell is updated
  float vz0, bzi, vxi, bxi, g;
  float sxx_p1, sxx_0, sxx_p2, sxx_m1;
  float sxz_0, sxz_mz, sxz_pz, sxz_m2z, sxz_m1, sxz_p1, sxz_m2;
  float szz_pz, szz_0, szz_p2z, szz_mz;
};

// Interior velocity update from the stencil values (wave2d.c:127-152); vzc is
// the cell's vz after injection.
__device__ __forceinline__ void ela_vel_update(const ElaVelIn& v, float vzc, float rdx, float rdz,
                                               float dt, float& vx, float& vz) {
  const float a = TSW_S1, b = TSW_S2;
  const float dsxx = (a * (v.sxx_p1 - v.sxx_0) + b * (v.sxx_p2 - v.sxx_m1)) * rdx;
  const float dsxz_z = (a * (v.sxz_0 - v.sxz_mz) + b * (v.sxz_pz - v.sxz_m2z)) * rdz;
  vx = v.vxi + (dt * v.bxi) * (dsxx + dsxz_z);
  const float dsxz_x = (a * (v.sxz_0 - v.sxz_m1) + b * (v.sxz_p1 - v.sxz_m2)) * rdx;
  const float dszz = (a * (v.szz_pz - v.szz_0) + b * (v.szz_p2z - v.szz_mz)) * rdz;
  vz = vzc + (dt * v.bzi) * (dsxz_x + dszz);
  vx = vx * v.g;
  vz = vz * v.g;
}

__device__ __forceinline__ ElaVelIn ela_vel_load(const ElaStep& s, int ix, int iz) {
  const int nx = s.nx;
  const int i = iz * nx + ix;
  const int z0 = s.fs ? 1 : 2;
  ElaVelIn v;
  v.upd = !(ix < 2 || ix >= nx - 2 || iz < z0 || iz >= s.nz - 2);
  const gfloat* sxx = as_global((const float*)s.sxx);
  const gfloat* szz = as_global((const float*)s.szz);
  const gfloat* sxz = as_global((const float*)s.sxz);
  v.vz0 = as_global((const float*)s.vz)[i];
  v.bzi = as_global(s.bz)[i];
  v.vxi = v.bxi = 0.0f;
  v.g = 1.0f;
  v.sxx_p1 = v.sxx_0 = v.sxx_p2 = v.sxx_m1 = 0.0f;
  v.sxz_0 = v.sxz_mz = v.sxz_pz = v.sxz_m2z = v.sxz_m1 = v.sxz_p1 = v.sxz_m2 = 0.0f;
  v.szz_pz = v.szz_0 = v.szz_p2z = v.szz_mz = 0.0f;
  if (v.upd) {
    v.vxi = as_global((const float*)s.vx)[i];
    v.bxi = as_global(s.bx)[i];
    v.sxx_p1 = sxx[i + 1];
    v.sxx_0 = sxx[i];
    v.sxx_p2 = sxx[i + 2];
    v.sxx_m1 = sxx[i - 1];
    v.sxz_0 = sxz[i];
    v.sxz_mz = sxz[i - nx];
    v.sxz_pz = sxz[i + nx];
    v.sxz_m2z = i - 2 * nx < 0 ? 0.0f : sxz[i - 2 * nx];  // rows above the grid read 0
    v.sxz_m1 = sxz[i - 1];
    v.sxz_p1 = sxz[i + 1];
    v.sxz_m2 = sxz[i - 2];
    v.szz_pz = szz[i + nx];
    v.szz_0 = szz[i];
    v.szz_p2z = szz[i + 2 * nx];
    v.szz_mz = szz[i - nx];
    v.g = ela_sponge(s, ix, iz, i);
  }
  return v;
}

__device__ __forceinline__ void ela_vel_finish(const ElaStep& s, int ix, int iz,
                                               const ElaVelIn& v) {
  const int nx = s.nx;
  const int i = iz * nx + ix;
  float vzc = v.vz0;
  bool injected = false;
  if (s.src_it >= 0 && i == s.src_cell) {
    vzc = vzc + (s.dt * v.bzi) * as_global(s.src)[s.src_it];
    injected = true;
  }
  if (s.rdata_it >= 0 && iz == s.riz) {
    const int r0 = as_global(s.roff)[ix], r1 = as_global(s.roff)[ix + 1];
    const gfloat* row = as_global(s.rdata) + (long long)s.rdata_it * s.nrec;
    for (int r = r0; r < r1; ++r) {
      vzc = vzc + (s.dt * v.bzi) * row[as_global(s.ridx)[r]];
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
  if (!v.upd) {
    if (injected) as_global(s.vz)[i] = vzc;
    return;
  }
  float vx, vz;
  ela_vel_update(v, vzc, s.rdx, s.rdz, s.dt, vx, vz);
  as_global(s.vx)[i] = vx;
  as_global(s.vz)[i] = vz;
}

__device__ __forceinline__ void ela_vel_cell(const ElaStep& s, int ix, int iz) {
  ela_vel_finish(s, ix, iz, ela_vel_load(s, ix, iz));
}

// Vertical strain rate (4th-order staggered D-z of vz) at (iz, ix); rows
// outside the grid read 0.
__device__ __forceinline__ float ela_strain_rate(const ElaStep& s, int iz, int ix) {
  const int nx = s.nx, nz = s.nz;
  auto at = [&](int z) { return (z < 0 || z >= nz) ? 0.0f : s.vz[z * nx + ix]; };
  return (TSW_S1 * (at(iz) - at(iz - 1)) + TSW_S2 * (at(iz + 1) - at(iz - 2))) * s.rdz;
}

// Stress pass at one cell (wave2d.c:158-195), free surface on row 0
// (wave2d.c:198-211) and receiver recording, in the same two halves: the
// interior update's reads (ela_str_load; the pass writes only stresses and
// receiver rows, which it does not read), then the rest (ela_str_finish).
struct ElaStrIn {
  bool upd;  // interior: the cell is updated
  float vx_0, vx_m1, vx_p1, vx_m2, vx_pz, vx_p2z, vx_mz;
  float vz_0, vz_mz, vz_pz, vz_m2z, vz_p1, vz_p2, vz_m1;
  float l, m, sxx, szz, sxz, g;
};

// Interior stress update from the stencil values (wave2d.c:162-192).
__device__ __forceinline__ void ela_str_update(const ElaStrIn& v, float rdx, float rdz, float dt,
                                               float& sxx, float& szz, float& sxz) {
  const float a = TSW_S1, b = TSW_S2;
  const float dvx_x = (a * (v.vx_0 - v.vx_m1) + b * (v.vx_p1 - v.vx_m2)) * rdx;
  const float dvz_z = (a * (v.vz_0 - v.vz_mz) + b * (v.vz_pz - v.vz_m2z)) * rdz;
  const float l2m = v.l + 2.0f * v.m;
  sxx = v.sxx + dt * (l2m * dvx_x + v.l * dvz_z);
  szz = v.szz + dt * (v.l * dvx_x + l2m * dvz_z);
  const float dvz_x = (a * (v.vz_p1 - v.vz_0) + b * (v.vz_p2 - v.vz_m1)) * rdx;
  const float dvx_z = (a * (v.vx_pz - v.vx_0) + b * (v.vx_p2z - v.vx_mz)) * rdz;
  sxz = v.sxz + (dt * v.m) * (dvz_x + dvx_z);
  sxx = sxx * v.g;
  szz = szz * v.g;
  sxz = sxz * v.g;
}

// sxx on the free-surface row from the updated vx of row 0 at ix-2..ix+1
// (wave2d.c:198-211): an assignment without dt, as the reference does.
__device__ __forceinline__ float ela_free_surface_sxx(float l, float m, float vx_m2, float vx_m1,
                                                      float vx_0, float vx_p1, float rdx) {
  const float a = TSW_S1, b = TSW_S2;
  const float d = (a * (vx_0 - vx_m1) + b * (vx_p1 - vx_m2)) * rdx;
  return (((4.0f * m) * (l + m)) / ((l + 2.0f * m) + 1e-8f)) * d;
}

__device__ __forceinline__ ElaStrIn ela_str_load(const ElaStep& s, int ix, int iz) {
  const int nx = s.nx;
  const int z0 = s.fs ? 1 : 2;
  ElaStrIn v;
  v.upd = !(iz == 0 || ix < 2 || ix >= nx - 2 || iz < z0 || iz >= s.nz - 2);
  v.vx_0 = v.vx_m1 = v.vx_p1 = v.vx_m2 = v.vx_pz = v.vx_p2z = v.vx_mz = 0.0f;
  v.vz_0 = v.vz_mz = v.vz_pz = v.vz_m2z = v.vz_p1 = v.vz_p2 = v.vz_m1 = 0.0f;
  v.l = v.m = v.sxx = v.szz = v.sxz = 0.0f;
  v.g = 1.0f;
  if (v.upd) {
    const int i = iz * nx + ix;
    const gfloat* vx = as_global((const float*)s.vx);
    const gfloat* vz = as_global((const float*)s.vz);
    v.vx_0 = vx[i];
    v.vx_m1 = vx[i - 1];
    v.vx_p1 = vx[i + 1];
    v.vx_m2 = vx[i - 2];
    v.vx_pz = vx[i + nx];
    v.vx_p2z = vx[i + 2 * nx];
    v.vx_mz = vx[i - nx];
    v.vz_0 = vz[i];
    v.vz_mz = vz[i - nx];
    v.vz_pz = vz[i + nx];
    v.vz_m2z = i - 2 * nx < 0 ? 0.0f : vz[i - 2 * nx];  // rows above the grid read 0
    v.vz_p1 = vz[i + 1];
    v.vz_p2 = vz[i + 2];
    v.vz_m1 = vz[i - 1];
    v.l = as_global(s.lam)[i];
    v.m = as_global(s.mu)[i];
    v.sxx = as_global((const float*)s.sxx)[i];
    v.szz = as_global((const float*)s.szz)[i];
    v.sxz = as_global((const float*)s.sxz)[i];
    v.g = ela_sponge(s, ix, iz, i);
  }
  return v;
}

__device__ __forceinline__ void ela_str_finish(const ElaStep& s, int ix, int iz,
                                               const ElaStrIn& v) {
  const int nx = s.nx;
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
    const gfloat* vx = as_global((const float*)s.vx);
    as_global(s.szz)[ix] = 0.0f;
    as_global(s.sxz)[ix] = 0.0f;
    if (ix >= 2 && ix < nx - 2)
      as_global(s.sxx)[ix] = ela_free_surface_sxx(as_global(s.lam)[ix], as_global(s.mu)[ix],
                                                  vx[ix - 2], vx[ix - 1], vx[ix], vx[ix + 1],
                                                  s.rdx);
    return;
  }
  if (!v.upd) return;
  const int i = iz * nx + ix;
  float sxx, szz, sxz;
  ela_str_update(v, s.rdx, s.rdz, s.dt, sxx, szz, sxz);
  as_global(s.sxx)[i] = sxx;
  as_global(s.szz)[i] = szz;
  as_global(s.sxz)[i] = sxz;
}

__device__ __forceinline__ void ela_str_cell(const ElaStep& s, int ix, int iz) {
  ela_str_finish(s, ix, iz, ela_str_load(s, ix, iz));
}

void launch_elastic_velocity(const ElaStep& s, hipStream_t st);
void launch_elastic_stress(const ElaStep& s, hipStream_t st);

}  // namespace tsw
