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
  const float* sponge;                   // nullptr: none
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

void launch_elastic_velocity(const ElaStep& s, hipStream_t st);
void launch_elastic_stress(const ElaStep& s, hipStream_t st);

}  // namespace tsw
