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
};

void launch_elastic_velocity(const ElaStep& s, hipStream_t st);
void launch_elastic_stress(const ElaStep& s, hipStream_t st);

}  // namespace tsw
