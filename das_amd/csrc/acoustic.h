// Acoustic propagation kernels: descriptors and launch helpers.
#pragma once
#include "common.h"

namespace tsw {

// Everything one acoustic leapfrog step needs, including the per-step side
// work that the reference performs in Python between kernel calls
// (tsw_rtm/wave_solver.py:157-183 and :210-242):
//   * injection of the NEXT iteration's source / receiver samples into `nxt`
//     (the reference adds them into p_curr at the top of the next iteration;
//     p_curr of the next iteration is this step's p_next buffer),
//   * recording of the new level at the receivers (before that injection),
//   * snapshot + illumination of `cur` (post-injection, pre-step),
//   * zero-lag imaging of `cur` against a stored snapshot.
struct AcuStep {
  int nx, nz, fs;
  float rdx2, rdz2;              // 1/(dx*dx), 1/(dz*dz) in FP32 (wave2d.c:32-33)
  const float* coef;             // (v*v)*dt2 per cell
  const float* sponge;           // nullptr: no absorbing strip
  const float* cur;
  const float* prv;
  float* nxt;
  float inj;                     // FP32(dt*dt) injection scale (numpy weak promotion)
  // point source (forward): nxt[src_cell] += src[src_it] * inj when src_it >= 0
  int src_cell;
  int src_it;
  const float* src;
  // receivers on row riz, CSR by column: roff[ix]..roff[ix+1] into ridx
  int riz, nrec;
  const int* roff;
  const int* ridx;
  const float* rdata;            // back-injection samples (nt, nrec); used when rdata_it >= 0
  int rdata_it;
  float* rec;                    // recorded traces (nt, nrec); used when rec_it >= 0
  int rec_it;
  float* snap;                   // non-null: snap = cur, illum += cur*cur
  float* illum;
  const float* isnap;            // non-null: image += (isnap*cur)*w
  float* image;
  float w;
};

// Tile shape of the per-step kernel: one wavefront spans 64 columns.
constexpr int kAcuBX = 64;
constexpr int kAcuBZ = 4;

void launch_acoustic_step(const AcuStep& s, hipStream_t st);
// Adds the injection described by `s` (source sample src_it and/or receiver
// samples rdata_it) into s.nxt, without stepping.  Used for iteration 0.
void launch_acoustic_inject(const AcuStep& s, hipStream_t st);
void launch_fill(float* p, long long n, float v, hipStream_t st);
void launch_coef(const float* vel, float* coef, long long n, float dt2, hipStream_t st);
void launch_scale(float* p, long long n, float w, hipStream_t st);
void launch_correlate(const float* a, const float* b, float* img, long long n, float w,
                      hipStream_t st);
void launch_sponge(float* sp, int nx, int nz, int n, int fs, const float* table, hipStream_t st);

}  // namespace tsw
