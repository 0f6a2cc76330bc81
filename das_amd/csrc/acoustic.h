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
  const float* sponge;           // per-cell sponge; nullptr: see sponge_table
  const float* sponge_table;     // used when sponge is null: table[depth], depth of
  int n_sponge;                  //   the cell into the strip (n_sponge wide); null: none
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

// Record / inject at cell (iz, ix) whose freshly stepped value is `v`.
// Returns true when an injection changed `v` (the cell must be written).
__device__ __forceinline__ bool acu_point_io(const AcuStep& s, int iz, int ix, float& v) {
  bool touched = false;
  if (iz == s.riz && s.roff != nullptr) {
    const int a = s.roff[ix], b = s.roff[ix + 1];
    if (a < b) {
      if (s.rec_it >= 0) {
        float* row = s.rec + (long long)s.rec_it * s.nrec;
        for (int r = a; r < b; ++r) row[s.ridx[r]] = v;
      }
      if (s.rdata_it >= 0) {
        const float* row = s.rdata + (long long)s.rdata_it * s.nrec;
        for (int r = a; r < b; ++r) v = v + row[s.ridx[r]] * s.inj;  // receiver order
        touched = true;
      }
    }
  }
  if (s.src_it >= 0 && iz * s.nx + ix == s.src_cell) {
    v = v + s.src[s.src_it] * s.inj;
    touched = true;
  }
  return touched;
}

// One cell of the fused step: side work on `cur`, then the leapfrog update
// into `nxt` (interior), the free-surface rows (done by the row-2 cell of each
// column), recording and next-iteration injection.  Cells outside
// [2, nz-2) x [2, nx-2) keep nxt's old content unless injected, as in the
// reference's three-buffer rotation (wave2d.c:44-81, wave_solver.py:183).
__device__ __forceinline__ void acu_cell(const AcuStep& s, int ix, int iz) {
  const int nx = s.nx;
  const int c = iz * nx + ix;
  const float pc = s.cur[c];
  if (s.snap != nullptr) {
    s.snap[c] = pc;
    s.illum[c] = s.illum[c] + pc * pc;
  }
  if (s.image != nullptr) s.image[c] = s.image[c] + (s.isnap[c] * pc) * s.w;
  if (s.fs && iz < 2) return;
  const bool interior = iz >= 2 && iz < s.nz - 2 && ix >= 2 && ix < nx - 2;
  float v;
  if (interior) {
    const float* p = s.cur;
    float lx = TSW_C2 * (p[c - 2] + p[c + 2]);
    lx = lx + TSW_C1 * (p[c - 1] + p[c + 1]);
    lx = (lx + TSW_C0 * pc) * s.rdx2;
    float lz = TSW_C2 * (p[c - 2 * nx] + p[c + 2 * nx]);
    lz = lz + TSW_C1 * (p[c - nx] + p[c + nx]);
    lz = (lz + TSW_C0 * pc) * s.rdz2;
    v = (2.0f * pc - s.prv[c]) + s.coef[c] * (lx + lz);
    if (s.sponge != nullptr) {
      v = v * s.sponge[c];
    } else if (s.sponge_table != nullptr) {
      const int d = sponge_depth(ix, iz, nx, s.nz, s.n_sponge, s.fs);
      if (d > 0) v = v * s.sponge_table[d];
    }
  } else {
    v = s.nxt[c];
  }
  if (s.fs && iz == 2) {
    float v0 = 0.0f;
    acu_point_io(s, 0, ix, v0);
    s.nxt[ix] = v0;
    float v1 = -v;
    acu_point_io(s, 1, ix, v1);
    s.nxt[nx + ix] = v1;
  }
  const bool injected = acu_point_io(s, iz, ix, v);
  if (interior || injected) s.nxt[c] = v;
}

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
