// Batched multi-shot acoustic RTM: device descriptors and the batch object.
#pragma once
#include <vector>

#include "engine.h"
#include "tswrtm.h"

namespace tsw {

// Device-side description of one shot of a batch.  All arrays are (nz, nx)
// row-major unless stated.  Shot time: global step g maps to local step
// m = g - start; forward iteration it = m for m < nt, then back-propagation
// iteration it = 2*nt - 1 - m for nt <= m < 2*nt.
struct ShotDev {
  int nx, nz, nt, k;
  int start;          // global step of the shot's first forward iteration
  int src_cell;       // siz*nx + six
  int riz, nrec;
  float* fwd[3];      // forward rotating buffers
  float* bwd[3];      // back-propagation rotating buffers
  const float* coef_f;  // (v*v)*dt2 of the forward model
  const float* coef_b;  // (v*v)*dt2 of the back-propagation model
  const float* src;     // (nt)
  const float* data;    // (nt, nrec)
  const int* roff;      // (nx+1)
  const int* ridx;      // (nrec)
  float* snaps;         // ((nt-1)/k+1, nz, nx)
  float* illum;         // sum of squared snapshots (scaled by dt*k at the end)
  float* image;
  int tiles_x, tiles;   // tile grid of this shot for the step kernel
};

// Launch-time list of the shots active during an epoch (a run of global
// steps with a fixed active set): shot ids and exclusive prefix of tiles.
struct Epoch {
  int g0, g1;         // global steps [g0, g1)
  int nshot;          // active shots
  int tiles;          // total tiles
  int* d_list;        // device: [nshot] shot ids then [nshot+1] tile prefix
};

// Blocked kernel configuration used when a caller does not choose one (the
// autotuner's choice for 300-cell-deep windows on MI355X).
constexpr int kDefaultBlockedKernel = 15;

// Stacks every shot's image and illumination of batch `h` into d_img / d_ill
// (gnz, gnx), in shot order at the shots' window offsets, on stream `st`.
int batch_stack_device(tsw_batch_t h, int gnx, int gnz, float* d_img, float* d_ill,
                       hipStream_t st);

// Constants shared by every shot of a batch.
struct BatchConst {
  int fs, n_sponge;
  float rdx2, rdz2, inj, w;
  const float* table;  // sponge table, n_sponge+1 entries
};

}  // namespace tsw
