// Elastic lane of a batch (see elabatch.hip).
#pragma once
#include <vector>

#include "engine.h"
#include "tswrtm.h"

namespace tsw {

struct EShot {
  int nx, nz, nt, k, start, src_cell, riz, nrec;
  float* f[5];          // forward vx, vz, sxx, szz, sxz: read by even steps
  float* fb[5];         // back-propagation vx, vz, sxx, szz, sxz: read by even steps
  float* f2[5];         // second forward set: read by odd steps
  float* fb2[5];        // second back-propagation set: read by odd steps
  const float* mf[4];   // forward model: lambda, mu, bx, bz
  const float* mb[4];   // back-propagation model
  const float* src;
  const float* data;
  const int* roff;
  const int* ridx;
  float* snaps;
  float* illum;
  float* image;
  int tiles_x, tiles;
};

struct EConst {
  int fs, n_sponge;
  float rdx, rdz, dt, w;
  const float* table;
};

// One launch of the batched elastic step (velocity and stress passes) over
// the shots in `list` (ids, then tile prefix) at global step g; shot time is
// m = g - start (m < 0: initialisation).  Step m of a propagation reads the
// set f/fb when m is even, f2/fb2 when odd, and writes the other.
void ela_batch_step(const EConst& c, const EShot* shots, const int* list, int nact, int ntiles,
                    int g, hipStream_t st);
int ela_tiles(int nx, int nz, int* tiles_x);
void ela_scale_launch(float* p, long long n, float w, hipStream_t st);

struct ElaLane;

int ela_lane_create(const std::vector<const tsw_shot_desc*>& shots, double dx, double dz,
                    double dt, int n_sponge, double alpha, int fs, int k, int slots,
                    int graph_steps, ElaLane** out);
// Enqueues a run on the lane's stream (TSW_RUN_BATCHED or TSW_RUN_PLAIN).
int ela_lane_launch(ElaLane* L, int mode);
int ela_lane_sync(ElaLane* L);
// One run with an event pair around every launch.
int ela_lane_time(ElaLane* L, double* ms, int* launches);
int ela_lane_result(ElaLane* L, int i, float* image_out, float* illum_out);
const float* ela_lane_image(ElaLane* L, int i);
const float* ela_lane_illum(ElaLane* L, int i);
long long ela_lane_cell_updates(ElaLane* L);
long long ela_lane_device_bytes(ElaLane* L);
int ela_lane_steps(ElaLane* L);
int ela_lane_graphs(ElaLane* L);
void ela_lane_destroy(ElaLane* L);

}  // namespace tsw
