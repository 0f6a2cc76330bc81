// Elastic lane of a batch (see elabatch.hip).
#pragma once
#include <vector>

#include "engine.h"
#include "tswrtm.h"

namespace tsw {

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
