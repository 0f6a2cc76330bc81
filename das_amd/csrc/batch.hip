// Batched acoustic RTM over many shots of uneven size and length.
//
// Every shot of a batch keeps its models, wavefields, snapshot arena and
// image in HBM.  Shots are packed into `slots` by longest-first list
// scheduling; a slot runs its shots back to back, each preceded by one
// initialisation step.  Each global time step is one kernel launch that
// covers the tiles of every shot active at that step (forward, back-
// propagation or initialisation), so one launch advances the whole batch.
// Launches are captured into hipGraphs of `graph_steps` steps and replayed.
// The temporally blocked kernels split the slots into launch groups, each
// with its own launch list and stream, so that the groups' launch boundaries
// do not coincide (see `kLaunchGroups`).
//
// Per-shot results are bit-identical to the single-shot path
// (tsw_rtm/rtm_engine.py:49-104 semantics): the cell update is the same
// device function.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "batch.h"
#include "elabatch.h"
#include "post.h"
#include "tb.h"
#include "tc.h"
#include "ts.h"
#include "tswrtm.h"

namespace tsw {

constexpr int kBX = 64;
constexpr int kBZ = 4;

// Launch groups of the temporally blocked kernels.  A launch ends with a tail
// in which fewer and fewer workgroups run, and the next launch on the stream
// starts with every workgroup in its load phase at once; with the slots in
// several groups on their own streams, one group's launches fill the others'
// ends.  On the batched 64-window workload (k_tc<5,6>) a step took 390 ms in
// one group and 372, 375, 364 and 368 ms in 2, 3, 8 and 16, results
// bit-identical.  The environment variable TSW_LAUNCH_GROUPS (1..64) sets
// another number (1: every launch on the batch's stream).
constexpr int kLaunchGroups = 8;
static int launch_groups() {
  const char* v = std::getenv("TSW_LAUNCH_GROUPS");
  if (!v || !*v) return kLaunchGroups;
  const int n = std::atoi(v);
  return n >= 1 && n <= 64 ? n : kLaunchGroups;
}

// Blocks are dispatched to the 8 XCDs round-robin; map them so that each XCD
// works on a contiguous range of tiles (neighbouring tiles share halos in
// that XCD's L2).
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, l = b >> 3;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

__global__ __launch_bounds__(kBX* kBZ) void k_batch_step(BatchConst bc, const ShotDev* shots,
                                                         const int* list, int nact, int ntiles,
                                                         int g) {
  const int b = xcd_tile(blockIdx.x, ntiles);
  const int* pre = list + nact;
  int lo = 0, hi = nact - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const ShotDev& S = shots[list[lo]];
  const int t = b - pre[lo];
  const int ix = (t % S.tiles_x) * kBX + threadIdx.x;
  const int iz = (t / S.tiles_x) * kBZ + threadIdx.y;
  if (ix >= S.nx || iz >= S.nz) return;
  const int nx = S.nx;
  const int c = iz * nx + ix;
  const int m = g - S.start;

  if (m < 0) {
    // initialisation step: zero state, then iteration-0 injections of both
    // propagations (source sample 0 forward, last data row backward)
    float f1 = 0.0f, b1 = 0.0f;
    if (c == S.src_cell) f1 = f1 + S.src[0] * bc.inj;
    if (iz == S.riz) {
      const int a = S.roff[ix], e = S.roff[ix + 1];
      const float* row = S.data + (long long)(S.nt - 1) * S.nrec;
      for (int r = a; r < e; ++r) b1 = b1 + row[S.ridx[r]] * bc.inj;
    }
    S.fwd[0][c] = 0.0f;
    S.fwd[1][c] = f1;
    S.fwd[2][c] = 0.0f;
    S.bwd[0][c] = 0.0f;
    S.bwd[1][c] = b1;
    S.bwd[2][c] = 0.0f;
    S.illum[c] = 0.0f;
    S.image[c] = 0.0f;
    return;
  }

  AcuStep s;
  s.nx = nx;
  s.nz = S.nz;
  s.fs = bc.fs;
  s.rdx2 = bc.rdx2;
  s.rdz2 = bc.rdz2;
  s.sponge = nullptr;
  s.sponge_table = bc.table;
  s.n_sponge = bc.n_sponge;
  s.inj = bc.inj;
  s.src_cell = S.src_cell;
  s.src = S.src;
  s.riz = S.riz;
  s.nrec = S.nrec;
  s.roff = S.roff;
  s.ridx = S.ridx;
  s.rdata = S.data;
  s.rec = nullptr;
  s.rec_it = -1;
  s.w = bc.w;
  if (m < S.nt) {
    const int it = m;
    s.coef = S.coef_f;
    s.prv = S.fwd[m % 3];
    s.cur = S.fwd[(m + 1) % 3];
    s.nxt = S.fwd[(m + 2) % 3];
    s.src_it = it + 1 < S.nt ? it + 1 : -1;
    s.rdata_it = -1;
    const bool snap = it % S.k == 0;
    s.snap = snap ? S.snaps + (long long)(it / S.k) * S.nx * S.nz : nullptr;
    s.illum = S.illum;
    s.isnap = nullptr;
    s.image = nullptr;
  } else {
    const int mb = m - S.nt;
    const int it = S.nt - 1 - mb;
    s.coef = S.coef_b;
    s.prv = S.bwd[mb % 3];
    s.cur = S.bwd[(mb + 1) % 3];
    s.nxt = S.bwd[(mb + 2) % 3];
    s.src_it = -1;
    s.rdata_it = it >= 1 ? it - 1 : -1;
    s.snap = nullptr;
    s.illum = nullptr;
    const bool img = it % S.k == 0;  // cursor == it/k (snapshot count (nt-1)/k+1)
    s.isnap = img ? S.snaps + (long long)(it / S.k) * S.nx * S.nz : nullptr;
    s.image = img ? S.image : nullptr;
  }
  acu_cell(s, ix, iz);
}

// illum *= w for one shot, then image and illumination added into the
// stacked arrays at the shot's window offset.
__global__ void k_finish_shot(float* illum, long long n, float w) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    illum[i] = illum[i] * w;
}

struct ShotHost {
  int nx, nz, nt, six, siz, riz, nrec, x0, z0, slot;
  long long cells() const { return (long long)nx * nz; }
  long long nsnap(int k) const { return (nt - 1) / k + 1; }
};

struct Batch {
  int nshot = 0, k = 1, slots = 1, graph_steps = 100;
  double dt = 0;
  BatchConst bc{};
  hipStream_t stream = nullptr;
  std::vector<ShotHost> hs;
  std::vector<ShotDev> sd;
  DevBuf d_shots, d_table, d_lists;
  // per shot: coef_f, coef_b, src, data, roff, ridx, illum, image
  std::vector<DevBuf> per_shot;
  // per slot: fwd[3], bwd[3], snaps
  std::vector<DevBuf> per_slot;
  std::vector<Epoch> epochs;
  int total_steps = 0;
  std::vector<hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_defs;
  long long cell_updates = 0;
  size_t device_bytes = 0;
  // temporally blocked kernel (kernel > 0): config, launch list, device tables
  int kernel = 0;
  std::vector<TbShot> tbs;
  struct TbLaunch {
    int ent_off, tile_off, ntiles;
  };
  // launch groups: slot s belongs to group s % groups; group q's launch L runs
  // the L-th block of each of its slots, on stream gstream[q] (group 0 on
  // `stream`), forked from `stream` and joined back into it by issue()
  int groups = 1;
  std::vector<std::vector<TbLaunch>> tbl;
  std::vector<hipStream_t> gstream;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  DevBuf d_tbshots, d_tbent, d_tbtile;
  // caller's shot index -> lane (0 acoustic, 1 elastic) and index in the lane
  int ntotal = 0;
  std::vector<int> lane_of, local_of;
  std::vector<ShotHost> ehs;  // elastic shots' geometry (stacking)
  ElaLane* ela = nullptr;
  ~Batch() {
    if (ela) ela_lane_destroy(ela);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto e : graphs) (void)hipGraphExecDestroy(e);
    for (auto gdef : graph_defs) (void)hipGraphDestroy(gdef);
    for (size_t q = 1; q < gstream.size(); ++q)
      if (gstream[q]) (void)hipStreamDestroy(gstream[q]);
    for (auto e : ev_join)
      if (e) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static int alloc(Batch* B, DevBuf& d, size_t bytes) {
  TSW_TRY(d.ensure(bytes));
  B->device_bytes += d.bytes;
  return kOk;
}

static void launch_step(Batch* B, const Epoch& e, int g, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_step, dim3(e.tiles), dim3(kBX, kBZ), 0, st, B->bc,
                     (const ShotDev*)B->d_shots.p, (const int*)e.d_list, e.nshot, e.tiles, g);
}

// Instantiated configurations of the temporally blocked kernel
// (T sub-steps, R rows per thread, WX waves across, WZ waves down).
typedef void (*TbLaunchFn)(const BatchConst&, const TbShot*, const TbEntry*, const TbTile*, int,
                           hipStream_t);
template <int T, int R, int WX, int WZ>
static void launch_tb(const BatchConst& bc, const TbShot* shots, const TbEntry* ents,
                      const TbTile* tiles, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL((k_tb<T, R, WX, WZ>), dim3(ntiles), dim3(64 * WX, WZ), 0, st, bc, shots,
                     ents, tiles);
}
struct TbCfg {
  int T, R, WX, WZ, W, H;  // k_ts: WZ = T + 1 wave rows (one per sub-step and the writer)
  TbLaunchFn fn;
  bool streaming;
  bool tc;  // time levels in LDS (tc.h)
};
#define TB_CFG(T, R, WX, WZ) \
  { T, R, WX, WZ, TbShape<T, R, WX, WZ>::W, TbShape<T, R, WX, WZ>::H, launch_tb<T, R, WX, WZ>, false, false }
// z-streaming configurations (ts.h): H rows per segment
TSW_TS_CONFIGS_A(TSW_TS_EXTERN)
TSW_TS_CONFIGS_B(TSW_TS_EXTERN)
#define TS_CFG(T, WX, H) {T, 2, WX, T + 1, TsShape<T, WX>::W, H, launch_ts<T, WX, H>, true, false},
#define TC_CFG(T, R) {T, R, 1, 16, TcShape<T, R>::W, TcShape<T, R>::H, launch_tc<T, R>, false, true},
static const TbCfg kTbCfgs[] = {
    TB_CFG(4, 8, 2, 8),  // 1: 128 x 64 extended, 112 x 48 core, 1024 threads
    TB_CFG(2, 8, 2, 4),  // 2: 128 x 32, 120 x 24 core
    TB_CFG(4, 8, 1, 8),  // 3: 64 x 64, 48 x 48 core
    TB_CFG(3, 8, 2, 8),  // 4: 128 x 64, 116 x 52 core
    TB_CFG(6, 8, 2, 8),  // 5: 128 x 64, 104 x 40 core
    TB_CFG(4, 6, 2, 8),  // 6: 128 x 48, 112 x 32 core
    TB_CFG(2, 4, 2, 8),  // 7: 128 x 32, 120 x 24 core
    TB_CFG(1, 4, 2, 4),  // 8: 128 x 16, 124 x 12 core (one step per launch)
    TB_CFG(2, 4, 1, 16),  // 9: 64 x 64, 56 x 56 core, 1024 threads
    TB_CFG(3, 4, 1, 16),  // 10: 64 x 64, 52 x 52 core
    TB_CFG(4, 4, 1, 16),  // 11: 64 x 64, 48 x 48 core
    TB_CFG(3, 4, 2, 8),   // 12: 128 x 32, 116 x 20 core
    TB_CFG(2, 6, 1, 16),  // 13: 64 x 96, 56 x 88 core
    TB_CFG(4, 6, 1, 16),  // 14: 64 x 96, 48 x 80 core
    TB_CFG(5, 4, 1, 16),  // 15: 64 x 64, 44 x 44 core
    TB_CFG(6, 4, 1, 16),  // 16: 64 x 64, 40 x 40 core
    // 17..24: z-streaming strips of 64 WX columns, depth segments of H rows
    TSW_TS_CONFIGS_A(TS_CFG) TSW_TS_CONFIGS_B(TS_CFG)
    // 25..28: time levels in LDS (tc.h), 64 x 16R extended, 1024 threads
    TC_CFG(5, 6)  // 25: 64 x 96, 44 x 76 core
    TC_CFG(4, 6)  // 26: 64 x 96, 48 x 80 core
    TC_CFG(5, 5)  // 27: 64 x 80, 44 x 60 core
    TC_CFG(6, 6)  // 28: 64 x 96, 40 x 72 core
};
constexpr int kNumTbCfgs = (int)(sizeof(kTbCfgs) / sizeof(kTbCfgs[0]));

int tb_config(int id, int* T, int* W, int* H) {
  if (id < 1 || id > kNumTbCfgs) return kErrArg;
  const TbCfg& c = kTbCfgs[id - 1];
  *T = c.T;
  *W = c.W;
  *H = c.H;
  return kOk;
}

int tb_max_sponge(int id) {
  return kTbCfgs[id - 1].streaming ? kTsMaxSponge : 1 << 30;  // k_ts keeps the table in LDS
}

void tb_launch(int id, const BatchConst& bc, const TbShot* shots, const TbEntry* ents,
               const TbTile* tiles, int ntiles, hipStream_t st) {
  kTbCfgs[id - 1].fn(bc, shots, ents, tiles, ntiles, st);
}

static void launch_tb_rec(Batch* B, const Batch::TbLaunch& L, hipStream_t st) {
  const TbCfg& c = kTbCfgs[B->kernel - 1];
  c.fn(B->bc, (const TbShot*)B->d_tbshots.p, (const TbEntry*)B->d_tbent.p + L.ent_off,
       (const TbTile*)B->d_tbtile.p + L.tile_off, L.ntiles, st);
}

// Longest-first list scheduling of shots onto slots; each shot gets one
// initialisation step before its first forward step.
static void schedule(Batch* B) {
  std::vector<int> order(B->nshot);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long long)B->hs[a].nt * B->hs[a].cells() > (long long)B->hs[b].nt * B->hs[b].cells();
  });
  std::vector<long long> free_at(B->slots, 0);
  for (int id : order) {
    int s = (int)(std::min_element(free_at.begin(), free_at.end()) - free_at.begin());
    B->hs[id].slot = s;
    B->sd[id].start = (int)free_at[s] + 1;
    free_at[s] += 1 + 2LL * B->hs[id].nt;
  }
  B->total_steps = (int)*std::max_element(free_at.begin(), free_at.end());
}

static int build_epochs(Batch* B) {
  std::vector<int> cuts{0, B->total_steps};
  for (int i = 0; i < B->nshot; ++i) {
    cuts.push_back(B->sd[i].start - 1);
    cuts.push_back(B->sd[i].start + 2 * B->hs[i].nt);
  }
  std::sort(cuts.begin(), cuts.end());
  cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
  std::vector<int> host_lists;
  std::vector<size_t> offs;
  for (size_t j = 0; j + 1 < cuts.size(); ++j) {
    const int g0 = cuts[j], g1 = cuts[j + 1];
    if (g0 >= B->total_steps) break;
    Epoch e{};
    e.g0 = g0;
    e.g1 = g1;
    std::vector<int> ids;
    for (int i = 0; i < B->nshot; ++i)
      if (g0 >= B->sd[i].start - 1 && g0 < B->sd[i].start + 2 * B->hs[i].nt) ids.push_back(i);
    e.nshot = (int)ids.size();
    std::vector<int> pre(ids.size() + 1, 0);
    for (size_t q = 0; q < ids.size(); ++q) pre[q + 1] = pre[q] + B->sd[ids[q]].tiles;
    e.tiles = pre.back();
    offs.push_back(host_lists.size());
    host_lists.insert(host_lists.end(), ids.begin(), ids.end());
    host_lists.insert(host_lists.end(), pre.begin(), pre.end());
    B->epochs.push_back(e);
  }
  TSW_TRY(alloc(B, B->d_lists, std::max<size_t>(host_lists.size(), 1) * sizeof(int)));
  TSW_HIP(hipMemcpy(B->d_lists.p, host_lists.data(), host_lists.size() * sizeof(int),
                    hipMemcpyHostToDevice));
  for (size_t j = 0; j < B->epochs.size(); ++j) B->epochs[j].d_list = B->d_lists.i() + offs[j];
  return kOk;
}

static const Epoch* epoch_at(const Batch* B, int g) {
  for (const auto& e : B->epochs)
    if (g >= e.g0 && g < e.g1) return &e;
  return nullptr;
}

static int issue(Batch* B, int g0, int g1, hipStream_t st);
static int launches_total(const Batch* B);

// hipGraphs of `graph_steps` time steps each (graph_steps/T blocked launches).
// A blocked graph holds every launch group's launches as parallel branches
// between a fork and a join (captured from issue()); on the batched 64-window
// workload that ran at 365 ms per step against 384 ms with each group's
// launches in graphs of their own, replayed on the group's stream.
static int capture_graphs(Batch* B) {
  const int total = launches_total(B);
  const int per = B->kernel > 0
                      ? std::max(1, B->graph_steps / kTbCfgs[B->kernel - 1].T)
                      : B->graph_steps;
  for (int g0 = 0; g0 < total; g0 += per) {
    const int g1 = std::min(total, g0 + per);
    TSW_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    const int rc = issue(B, g0, g1, B->stream);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(B->stream, &gr);
    if (rc != kOk || e != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      if (rc != kOk) return rc;
      set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
      return kErrHip;
    }
    B->graph_defs.push_back(gr);
    hipGraphExec_t ex;
    TSW_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
    B->graphs.push_back(ex);
  }
  return kOk;
}

static int replay_graphs(Batch* B, hipStream_t st) {
  for (auto ex : B->graphs) TSW_HIP(hipGraphLaunch(ex, st));
  return kOk;
}

static int finish_shots(Batch* B, hipStream_t st) {
  for (int i = 0; i < B->nshot; ++i) {
    const long long n = B->hs[i].cells();
    long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_finish_shot, dim3((unsigned)blocks), dim3(256), 0, st, B->sd[i].illum, n,
                       B->bc.w);
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}

// Launch list of the temporally blocked kernel.  Each shot is a forward
// phase of ceil(nt/T) blocks then a back-propagation phase of as many; the
// shots of a slot follow each other (slot order of the per-step schedule) and
// launch L advances every slot by its L-th block.
static int build_tb(Batch* B) {
  const TbCfg& cf = kTbCfgs[B->kernel - 1];
  const int T = cf.T;
  B->tbs.resize(B->nshot);
  std::vector<std::vector<int>> slot_shots(B->slots);
  {
    std::vector<int> order(B->nshot);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return B->sd[a].start < B->sd[b].start; });
    for (int id : order) slot_shots[B->hs[id].slot].push_back(id);
  }
  std::vector<int> tiles_x(B->nshot), tiles(B->nshot);
  for (int i = 0; i < B->nshot; ++i) {
    const ShotDev& S = B->sd[i];
    TbShot& P = B->tbs[i];
    std::memset(&P, 0, sizeof P);
    P.nx = S.nx;
    P.nz = S.nz;
    P.nt = S.nt;
    P.k = S.k;
    P.src_cell = S.src_cell;
    P.riz = S.riz;
    P.nrec = S.nrec;
    for (int j = 0; j < 3; ++j) {
      P.set[0][j] = S.fwd[j];
      P.set[1][j] = S.bwd[j];
    }
    P.coef_f = S.coef_f;
    P.coef_b = S.coef_b;
    P.src = S.src;
    P.data = S.data;
    P.roff = S.roff;
    P.ridx = S.ridx;
    P.snaps = S.snaps;
    P.illum = S.illum;
    P.image = S.image;
    tiles_x[i] = (S.nx + cf.W - 1) / cf.W;
    const int tiles_z = (S.nz + cf.H - 1) / cf.H;
    TSW_ARG(tiles_x[i] < (1 << 16) && tiles_z < (1 << 15), "shot %d: too many tiles for the blocked kernel", i);
    tiles[i] = tiles_x[i] * tiles_z;
  }
  // per-slot block sequences
  std::vector<std::vector<TbEntry>> seq(B->slots);
  for (int s = 0; s < B->slots; ++s) {
    for (int id : slot_shots[s]) {
      const int nt = B->hs[id].nt, k = B->sd[id].k;
      const int nb = (nt + T - 1) / T;
      for (int ph = 0; ph < 2; ++ph)
        for (int b = 0; b < nb; ++b) {
          TbEntry e{};
          e.shot = id;
          e.phase = ph;
          e.it0 = ph == 0 ? b * T : nt - 1 - b * T;
          e.nsub = ph == 0 ? std::min(T, nt - e.it0) : std::min(T, e.it0 + 1);
          e.flags = (b == 0 ? kTbFirst : 0) | (b == nb - 1 ? kTbLast : 0);
          e.rd = b & 1;
          e.tiles_x = tiles_x[id];
          tb_schedule(e, nt, k);
          seq[s].push_back(e);
        }
    }
  }
  // Launch L of a group runs the L-th block of each of the group's slots.  Its
  // workgroup table lists
  // (entry, tile) per workgroup in the XCD-aware order; it depends only on
  // the launch's sequence of shots, so launches with the same shots share it.
  size_t nl = 0;
  for (auto& q : seq) nl = std::max(nl, q.size());
  std::vector<TbEntry> ents;
  std::vector<TbTile> tab;
  std::map<std::vector<int>, int> tab_of;  // shots of a launch -> offset in `tab`
  B->tbl.assign(B->groups, {});
  for (size_t L = 0; L < nl; ++L)
    for (int grp = 0; grp < B->groups; ++grp) {
      Batch::TbLaunch rec{(int)ents.size(), 0, 0};
      std::vector<int> shots;
      for (int s = grp; s < B->slots; s += B->groups) {
        if (L >= seq[s].size()) continue;
        ents.push_back(seq[s][L]);
        shots.push_back(seq[s][L].shot);
      }
      if (shots.empty()) continue;  // the group's slots are all done
      std::vector<TbTile> lin;
      for (int e = 0; e < (int)shots.size(); ++e) {
        const int id = shots[e];
        for (int t = 0; t < tiles[id]; ++t) lin.push_back({e, (t % tiles_x[id]) | (t / tiles_x[id]) << 16});
      }
      rec.ntiles = (int)lin.size();
      auto it = tab_of.find(shots);
      if (it == tab_of.end()) {
        it = tab_of.emplace(shots, (int)tab.size()).first;
        for (int b = 0; b < rec.ntiles; ++b) tab.push_back(lin[xcd_tile_tb(b, rec.ntiles)]);
      }
      rec.tile_off = it->second;
      B->tbl[grp].push_back(rec);
    }
  TSW_TRY(alloc(B, B->d_tbshots, B->tbs.size() * sizeof(TbShot)));
  TSW_TRY(alloc(B, B->d_tbent, std::max<size_t>(ents.size(), 1) * sizeof(TbEntry)));
  TSW_TRY(alloc(B, B->d_tbtile, std::max<size_t>(tab.size(), 1) * sizeof(TbTile)));
  TSW_HIP(hipMemcpy(B->d_tbshots.p, B->tbs.data(), B->tbs.size() * sizeof(TbShot),
                    hipMemcpyHostToDevice));
  TSW_HIP(hipMemcpy(B->d_tbent.p, ents.data(), ents.size() * sizeof(TbEntry),
                    hipMemcpyHostToDevice));
  TSW_HIP(hipMemcpy(B->d_tbtile.p, tab.data(), tab.size() * sizeof(TbTile), hipMemcpyHostToDevice));
  return kOk;
}

// Issues every launch of a batched run (steps [g0, g1) of the per-step
// kernel, or launches [g0, g1) of the blocked kernel).
// A launch of the blocked kernel is the L-th block of every slot: one kernel
// per launch group, the groups on their own streams between a fork from `st`
// and a join back into it.
static int launches_total(const Batch* B) {
  if (B->kernel == 0) return B->total_steps;
  size_t n = 0;
  for (const auto& t : B->tbl) n = std::max(n, t.size());
  return (int)n;
}
static int issue(Batch* B, int g0, int g1, hipStream_t st) {
  if (B->kernel > 0) {
    const int G = B->groups;
    if (G > 1) {
      TSW_HIP(hipEventRecord(B->ev_fork, st));
      for (int q = 1; q < G; ++q) TSW_HIP(hipStreamWaitEvent(B->gstream[q], B->ev_fork, 0));
    }
    for (int g = g0; g < g1; ++g)
      for (int q = 0; q < G; ++q)
        if (g < (int)B->tbl[q].size()) launch_tb_rec(B, B->tbl[q][g], q ? B->gstream[q] : st);
    for (int q = 1; q < G; ++q) {
      TSW_HIP(hipEventRecord(B->ev_join[q], B->gstream[q]));
      TSW_HIP(hipStreamWaitEvent(st, B->ev_join[q], 0));
    }
    return kOk;
  }
  for (int g = g0; g < g1; ++g) {
    const Epoch* e = epoch_at(B, g);
    if (e && e->tiles > 0) launch_step(B, *e, g, st);
  }
  return kOk;
}

}  // namespace tsw

using namespace tsw;

extern "C" {

static int batch_create_acoustic(Batch* B, int nshot, const tsw_shot_desc* shots, double dx,
                                 double dz, double dt, int n_sponge, double sponge_alpha,
                                 int free_surface_top, int k, int slots, int graph_steps,
                                 int kernel);

int tsw_batch_create(int nshot, const tsw_shot_desc* shots, double dx, double dz, double dt,
                     int n_sponge, double sponge_alpha, int free_surface_top, int k, int slots,
                     int graph_steps, int kernel, tsw_batch_t* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_TRY(fail_injected());
  TSW_ARG(out && shots && nshot >= 1, "need at least one shot");
  TSW_ARG(k >= 1 && n_sponge >= 0 && dx > 0 && dz > 0 && dt > 0, "bad batch constants");
  TSW_ARG(kernel >= 0 && kernel <= kNumTbCfgs, "kernel %d outside [0, %d]", kernel, kNumTbCfgs);
  TSW_ARG(kernel == 0 || n_sponge <= tb_max_sponge(kernel),
          "kernel %d holds a sponge of at most %d cells", kernel, tb_max_sponge(kernel));
  *out = nullptr;
  std::vector<tsw_shot_desc> ac;
  std::vector<const tsw_shot_desc*> el;
  std::vector<int> lane_of(nshot), local_of(nshot);
  for (int i = 0; i < nshot; ++i) {
    const tsw_shot_desc& d = shots[i];
    TSW_ARG(d.kind == TSW_SHOT_ACOUSTIC || d.kind == TSW_SHOT_ELASTIC, "shot %d: bad kind", i);
    TSW_ARG(d.nx >= 5 && d.nz >= 5, "shot %d: grid %d x %d too small", i, d.nx, d.nz);
    TSW_ARG(d.nt >= 1 && d.src, "shot %d: missing source", i);
    TSW_ARG(d.six >= 0 && d.six < d.nx && d.siz >= 0 && d.siz < d.nz,
            "shot %d: source outside grid", i);
    TSW_ARG(d.riz >= 0 && d.riz < d.nz && d.nrec >= 0, "shot %d: bad receiver row", i);
    TSW_ARG(d.nrec == 0 || (d.rix && d.data), "shot %d: receivers need rix and data", i);
    for (int r = 0; r < d.nrec; ++r)
      TSW_ARG(d.rix[r] >= 0 && d.rix[r] < d.nx, "shot %d: receiver %d column %d outside [0, %d)",
              i, r, d.rix[r], d.nx);
    // every field stays under 4 GiB: the blocked kernels address it in bytes
    // with 32-bit offsets
    TSW_ARG((long long)d.nx * d.nz < (1LL << 30), "shot %d: grid too large", i);
    if (d.kind == TSW_SHOT_ACOUSTIC) {
      TSW_ARG(d.vel_fwd && d.vel_bwd, "shot %d: missing velocity models", i);
      lane_of[i] = 0;
      local_of[i] = (int)ac.size();
      ac.push_back(d);
    } else {
      for (int j = 0; j < 4; ++j)
        TSW_ARG(d.ela_fwd[j] && d.ela_bwd[j], "shot %d: missing elastic model", i);
      lane_of[i] = 1;
      local_of[i] = (int)el.size();
      el.push_back(&d);
    }
  }
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  auto* B = new Batch();
  B->ntotal = nshot;
  B->lane_of = lane_of;
  B->local_of = local_of;
  B->kernel = kernel;
  B->k = k;
  int rc = kOk;
  if (!ac.empty())
    rc = batch_create_acoustic(B, (int)ac.size(), ac.data(), dx, dz, dt, n_sponge, sponge_alpha,
                               free_surface_top, k, slots, graph_steps, kernel);
  if (rc == kOk && !el.empty()) {
    rc = ela_lane_create(el, dx, dz, dt, n_sponge, sponge_alpha, free_surface_top, k, slots,
                         graph_steps, &B->ela);
    for (const auto* d : el)
      B->ehs.push_back(ShotHost{d->nx, d->nz, d->nt, d->six, d->siz, d->riz, d->nrec, d->x0,
                                d->z0, 0});
  }
  if (rc != kOk) {
    delete B;
    return rc;
  }
  *out = reinterpret_cast<tsw_batch_t>(B);
  return kOk;
}

static int batch_create_acoustic(Batch* B, int nshot, const tsw_shot_desc* shots, double dx,
                                 double dz, double dt, int n_sponge, double sponge_alpha,
                                 int free_surface_top, int k, int slots, int graph_steps,
                                 int kernel) {
  auto fail = [&](int rc) { return rc; };  // the caller deletes the batch
  B->nshot = nshot;
  B->k = k;
  B->dt = dt;
  B->slots = std::max(1, std::min(slots <= 0 ? nshot : slots, nshot));
  B->graph_steps = graph_steps;
  B->kernel = kernel;
  if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess)
    return fail((set_error("hipStreamCreate failed"), kErrHip));
  if (kernel > 0) {
    B->groups = std::min(launch_groups(), B->slots);
    B->gstream.assign(B->groups, nullptr);
    B->ev_join.assign(B->groups, nullptr);
    B->gstream[0] = B->stream;
    for (int q = 1; q < B->groups; ++q)
      if (hipStreamCreateWithFlags(&B->gstream[q], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&B->ev_join[q], hipEventDisableTiming) != hipSuccess)
        return fail((set_error("launch group stream or event creation failed"), kErrHip));
    if (B->groups > 1 && hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming) != hipSuccess)
      return fail((set_error("launch group event creation failed"), kErrHip));
  }
  const float dxf = (float)dx, dzf = (float)dz, dtf = (float)dt;
  B->bc.fs = free_surface_top ? 1 : 0;
  B->bc.n_sponge = n_sponge;
  B->bc.rdx2 = 1.0f / (dxf * dxf);
  B->bc.rdz2 = 1.0f / (dzf * dzf);
  B->bc.inj = (float)(dt * dt);
  B->bc.w = (float)(dt * k);
  std::vector<float> tab;
  sponge_table(tab, n_sponge, (float)sponge_alpha);
  int rc = alloc(B, B->d_table, tab.size() * sizeof(float));
  if (rc) return fail(rc);
  if (hipMemcpy(B->d_table.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) !=
      hipSuccess)
    return fail((set_error("table upload failed"), kErrHip));
  B->bc.table = B->d_table.f();

  B->hs.resize(nshot);
  B->sd.resize(nshot);
  std::memset(B->sd.data(), 0, nshot * sizeof(ShotDev));
  for (int i = 0; i < nshot; ++i) {
    const tsw_shot_desc& d = shots[i];
    B->hs[i] = ShotHost{d.nx, d.nz, d.nt, d.six, d.siz, d.riz, d.nrec, d.x0, d.z0, 0};
    ShotDev& S = B->sd[i];
    S.nx = d.nx;
    S.nz = d.nz;
    S.nt = d.nt;
    S.k = k;
    S.src_cell = d.siz * d.nx + d.six;
    S.riz = d.riz;
    S.nrec = d.nrec;
    S.tiles_x = (d.nx + kBX - 1) / kBX;
    S.tiles = S.tiles_x * ((d.nz + kBZ - 1) / kBZ);
    B->cell_updates += 2LL * d.nt * d.nx * d.nz;
  }
  schedule(B);

  // per-slot state sized for the largest shot of the slot
  std::vector<long long> slot_cells(B->slots, 0), slot_snap(B->slots, 0);
  for (int i = 0; i < nshot; ++i) {
    const int s = B->hs[i].slot;
    slot_cells[s] = std::max(slot_cells[s], B->hs[i].cells());
    slot_snap[s] = std::max(slot_snap[s], B->hs[i].nsnap(k) * B->hs[i].cells());
  }
  B->per_slot.resize((size_t)B->slots * 7);
  for (int s = 0; s < B->slots; ++s) {
    for (int j = 0; j < 6; ++j)
      if ((rc = alloc(B, B->per_slot[s * 7 + j], slot_cells[s] * sizeof(float)))) return fail(rc);
    if ((rc = alloc(B, B->per_slot[s * 7 + 6], slot_snap[s] * sizeof(float)))) return fail(rc);
  }
  B->per_shot.resize((size_t)nshot * 8);
  DevBuf vel;
  for (int i = 0; i < nshot; ++i) {
    const tsw_shot_desc& d = shots[i];
    ShotDev& S = B->sd[i];
    const long long n = B->hs[i].cells();
    const int s = B->hs[i].slot;
    DevBuf* P = &B->per_shot[(size_t)i * 8];
    for (int j = 0; j < 3; ++j) {
      S.fwd[j] = B->per_slot[s * 7 + j].f();
      S.bwd[j] = B->per_slot[s * 7 + 3 + j].f();
    }
    S.snaps = B->per_slot[s * 7 + 6].f();
    RecvMap m;
    if ((rc = m.build(d.rix, d.nrec, d.nx))) return fail(rc);
    const size_t nb = (size_t)n * sizeof(float);
    if ((rc = alloc(B, P[0], nb)) || (rc = alloc(B, P[1], nb)) ||
        (rc = alloc(B, P[2], (size_t)d.nt * sizeof(float))) ||
        (rc = alloc(B, P[3], (size_t)d.nt * std::max(d.nrec, 1) * sizeof(float))) ||
        (rc = alloc(B, P[4], m.off.size() * sizeof(int))) ||
        (rc = alloc(B, P[5], m.idx.size() * sizeof(int))) || (rc = alloc(B, P[6], nb)) ||
        (rc = alloc(B, P[7], nb)) || (rc = vel.ensure(nb)))
      return fail(rc);
    const float* vels[2] = {d.vel_fwd, d.vel_bwd};
    for (int j = 0; j < 2; ++j) {
      if (hipMemcpy(vel.p, vels[j], nb, hipMemcpyHostToDevice) != hipSuccess)
        return fail((set_error("model upload failed"), kErrHip));
      launch_coef(vel.f(), P[j].f(), n, dtf * dtf, B->stream);
      if (hipStreamSynchronize(B->stream) != hipSuccess)
        return fail((set_error("coef kernel failed"), kErrHip));
    }
    if (hipMemcpy(P[2].p, d.src, (size_t)d.nt * sizeof(float), hipMemcpyHostToDevice) ||
        (d.nrec && hipMemcpy(P[3].p, d.data, (size_t)d.nt * d.nrec * sizeof(float),
                             hipMemcpyHostToDevice)) ||
        hipMemcpy(P[4].p, m.off.data(), m.off.size() * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(P[5].p, m.idx.data(), m.idx.size() * sizeof(int), hipMemcpyHostToDevice))
      return fail((set_error("shot %d upload failed", i), kErrHip));
    S.coef_f = P[0].f();
    S.coef_b = P[1].f();
    S.src = P[2].f();
    S.data = P[3].f();
    S.roff = P[4].i();
    S.ridx = P[5].i();
    S.illum = P[6].f();
    S.image = P[7].f();
  }
  if ((rc = alloc(B, B->d_shots, nshot * sizeof(ShotDev)))) return fail(rc);
  if (hipMemcpy(B->d_shots.p, B->sd.data(), nshot * sizeof(ShotDev), hipMemcpyHostToDevice))
    return fail((set_error("descriptor upload failed"), kErrHip));
  if ((rc = build_epochs(B))) return fail(rc);
  if (B->kernel > 0 && (rc = build_tb(B))) return fail(rc);
  if (B->graph_steps > 0 && (rc = capture_graphs(B))) return fail(rc);
  return kOk;
}

static int run_acoustic(Batch* B, int mode) {
  hipStream_t st = B->stream;
  bool per_step = true;
  if (mode == TSW_RUN_BATCHED) {
    if (!B->graphs.empty()) {
      TSW_TRY(replay_graphs(B, st));
    } else {
      TSW_TRY(issue(B, 0, launches_total(B), st));
    }
    per_step = B->kernel == 0;
  } else {
    // one shot at a time, one launch per step covering only that shot
    DevBuf one;
    int lst[3];
    TSW_TRY(one.ensure(sizeof lst));
    for (int i = 0; i < B->nshot; ++i) {
      lst[0] = i;  // shot id, then its tile prefix {0, tiles}
      lst[1] = 0;
      lst[2] = B->sd[i].tiles;
      TSW_HIP(hipMemcpyAsync(one.p, lst, sizeof lst, hipMemcpyHostToDevice, st));
      TSW_HIP(hipStreamSynchronize(st));
      Epoch e{};
      e.nshot = 1;
      e.tiles = B->sd[i].tiles;
      e.d_list = one.i();
      for (int m = -1; m < 2 * B->hs[i].nt; ++m) launch_step(B, e, B->sd[i].start + m, st);
      TSW_HIP(hipStreamSynchronize(st));
    }
  }
  // the per-step kernel leaves illumination unscaled; the blocked kernel
  // scales it in the last forward block
  if (per_step) TSW_TRY(finish_shots(B, st));
  TSW_HIP(hipGetLastError());
  return kOk;
}

static const ShotHost& shot_host(const Batch* B, int i);
static const float* shot_image(Batch* B, int i);
static const float* shot_illum(Batch* B, int i);

int tsw_batch_run(tsw_batch_t h, int mode) {
  Batch* B = reinterpret_cast<Batch*>(h);
  TSW_ARG(B, "null batch");
  TSW_ARG(mode == TSW_RUN_BATCHED || mode == TSW_RUN_PLAIN, "unknown run mode %d", mode);
  TSW_TRY(fail_injected());
  // the elastic lane is enqueued first on its own stream so both lanes overlap
  if (B->ela) TSW_TRY(ela_lane_launch(B->ela, mode));
  if (B->nshot > 0) TSW_TRY(run_acoustic(B, mode));
  if (B->nshot > 0) TSW_HIP(hipStreamSynchronize(B->stream));
  if (B->ela) TSW_TRY(ela_lane_sync(B->ela));
  // divergence sentinel: every shot's image and illumination must be finite
  std::vector<const float*> arr;
  std::vector<long long> len;
  for (int i = 0; i < B->ntotal; ++i) {
    arr.push_back(shot_image(B, i));
    arr.push_back(shot_illum(B, i));
    len.push_back(shot_host(B, i).cells());
    len.push_back(shot_host(B, i).cells());
  }
  return check_finite(arr.data(), len.data(), (int)arr.size(), 2, 0, "shot",
                      B->nshot > 0 ? B->stream : nullptr);
}

int tsw_batch_time_kernels(tsw_batch_t h, int lane, double* kernel_ms, int* launches) {
  Batch* B = reinterpret_cast<Batch*>(h);
  TSW_ARG(B && kernel_ms && launches, "null argument");
  TSW_ARG(lane == 0 || lane == 1, "lane must be 0 (acoustic) or 1 (elastic)");
  if (lane == 1) {
    TSW_ARG(B->ela, "batch has no elastic shots");
    return ela_lane_time(B->ela, kernel_ms, launches);
  }
  TSW_ARG(B->nshot > 0, "batch has no acoustic shots");
  hipStream_t st = B->stream;
  if (B->kernel > 0 && !B->graphs.empty()) {
    // the blocked kernel as a batched run issues it: the batch's hipGraphs
    // (the launch groups as parallel branches) inside one event pair
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess ? kOk : kErrHip;
    float ms = 0.0f;
    if (rc == kOk && hipEventRecord(e0, st) != hipSuccess) rc = kErrHip;
    if (rc == kOk) rc = replay_graphs(B, st);
    if (rc == kOk && hipEventRecord(e1, st) != hipSuccess) rc = kErrHip;
    if (rc == kOk && (hipEventSynchronize(e1) != hipSuccess ||
                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = kErrHip;
    (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc != kOk) {
      set_error("kernel timing run failed");
      return rc;
    }
    *kernel_ms = ms;
    *launches = launches_total(B);
    return kOk;
  }
  // one event pair around each run of kTimedRun back-to-back launches (an
  // event between every two launches would add its own cost to short ones);
  // the pairs are drained in chunks to bound the pool.  The blocked kernel's
  // launch groups run a whole run between one fork and one join, as in the
  // hipGraphs, so consecutive launches overlap as they do there.
  const int chunk = 512;
  std::vector<hipEvent_t> ev(2 * chunk);
  for (auto& e : ev) TSW_HIP(hipEventCreate(&e));
  double total = 0.0;
  int count = 0, used = 0, inrun = 0;
  auto drain = [&]() -> int {
    TSW_HIP(hipEventSynchronize(ev[2 * used - 1]));
    for (int j = 0; j < used; ++j) {
      float ms = 0.0f;
      TSW_HIP(hipEventElapsedTime(&ms, ev[2 * j], ev[2 * j + 1]));
      total += ms;
    }
    used = 0;
    return kOk;
  };
  int rc = kOk;
  const int nlaunch = launches_total(B);
  for (int g = 0; g < nlaunch && rc == kOk; ++g) {
    if (B->kernel == 0) {
      const Epoch* e = epoch_at(B, g);
      if (!e || e->tiles == 0) continue;
    }
    if (inrun == 0 && hipEventRecord(ev[2 * used], st) != hipSuccess) rc = kErrHip;
    if (rc == kOk && B->kernel == 0) {
      rc = issue(B, g, g + 1, st);
    } else if (rc == kOk && inrun == 0) {
      rc = issue(B, g, std::min(nlaunch, g + kTimedRun), st);
    }
    ++count;
    if (++inrun == kTimedRun) {
      if (hipEventRecord(ev[2 * used + 1], st) != hipSuccess) rc = kErrHip;
      inrun = 0;
      if (++used == chunk && rc == kOk) rc = drain();
    }
  }
  if (rc == kOk && inrun) {
    if (hipEventRecord(ev[2 * used + 1], st) != hipSuccess) rc = kErrHip;
    ++used;
  }
  if (rc == kOk && used) rc = drain();
  if (rc == kOk && B->kernel == 0) rc = finish_shots(B, st);
  (void)hipStreamSynchronize(st);
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc != kOk) {
    set_error("kernel timing run failed");
    return rc;
  }
  *kernel_ms = total;
  *launches = count;
  return kOk;
}

static const ShotHost& shot_host(const Batch* B, int i) {
  return B->lane_of[i] == 0 ? B->hs[B->local_of[i]] : B->ehs[B->local_of[i]];
}
static const float* shot_image(Batch* B, int i) {
  const int l = B->local_of[i];
  return B->lane_of[i] == 0 ? B->sd[l].image : ela_lane_image(B->ela, l);
}
static const float* shot_illum(Batch* B, int i) {
  const int l = B->local_of[i];
  return B->lane_of[i] == 0 ? B->sd[l].illum : ela_lane_illum(B->ela, l);
}

int tsw_batch_shot_result(tsw_batch_t h, int shot, float* image_out, float* illum_out) {
  Batch* B = reinterpret_cast<Batch*>(h);
  TSW_ARG(B && shot >= 0 && shot < B->ntotal, "bad shot index");
  const size_t nb = (size_t)shot_host(B, shot).cells() * sizeof(float);
  if (image_out) TSW_HIP(hipMemcpy(image_out, shot_image(B, shot), nb, hipMemcpyDeviceToHost));
  if (illum_out) TSW_HIP(hipMemcpy(illum_out, shot_illum(B, shot), nb, hipMemcpyDeviceToHost));
  return kOk;
}

}  // extern "C"

namespace tsw {
int batch_stack_device(tsw_batch_t h, int gnx, int gnz, float* d_img, float* d_ill,
                       hipStream_t st) {
  Batch* B = reinterpret_cast<Batch*>(h);
  TSW_ARG(B && gnx >= 1 && gnz >= 1 && d_img && d_ill, "bad argument");
  for (int i = 0; i < B->ntotal; ++i) {
    const ShotHost& s = shot_host(B, i);
    TSW_ARG(s.x0 >= 0 && s.z0 >= 0 && s.x0 + s.nx <= gnx && s.z0 + s.nz <= gnz,
            "shot window outside the %d x %d stack", gnx, gnz);
  }
  const long long n = (long long)gnx * gnz;
  launch_fill(d_img, n, 0.0f, st);
  launch_fill(d_ill, n, 0.0f, st);
  for (int i = 0; i < B->ntotal; ++i) {  // shot order, as the reference's stack loop
    const ShotHost& s = shot_host(B, i);
    launch_stack_window(d_img + (long long)s.z0 * gnx, gnx, shot_image(B, i), s.nx, s.nz, s.x0, 0,
                        st);
    launch_stack_window(d_ill + (long long)s.z0 * gnx, gnx, shot_illum(B, i), s.nx, s.nz, s.x0, 0,
                        st);
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}
}  // namespace tsw

extern "C" {

int tsw_batch_stack(tsw_batch_t h, int gnx, int gnz, float* image_out, float* illum_out) {
  TSW_ARG(image_out && illum_out, "bad argument");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = (long long)gnx * gnz;
  DevBuf img, ill;
  TSW_TRY(img.ensure(n * sizeof(float)));
  TSW_TRY(ill.ensure(n * sizeof(float)));
  TSW_TRY(batch_stack_device(h, gnx, gnz, img.f(), ill.f(), st));
  TSW_HIP(hipMemcpyAsync(image_out, img.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipMemcpyAsync(illum_out, ill.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_batch_stack_device(tsw_batch_t h, int gnx, int gnz, float* d_image, float* d_illum,
                           void* stream) {
  TSW_ARG(d_image && d_illum, "bad argument");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return batch_stack_device(h, gnx, gnz, d_image, d_illum, st);
}

int tsw_batch_info(tsw_batch_t h, long long* cell_updates, long long* device_bytes,
                   int* total_steps, int* n_graphs, long long* elastic_cell_updates) {
  Batch* B = reinterpret_cast<Batch*>(h);
  TSW_ARG(B, "null batch");
  const long long ecu = B->ela ? ela_lane_cell_updates(B->ela) : 0;
  if (cell_updates) *cell_updates = B->cell_updates + ecu;
  if (device_bytes)
    *device_bytes = (long long)B->device_bytes + (B->ela ? ela_lane_device_bytes(B->ela) : 0);
  if (total_steps)
    *total_steps = std::max(B->total_steps, B->ela ? ela_lane_steps(B->ela) : 0);
  if (n_graphs) *n_graphs = (int)B->graphs.size() + (B->ela ? ela_lane_graphs(B->ela) : 0);
  if (elastic_cell_updates) *elastic_cell_updates = ecu;
  return kOk;
}

int tsw_kernel_configs(void) { return kNumTbCfgs; }

int tsw_kernel_name(int id, char* buf, int len) {
  TSW_ARG(id >= 1 && id <= kNumTbCfgs, "kernel config %d outside [1, %d]", id, kNumTbCfgs);
  TSW_ARG(buf && len > 0, "null buffer");
  const TbCfg& c = kTbCfgs[id - 1];
  if (c.streaming) snprintf(buf, len, "k_ts<%d,%d,%d>", c.T, c.WX, c.H);
  else if (c.tc) snprintf(buf, len, "k_tc<%d,%d>", c.T, c.R);
  else snprintf(buf, len, "k_tb<%d,%d,%d,%d>", c.T, c.R, c.WX, c.WZ);
  return kOk;
}

int tsw_kernel_config(int id, int* T, int* W, int* H, int* threads) {
  TSW_ARG(id >= 1 && id <= kNumTbCfgs, "kernel config %d outside [1, %d]", id, kNumTbCfgs);
  const TbCfg& c = kTbCfgs[id - 1];
  if (T) *T = c.T;
  if (W) *W = c.W;
  if (H) *H = c.H;
  if (threads) *threads = 64 * c.WX * c.WZ;
  return kOk;
}

int tsw_batch_destroy(tsw_batch_t h) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  delete reinterpret_cast<Batch*>(h);
  return kOk;
}

}  // extern "C"
