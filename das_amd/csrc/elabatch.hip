// Elastic lane of a batch: many elastic shots of uneven size and length,
// resident in HBM, advanced together.  Each global time step is two launches
// (velocity pass, then stress pass) covering the tiles of every active elastic
// shot; the steps are captured into hipGraphs and replayed on the lane's own
// stream, so an elastic lane overlaps with the acoustic lane of the same batch.
// The per-cell updates are the device functions of the single-shot elastic
// path, so per-shot results are bit-identical to tsw_migrate_shot_elastic.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "elabatch.h"
#include "elastic.h"

namespace tsw {

constexpr int kEBX = 64;
constexpr int kEBZ = 4;
// rows per thread, kEBZ apart: a tile is kEBX x (kEBZ * kERows) cells
constexpr int kERows = 2;
constexpr int kETileZ = kEBZ * kERows;

__device__ __forceinline__ int xcd_tile_e(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, l = b >> 3;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

// pass 0: velocities (with injection and side work), pass 1: stresses.
__global__ __launch_bounds__(kEBX* kEBZ) void k_ela_batch(EConst c, const EShot* shots,
                                                         const int* list, int nact, int ntiles,
                                                         int g, int pass) {
  const int b = xcd_tile_e(blockIdx.x, ntiles);
  const int* pre = list + nact;
  int lo = 0, hi = nact - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const EShot& S = shots[list[lo]];
  const int t = b - pre[lo];
  const int ix = (t % S.tiles_x) * kEBX + threadIdx.x;
  const int iz0 = (t / S.tiles_x) * kETileZ + threadIdx.y;
  if (ix >= S.nx || iz0 >= S.nz) return;
  const int m = g - S.start;
  if (m < 0) {  // initialisation step of the shot
    if (pass == 0) {
      for (int r = 0; r < kERows; ++r) {
        const int iz = iz0 + r * kEBZ;
        if (iz >= S.nz) break;
        const int i = iz * S.nx + ix;
        for (int j = 0; j < 5; ++j) {
          S.f[j][i] = 0.0f;
          S.fb[j][i] = 0.0f;
        }
        S.illum[i] = 0.0f;
        S.image[i] = 0.0f;
      }
    }
    return;
  }
  ElaStep s;
  s.nx = S.nx;
  s.nz = S.nz;
  s.fs = c.fs;
  s.rdx = c.rdx;
  s.rdz = c.rdz;
  s.dt = c.dt;
  const bool fwd = m < S.nt;
  float* const* fld = fwd ? S.f : S.fb;
  s.vx = fld[0];
  s.vz = fld[1];
  s.sxx = fld[2];
  s.szz = fld[3];
  s.sxz = fld[4];
  s.sponge = nullptr;
  s.sponge_table = c.table;
  s.n_sponge = c.n_sponge;
  s.src_cell = S.src_cell;
  s.src = S.src;
  s.src_it = -1;
  s.riz = S.riz;
  s.nrec = S.nrec;
  s.roff = S.roff;
  s.ridx = S.ridx;
  s.rdata = S.data;
  s.rdata_it = -1;
  s.rec = nullptr;
  s.rec_it = -1;
  s.snap = nullptr;
  s.illum = S.illum;
  s.isnap = nullptr;
  s.image = nullptr;
  s.w = c.w;
  const long long n = (long long)S.nx * S.nz;
  const float* const* model;
  if (fwd) {
    const int it = m;
    model = S.mf;
    if (pass == 0) {
      s.src_it = it;
      if (it % S.k == 0) s.snap = S.snaps + (long long)(it / S.k) * n;
    }
  } else {
    const int it = 2 * S.nt - 1 - m;
    model = S.mb;
    if (pass == 0) {
      s.rdata_it = it;
      if (it % S.k == 0) {  // cursor == it/k for (nt-1)/k+1 snapshots
        s.isnap = S.snaps + (long long)(it / S.k) * n;
        s.image = S.image;
      }
    }
  }
  s.lam = model[0];
  s.mu = model[1];
  s.bx = model[2];
  s.bz = model[3];
  // every row's reads first, then the rows' updates (see ela_vel_load)
  if (pass == 0) {
    ElaVelIn in[kERows];
#pragma unroll
    for (int r = 0; r < kERows; ++r)
      if (iz0 + r * kEBZ < S.nz) in[r] = ela_vel_load(s, ix, iz0 + r * kEBZ);
#pragma unroll
    for (int r = 0; r < kERows; ++r)
      if (iz0 + r * kEBZ < S.nz) ela_vel_finish(s, ix, iz0 + r * kEBZ, in[r]);
  } else {
    ElaStrIn in[kERows];
#pragma unroll
    for (int r = 0; r < kERows; ++r)
      if (iz0 + r * kEBZ < S.nz) in[r] = ela_str_load(s, ix, iz0 + r * kEBZ);
#pragma unroll
    for (int r = 0; r < kERows; ++r)
      if (iz0 + r * kEBZ < S.nz) ela_str_finish(s, ix, iz0 + r * kEBZ, in[r]);
  }
}

void ela_batch_launch(const EConst& c, const EShot* shots, const int* list, int nact,
                      int ntiles, int g, int pass, hipStream_t st) {
  hipLaunchKernelGGL(k_ela_batch, dim3(ntiles), dim3(kEBX, kEBZ), 0, st, c, shots, list, nact,
                     ntiles, g, pass);
}

int ela_tiles(int nx, int nz, int* tiles_x) {
  *tiles_x = (nx + kEBX - 1) / kEBX;
  return *tiles_x * ((nz + kETileZ - 1) / kETileZ);
}

__global__ void k_ela_scale(float* p, long long n, float w) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = p[i] * w;
}

struct EEpoch {
  int g0, g1, nshot, tiles;
  int* d_list;
};

struct ElaLane {
  int nshot = 0, slots = 1, graph_steps = 100, k = 1;
  EConst c{};
  hipStream_t stream = nullptr;
  std::vector<EShot> sh;
  std::vector<int> slot_of;
  std::vector<DevBuf> per_shot, per_slot;
  DevBuf d_shots, d_table, d_lists;
  std::vector<EEpoch> epochs;
  int total_steps = 0;
  std::vector<hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_defs;
  long long cell_updates = 0;
  size_t device_bytes = 0;
  ~ElaLane() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto e : graphs) (void)hipGraphExecDestroy(e);
    for (auto gd : graph_defs) (void)hipGraphDestroy(gd);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static int ealloc(ElaLane* L, DevBuf& d, size_t bytes) {
  TSW_TRY(d.ensure(bytes));
  L->device_bytes += d.bytes;
  return kOk;
}

static void e_launch(ElaLane* L, const EEpoch& e, int g, hipStream_t st) {
  for (int pass = 0; pass < 2; ++pass)
    hipLaunchKernelGGL(k_ela_batch, dim3(e.tiles), dim3(kEBX, kEBZ), 0, st, L->c,
                       (const EShot*)L->d_shots.p, (const int*)e.d_list, e.nshot, e.tiles, g,
                       pass);
}

static const EEpoch* e_epoch_at(const ElaLane* L, int g) {
  for (const auto& e : L->epochs)
    if (g >= e.g0 && g < e.g1) return &e;
  return nullptr;
}

static void e_issue(ElaLane* L, int g0, int g1, hipStream_t st) {
  for (int g = g0; g < g1; ++g) {
    const EEpoch* e = e_epoch_at(L, g);
    if (e && e->tiles > 0) e_launch(L, *e, g, st);
  }
}

static int e_build(ElaLane* L) {
  // longest-first list scheduling onto slots, one initialisation step each
  std::vector<int> order(L->nshot);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long long)L->sh[a].nt * L->sh[a].nx * L->sh[a].nz >
           (long long)L->sh[b].nt * L->sh[b].nx * L->sh[b].nz;
  });
  std::vector<long long> free_at(L->slots, 0);
  L->slot_of.assign(L->nshot, 0);
  for (int id : order) {
    const int s = (int)(std::min_element(free_at.begin(), free_at.end()) - free_at.begin());
    L->slot_of[id] = s;
    L->sh[id].start = (int)free_at[s] + 1;
    free_at[s] += 1 + 2LL * L->sh[id].nt;
  }
  L->total_steps = (int)*std::max_element(free_at.begin(), free_at.end());
  std::vector<int> cuts{0, L->total_steps};
  for (const auto& S : L->sh) {
    cuts.push_back(S.start - 1);
    cuts.push_back(S.start + 2 * S.nt);
  }
  std::sort(cuts.begin(), cuts.end());
  cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
  std::vector<int> lists;
  std::vector<size_t> offs;
  for (size_t j = 0; j + 1 < cuts.size(); ++j) {
    if (cuts[j] >= L->total_steps) break;
    EEpoch e{};
    e.g0 = cuts[j];
    e.g1 = cuts[j + 1];
    std::vector<int> ids;
    for (int i = 0; i < L->nshot; ++i)
      if (e.g0 >= L->sh[i].start - 1 && e.g0 < L->sh[i].start + 2 * L->sh[i].nt) ids.push_back(i);
    e.nshot = (int)ids.size();
    std::vector<int> pre(ids.size() + 1, 0);
    for (size_t q = 0; q < ids.size(); ++q) pre[q + 1] = pre[q] + L->sh[ids[q]].tiles;
    e.tiles = pre.back();
    offs.push_back(lists.size());
    lists.insert(lists.end(), ids.begin(), ids.end());
    lists.insert(lists.end(), pre.begin(), pre.end());
    L->epochs.push_back(e);
  }
  TSW_TRY(ealloc(L, L->d_lists, std::max<size_t>(lists.size(), 1) * sizeof(int)));
  TSW_HIP(hipMemcpy(L->d_lists.p, lists.data(), lists.size() * sizeof(int),
                    hipMemcpyHostToDevice));
  for (size_t j = 0; j < L->epochs.size(); ++j) L->epochs[j].d_list = L->d_lists.i() + offs[j];
  return kOk;
}

void ela_scale_launch(float* p, long long n, float w, hipStream_t st) {
  hipLaunchKernelGGL(k_ela_scale, dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)),
                     dim3(256), 0, st, p, n, w);
}

static int e_finish(ElaLane* L, hipStream_t st) {
  for (const auto& S : L->sh) {
    const long long n = (long long)S.nx * S.nz;
    hipLaunchKernelGGL(k_ela_scale, dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)),
                       dim3(256), 0, st, S.illum, n, L->c.w);
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}

int ela_lane_create(const std::vector<const tsw_shot_desc*>& shots, double dx, double dz,
                    double dt, int n_sponge, double alpha, int fs, int k, int slots,
                    int graph_steps, ElaLane** out) {
  *out = nullptr;
  auto* L = new ElaLane();
  auto fail = [&](int rc) {
    delete L;
    return rc;
  };
  L->nshot = (int)shots.size();
  L->k = k;
  L->slots = std::max(1, std::min(slots <= 0 ? L->nshot : slots, L->nshot));
  L->graph_steps = graph_steps;
  if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess)
    return fail((set_error("hipStreamCreate failed"), kErrHip));
  L->c.fs = fs ? 1 : 0;
  L->c.n_sponge = n_sponge;
  L->c.rdx = 1.0f / (float)dx;
  L->c.rdz = 1.0f / (float)dz;
  L->c.dt = (float)dt;
  L->c.w = (float)(dt * k);
  std::vector<float> tab;
  sponge_table(tab, n_sponge, (float)alpha);
  int rc;
  if ((rc = ealloc(L, L->d_table, tab.size() * sizeof(float)))) return fail(rc);
  if (hipMemcpy(L->d_table.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice))
    return fail((set_error("table upload failed"), kErrHip));
  L->c.table = L->d_table.f();
  L->sh.resize(L->nshot);
  for (int i = 0; i < L->nshot; ++i) {
    const tsw_shot_desc& d = *shots[i];
    EShot& S = L->sh[i];
    std::memset(&S, 0, sizeof S);
    S.nx = d.nx;
    S.nz = d.nz;
    S.nt = d.nt;
    S.k = k;
    S.src_cell = d.siz * d.nx + d.six;
    S.riz = d.riz;
    S.nrec = d.nrec;
    S.tiles_x = (d.nx + kEBX - 1) / kEBX;
    S.tiles = S.tiles_x * ((d.nz + kETileZ - 1) / kETileZ);
    L->cell_updates += 2LL * d.nt * d.nx * d.nz;
  }
  if ((rc = e_build(L))) return fail(rc);
  // per slot: 5 forward + 5 back-propagation fields + vz snapshot arena, sized
  // for the slot's largest shot
  std::vector<long long> cells(L->slots, 0), snap(L->slots, 0);
  for (int i = 0; i < L->nshot; ++i) {
    const long long n = (long long)L->sh[i].nx * L->sh[i].nz;
    const int s = L->slot_of[i];
    cells[s] = std::max(cells[s], n);
    snap[s] = std::max(snap[s], (long long)((L->sh[i].nt - 1) / k + 1) * n);
  }
  L->per_slot.resize((size_t)L->slots * 11);
  for (int s = 0; s < L->slots; ++s) {
    for (int j = 0; j < 10; ++j)
      if ((rc = ealloc(L, L->per_slot[s * 11 + j], cells[s] * sizeof(float)))) return fail(rc);
    if ((rc = ealloc(L, L->per_slot[s * 11 + 10], snap[s] * sizeof(float)))) return fail(rc);
  }
  // per shot: 8 model arrays, src, data, roff, ridx, illum, image
  L->per_shot.resize((size_t)L->nshot * 14);
  for (int i = 0; i < L->nshot; ++i) {
    const tsw_shot_desc& d = *shots[i];
    EShot& S = L->sh[i];
    DevBuf* P = &L->per_shot[(size_t)i * 14];
    const size_t nb = (size_t)d.nx * d.nz * sizeof(float);
    RecvMap m;
    if ((rc = m.build(d.rix, d.nrec, d.nx))) return fail(rc);
    for (int j = 0; j < 4; ++j) {
      if (!d.ela_fwd[j] || !d.ela_bwd[j])
        return fail((set_error("elastic shot %d: model array missing", i), kErrArg));
      if ((rc = ealloc(L, P[j], nb)) || (rc = ealloc(L, P[4 + j], nb))) return fail(rc);
      if (hipMemcpy(P[j].p, d.ela_fwd[j], nb, hipMemcpyHostToDevice) ||
          hipMemcpy(P[4 + j].p, d.ela_bwd[j], nb, hipMemcpyHostToDevice))
        return fail((set_error("elastic model upload failed"), kErrHip));
      S.mf[j] = P[j].f();
      S.mb[j] = P[4 + j].f();
    }
    if ((rc = ealloc(L, P[8], (size_t)d.nt * sizeof(float))) ||
        (rc = ealloc(L, P[9], (size_t)d.nt * std::max(d.nrec, 1) * sizeof(float))) ||
        (rc = ealloc(L, P[10], m.off.size() * sizeof(int))) ||
        (rc = ealloc(L, P[11], m.idx.size() * sizeof(int))) || (rc = ealloc(L, P[12], nb)) ||
        (rc = ealloc(L, P[13], nb)))
      return fail(rc);
    if (hipMemcpy(P[8].p, d.src, (size_t)d.nt * sizeof(float), hipMemcpyHostToDevice) ||
        (d.nrec && hipMemcpy(P[9].p, d.data, (size_t)d.nt * d.nrec * sizeof(float),
                             hipMemcpyHostToDevice)) ||
        hipMemcpy(P[10].p, m.off.data(), m.off.size() * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(P[11].p, m.idx.data(), m.idx.size() * sizeof(int), hipMemcpyHostToDevice))
      return fail((set_error("elastic shot %d upload failed", i), kErrHip));
    S.src = P[8].f();
    S.data = P[9].f();
    S.roff = P[10].i();
    S.ridx = P[11].i();
    S.illum = P[12].f();
    S.image = P[13].f();
    const int s = L->slot_of[i];
    for (int j = 0; j < 5; ++j) {
      S.f[j] = L->per_slot[s * 11 + j].f();
      S.fb[j] = L->per_slot[s * 11 + 5 + j].f();
    }
    S.snaps = L->per_slot[s * 11 + 10].f();
  }
  if ((rc = ealloc(L, L->d_shots, L->sh.size() * sizeof(EShot)))) return fail(rc);
  if (hipMemcpy(L->d_shots.p, L->sh.data(), L->sh.size() * sizeof(EShot), hipMemcpyHostToDevice))
    return fail((set_error("descriptor upload failed"), kErrHip));
  if (L->graph_steps > 0) {
    for (int g0 = 0; g0 < L->total_steps; g0 += L->graph_steps) {
      const int g1 = std::min(L->total_steps, g0 + L->graph_steps);
      if (hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal))
        return fail((set_error("graph capture failed"), kErrHip));
      e_issue(L, g0, g1, L->stream);
      hipGraph_t gr;
      hipGraphExec_t ex;
      if (hipStreamEndCapture(L->stream, &gr) ||
          hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0))
        return fail((set_error("graph instantiation failed"), kErrHip));
      L->graph_defs.push_back(gr);
      L->graphs.push_back(ex);
    }
  }
  *out = L;
  return kOk;
}

int ela_lane_launch(ElaLane* L, int mode) {
  hipStream_t st = L->stream;
  if (mode == TSW_RUN_BATCHED) {
    if (!L->graphs.empty()) {
      for (auto ex : L->graphs) TSW_HIP(hipGraphLaunch(ex, st));
    } else {
      e_issue(L, 0, L->total_steps, st);
    }
  } else {
    DevBuf one;
    int lst[3];
    TSW_TRY(one.ensure(sizeof lst));
    for (int i = 0; i < L->nshot; ++i) {
      lst[0] = i;
      lst[1] = 0;
      lst[2] = L->sh[i].tiles;
      TSW_HIP(hipMemcpyAsync(one.p, lst, sizeof lst, hipMemcpyHostToDevice, st));
      TSW_HIP(hipStreamSynchronize(st));
      EEpoch e{};
      e.nshot = 1;
      e.tiles = L->sh[i].tiles;
      e.d_list = one.i();
      for (int m = -1; m < 2 * L->sh[i].nt; ++m) e_launch(L, e, L->sh[i].start + m, st);
      TSW_HIP(hipStreamSynchronize(st));
    }
  }
  TSW_TRY(e_finish(L, st));
  TSW_HIP(hipGetLastError());
  return kOk;
}

int ela_lane_sync(ElaLane* L) {
  TSW_HIP(hipStreamSynchronize(L->stream));
  return kOk;
}

int ela_lane_time(ElaLane* L, double* ms, int* launches) {
  hipStream_t st = L->stream;
  const int chunk = 512;
  std::vector<hipEvent_t> ev(2 * chunk);
  for (auto& e : ev) TSW_HIP(hipEventCreate(&e));
  double total = 0.0;
  int count = 0, used = 0, inrun = 0, rc = kOk;
  auto drain = [&]() -> int {
    TSW_HIP(hipEventSynchronize(ev[2 * used - 1]));
    for (int j = 0; j < used; ++j) {
      float t = 0.0f;
      TSW_HIP(hipEventElapsedTime(&t, ev[2 * j], ev[2 * j + 1]));
      total += t;
    }
    used = 0;
    return kOk;
  };
  for (int g = 0; g < L->total_steps && rc == kOk; ++g) {
    const EEpoch* e = e_epoch_at(L, g);
    if (!e || e->tiles == 0) continue;
    for (int pass = 0; pass < 2 && rc == kOk; ++pass) {
      if (inrun == 0 && hipEventRecord(ev[2 * used], st) != hipSuccess) rc = kErrHip;
      hipLaunchKernelGGL(k_ela_batch, dim3(e->tiles), dim3(kEBX, kEBZ), 0, st, L->c,
                         (const EShot*)L->d_shots.p, (const int*)e->d_list, e->nshot, e->tiles,
                         g, pass);
      ++count;
      if (++inrun == kTimedRun) {
        if (hipEventRecord(ev[2 * used + 1], st) != hipSuccess) rc = kErrHip;
        inrun = 0;
        if (++used == chunk && rc == kOk) rc = drain();
      }
    }
  }
  if (rc == kOk && inrun) {
    if (hipEventRecord(ev[2 * used + 1], st) != hipSuccess) rc = kErrHip;
    ++used;
  }
  if (rc == kOk && used) rc = drain();
  if (rc == kOk) rc = e_finish(L, st);
  (void)hipStreamSynchronize(st);
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc != kOk) {
    set_error("elastic kernel timing run failed");
    return rc;
  }
  *ms = total;
  *launches = count;
  return kOk;
}

int ela_lane_result(ElaLane* L, int i, float* image_out, float* illum_out) {
  const size_t nb = (size_t)L->sh[i].nx * L->sh[i].nz * sizeof(float);
  if (image_out) TSW_HIP(hipMemcpy(image_out, L->sh[i].image, nb, hipMemcpyDeviceToHost));
  if (illum_out) TSW_HIP(hipMemcpy(illum_out, L->sh[i].illum, nb, hipMemcpyDeviceToHost));
  return kOk;
}

const float* ela_lane_image(ElaLane* L, int i) { return L->sh[i].image; }
const float* ela_lane_illum(ElaLane* L, int i) { return L->sh[i].illum; }
long long ela_lane_cell_updates(ElaLane* L) { return L->cell_updates; }
long long ela_lane_device_bytes(ElaLane* L) { return (long long)L->device_bytes; }
int ela_lane_steps(ElaLane* L) { return L->total_steps; }
int ela_lane_graphs(ElaLane* L) { return (int)L->graphs.size(); }
void ela_lane_destroy(ElaLane* L) { delete L; }

}  // namespace tsw
