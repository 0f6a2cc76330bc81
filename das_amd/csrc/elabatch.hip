// Elastic lane of a batch: many elastic shots of uneven size and length,
// resident in HBM, advanced together.  Each global time step is one launch
// (velocity and stress passes fused through LDS) covering the tiles of every
// active elastic shot; the steps are captured into hipGraphs and replayed on
// the lane's own stream, so an elastic lane overlaps with the acoustic lane of
// the same batch.  The per-cell arithmetic is the device functions of the
// single-shot elastic path, so per-shot results are bit-identical to
// tsw_migrate_shot_elastic.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "elabatch.h"
#include "elastic.h"

namespace tsw {

// One launch advances every active shot one full time step: the velocity
// pass, then the stress pass, with the per-cell arithmetic of the two-pass
// single-shot kernels.  A workgroup owns a core of kFW x kFH cells.  It stages
// the old stresses of the core plus 4 cells on each side in LDS, computes the
// new velocities of the core plus 2 cells on each side into LDS (the halo
// velocities are recomputed by every tile that needs them; only core cells
// are written out), then the new stresses of the core.  A step reads one
// buffer set of a shot and writes the other, so no tile reads what another
// tile of the same launch writes.
constexpr int kFW = 56;            // core columns
constexpr int kFC = kFW + 8;       // LDS columns: one wave per row
constexpr int kFH = 8;             // core rows
constexpr int kFWaves = 4;
constexpr int kFSR = kFH + 8;      // stress rows staged
constexpr int kFVR = kFH + 4;      // velocity rows computed
constexpr int kFSQ = kFSR / kFWaves;
constexpr int kFVQ = kFVR / kFWaves;
constexpr int kFCQ = kFH / kFWaves;
static_assert(kFC == 64, "one wave per LDS row");
static_assert(kFSR % kFWaves == 0 && kFVR % kFWaves == 0 && kFH % kFWaves == 0,
              "rows split evenly over the waves");

__device__ __forceinline__ int xcd_tile_e(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, l = b >> 3;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

__global__ __launch_bounds__(kFC* kFWaves) void k_ela_fused(EConst c, const EShot* shots,
                                                           const int* list, int nact, int ntiles,
                                                           int g) {
  __shared__ float s_xx[kFSR][kFC], s_zz[kFSR][kFC], s_xz[kFSR][kFC];
  __shared__ float s_vx[kFVR][kFC], s_vz[kFVR][kFC];
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
  // a wave is one threadIdx.y: its row index is wave-uniform
  const int lane = threadIdx.x, wv = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int nx = S.nx, nz = S.nz;
  const int x0 = (t % S.tiles_x) * kFW, z0 = (t / S.tiles_x) * kFH;
  const int gx = x0 - 4 + lane;  // this lane's column
  const bool xin = gx >= 0 && gx < nx;
  const bool core_x = lane >= 4 && lane < 4 + kFW && gx < nx;
  const int m = g - S.start;
  if (m < 0) {  // initialisation step of the shot: the first read set, illum, image
    if (core_x)
      for (int r = wv; r < kFH && z0 + r < nz; r += kFWaves) {
        const int i = (z0 + r) * nx + gx;
        for (int j = 0; j < 5; ++j) {
          S.f[j][i] = 0.0f;
          S.fb[j][i] = 0.0f;
        }
        S.illum[i] = 0.0f;
        S.image[i] = 0.0f;
      }
    return;
  }
  const bool fwd = m < S.nt;
  const int ms = fwd ? m : m - S.nt;  // step within this propagation
  const bool odd = ms & 1;
  float* const* cur = fwd ? (odd ? S.f2 : S.f) : (odd ? S.fb2 : S.fb);
  float* const* nxt = fwd ? (odd ? S.f : S.f2) : (odd ? S.fb : S.fb2);
  const float* const* model = fwd ? S.mf : S.mb;
  const int it = fwd ? m : 2 * S.nt - 1 - m;
  const long long n = (long long)nx * nz;
  // imaging steps: forward vz snapshot + illumination, backward correlation
  // with the snapshot of the same step (cursor == it/k)
  gfloat* snap = nullptr;
  const gfloat* isnap = nullptr;
  if (it % S.k == 0) {
    if (fwd) snap = as_global(S.snaps) + (long long)(it / S.k) * n;
    else isnap = as_global((const float*)S.snaps) + (long long)(it / S.k) * n;
  }
  const int zlo = c.fs ? 1 : 2;  // first interior row
  // Every global read of the step is issued here, before the first barrier,
  // from clamped addresses (so no load sits behind a branch and waits alone);
  // cells outside the grid are then set to 0.
  const int gxc = min(max(gx, 0), nx - 1);
  {
    const gfloat* oxx = as_global((const float*)cur[2]);
    const gfloat* ozz = as_global((const float*)cur[3]);
    const gfloat* oxz = as_global((const float*)cur[4]);
    float a[kFSQ], bb[kFSQ], d[kFSQ];
#pragma unroll
    for (int q = 0; q < kFSQ; ++q) {
      const int gz = z0 - 4 + wv + q * kFWaves;
      const int i = min(max(gz, 0), nz - 1) * nx + gxc;
      a[q] = oxx[i];
      bb[q] = ozz[i];
      d[q] = oxz[i];
    }
#pragma unroll
    for (int q = 0; q < kFSQ; ++q) {
      const int r = wv + q * kFWaves, gz = z0 - 4 + r;
      const bool in = xin && gz >= 0 && gz < nz;
      s_xx[r][lane] = in ? a[q] : 0.0f;
      s_zz[r][lane] = in ? bb[q] : 0.0f;
      s_xz[r][lane] = in ? d[q] : 0.0f;
    }
  }
  // velocity rows: old velocities, buoyancies, sponge factors and the imaging
  // operands; core rows: Lame parameters and sponge factors
  float ovx[kFVQ], ovz[kFVQ], obx[kFVQ], obz[kFVQ], gv[kFVQ], side[kFVQ], isn[kFVQ];
  float lam[kFCQ], mu[kFCQ], gc[kFCQ];
  const gfloat* tab = as_global(c.table);
  {
    const gfloat* vx = as_global((const float*)cur[0]);
    const gfloat* vz = as_global((const float*)cur[1]);
    const gfloat* bx = as_global(model[2]);
    const gfloat* bz = as_global(model[3]);
    // imaging steps read the illumination (forward) or the image (backward)
    const gfloat* sp = as_global((const float*)(snap != nullptr ? S.illum : S.image));
    const bool imaging = snap != nullptr || isnap != nullptr;
#pragma unroll
    for (int q = 0; q < kFVQ; ++q) {
      const int gzc = min(max(z0 - 2 + wv + q * kFWaves, 0), nz - 1);
      const int i = gzc * nx + gxc;
      ovx[q] = vx[i];
      ovz[q] = vz[i];
      obx[q] = bx[i];
      obz[q] = bz[i];
      const int dep = sponge_depth(gxc, gzc, nx, nz, c.n_sponge, c.fs);
      const float t = tab[dep];
      gv[q] = dep > 0 ? t : 1.0f;
      side[q] = isn[q] = 0.0f;
      if (imaging) side[q] = sp[i];
      if (isnap != nullptr) isn[q] = __builtin_nontemporal_load(isnap + i);  // read once
    }
    const gfloat* L = as_global(model[0]);
    const gfloat* M = as_global(model[1]);
#pragma unroll
    for (int q = 0; q < kFCQ; ++q) {
      const int gzc = min(z0 + wv + q * kFWaves, nz - 1);
      const int i = gzc * nx + gxc;
      lam[q] = L[i];
      mu[q] = M[i];
      const int dep = sponge_depth(gxc, gzc, nx, nz, c.n_sponge, c.fs);
      const float t = tab[dep];
      gc[q] = dep > 0 ? t : 1.0f;
    }
  }
  // injection operands: the source sample (forward); on the receiver row, the
  // lane's receiver range and its first sample (back-propagation)
  const float sval = fwd ? as_global(S.src)[it] : 0.0f;
  const int rvr = S.riz - (z0 - 2);  // velocity row of the receiver row
  const bool rwave = !fwd && rvr >= 0 && rvr < kFVR && rvr % kFWaves == wv;
  int r0 = 0, r1 = 0;
  float rs0 = 0.0f;
  const gfloat* rrow = as_global(S.data) + (long long)it * S.nrec;
  if (rwave) {
    r0 = as_global(S.roff)[gxc];
    r1 = as_global(S.roff)[gxc + 1];
    if (r0 < r1) rs0 = rrow[as_global(S.ridx)[r0]];
  }
  __syncthreads();
  // velocity pass on the core plus 2 cells each side (wave2d.c:123-155), with
  // the injection and, on core cells, the imaging side work
  {
    gfloat* nvx = as_global(nxt[0]);
    gfloat* nvz = as_global(nxt[1]);
#pragma unroll
    for (int q = 0; q < kFVQ; ++q) {
      const int vr = wv + q * kFWaves, lz = vr + 2, gz = z0 - 2 + vr;
      float vxn = 0.0f, vzn = 0.0f;
      if (lane >= 2 && lane < kFC - 2 && xin && gz >= 0 && gz < nz) {
        const int i = gz * nx + gx;
        const float dtb = c.dt * obz[q];
        float vzc = ovz[q];
        if (fwd) {
          if (i == S.src_cell) vzc = vzc + dtb * sval;
        } else if (vr == rvr && r0 < r1) {
          vzc = vzc + dtb * rs0;
          for (int r = r0 + 1; r < r1; ++r) vzc = vzc + dtb * rrow[as_global(S.ridx)[r]];
        }
        if (lane >= 4 && lane < 4 + kFW && vr >= 2 && vr < 2 + kFH) {
          if (snap != nullptr) {
            __builtin_nontemporal_store(vzc, snap + i);  // read back steps later
            as_global(S.illum)[i] = side[q] + vzc * vzc;
          }
          if (isnap != nullptr) as_global(S.image)[i] = side[q] + (isn[q] * vzc) * c.w;
        }
        if (!(gx < 2 || gx >= nx - 2 || gz < zlo || gz >= nz - 2)) {
          ElaVelIn v;
          v.upd = true;
          v.vz0 = ovz[q];
          v.vxi = ovx[q];
          v.bxi = obx[q];
          v.bzi = obz[q];
          v.sxx_p1 = s_xx[lz][lane + 1];
          v.sxx_0 = s_xx[lz][lane];
          v.sxx_p2 = s_xx[lz][lane + 2];
          v.sxx_m1 = s_xx[lz][lane - 1];
          v.sxz_0 = s_xz[lz][lane];
          v.sxz_mz = s_xz[lz - 1][lane];
          v.sxz_pz = s_xz[lz + 1][lane];
          v.sxz_m2z = s_xz[lz - 2][lane];
          v.sxz_m1 = s_xz[lz][lane - 1];
          v.sxz_p1 = s_xz[lz][lane + 1];
          v.sxz_m2 = s_xz[lz][lane - 2];
          v.szz_pz = s_zz[lz + 1][lane];
          v.szz_0 = s_zz[lz][lane];
          v.szz_p2z = s_zz[lz + 2][lane];
          v.szz_mz = s_zz[lz - 1][lane];
          v.g = gv[q];
          ela_vel_update(v, vzc, c.rdx, c.rdz, c.dt, vxn, vzn);
        } else {
          vxn = ovx[q];
          vzn = vzc;
        }
        if (lane >= 4 && lane < 4 + kFW && vr >= 2 && vr < 2 + kFH) {
          nvx[i] = vxn;
          nvz[i] = vzn;
        }
      }
      s_vx[vr][lane] = vxn;
      s_vz[vr][lane] = vzn;
    }
  }
  __syncthreads();
  // stress pass on the core (wave2d.c:158-195) and the free surface on row 0
  // (wave2d.c:198-211); cells the pass does not update keep their values
  {
    gfloat* nxx = as_global(nxt[2]);
    gfloat* nzz = as_global(nxt[3]);
    gfloat* nxz = as_global(nxt[4]);
#pragma unroll
    for (int q = 0; q < kFCQ; ++q) {
      const int cr = wv + q * kFWaves, gz = z0 + cr, lz = cr + 4, vr = cr + 2;
      if (!core_x || gz >= nz) continue;
      const int i = gz * nx + gx;
      float sxx = s_xx[lz][lane], szz = s_zz[lz][lane], sxz = s_xz[lz][lane];
      if (gz == 0) {
        if (c.fs) {
          szz = 0.0f;
          sxz = 0.0f;
          if (gx >= 2 && gx < nx - 2)
            sxx = ela_free_surface_sxx(lam[q], mu[q], s_vx[vr][lane - 2], s_vx[vr][lane - 1],
                                       s_vx[vr][lane], s_vx[vr][lane + 1], c.rdx);
        }
      } else if (!(gx < 2 || gx >= nx - 2 || gz < zlo || gz >= nz - 2)) {
        ElaStrIn v;
        v.upd = true;
        v.vx_0 = s_vx[vr][lane];
        v.vx_m1 = s_vx[vr][lane - 1];
        v.vx_p1 = s_vx[vr][lane + 1];
        v.vx_m2 = s_vx[vr][lane - 2];
        v.vx_pz = s_vx[vr + 1][lane];
        v.vx_p2z = s_vx[vr + 2][lane];
        v.vx_mz = s_vx[vr - 1][lane];
        v.vz_0 = s_vz[vr][lane];
        v.vz_mz = s_vz[vr - 1][lane];
        v.vz_pz = s_vz[vr + 1][lane];
        v.vz_m2z = s_vz[vr - 2][lane];
        v.vz_p1 = s_vz[vr][lane + 1];
        v.vz_p2 = s_vz[vr][lane + 2];
        v.vz_m1 = s_vz[vr][lane - 1];
        v.l = lam[q];
        v.m = mu[q];
        v.sxx = sxx;
        v.szz = szz;
        v.sxz = sxz;
        v.g = gc[q];
        ela_str_update(v, c.rdx, c.rdz, c.dt, sxx, szz, sxz);
      }
      nxx[i] = sxx;
      nzz[i] = szz;
      nxz[i] = sxz;
    }
  }
}

void ela_batch_step(const EConst& c, const EShot* shots, const int* list, int nact, int ntiles,
                    int g, hipStream_t st) {
  hipLaunchKernelGGL(k_ela_fused, dim3(ntiles), dim3(kFC, kFWaves), 0, st, c, shots, list, nact,
                     ntiles, g);
}

int ela_tiles(int nx, int nz, int* tiles_x) {
  *tiles_x = (nx + kFW - 1) / kFW;
  return *tiles_x * ((nz + kFH - 1) / kFH);
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
  ela_batch_step(L->c, (const EShot*)L->d_shots.p, (const int*)e.d_list, e.nshot, e.tiles, g, st);
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
    S.tiles = ela_tiles(d.nx, d.nz, &S.tiles_x);
    L->cell_updates += 2LL * d.nt * d.nx * d.nz;
  }
  if ((rc = e_build(L))) return fail(rc);
  // per slot: two sets each of the 5 forward and the 5 back-propagation fields
  // + vz snapshot arena, sized for the slot's largest shot
  std::vector<long long> cells(L->slots, 0), snap(L->slots, 0);
  for (int i = 0; i < L->nshot; ++i) {
    const long long n = (long long)L->sh[i].nx * L->sh[i].nz;
    const int s = L->slot_of[i];
    cells[s] = std::max(cells[s], n);
    snap[s] = std::max(snap[s], (long long)((L->sh[i].nt - 1) / k + 1) * n);
  }
  L->per_slot.resize((size_t)L->slots * 21);
  for (int s = 0; s < L->slots; ++s) {
    for (int j = 0; j < 20; ++j)
      if ((rc = ealloc(L, L->per_slot[s * 21 + j], cells[s] * sizeof(float)))) return fail(rc);
    if ((rc = ealloc(L, L->per_slot[s * 21 + 20], snap[s] * sizeof(float)))) return fail(rc);
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
      S.f[j] = L->per_slot[s * 21 + j].f();
      S.fb[j] = L->per_slot[s * 21 + 5 + j].f();
      S.f2[j] = L->per_slot[s * 21 + 10 + j].f();
      S.fb2[j] = L->per_slot[s * 21 + 15 + j].f();
    }
    S.snaps = L->per_slot[s * 21 + 20].f();
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
    if (inrun == 0 && hipEventRecord(ev[2 * used], st) != hipSuccess) rc = kErrHip;
    e_launch(L, *e, g, st);
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
