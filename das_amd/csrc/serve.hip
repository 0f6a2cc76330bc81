// Serving loop for virtual-shot migration requests (continuous batching).
//
// Requests (acoustic or elastic shots of any grid size and length) are
// submitted at any time and queued in arrival order.  A worker thread owns the
// device: at every tick it admits queued requests while a slot is free and the
// request's HBM footprint (models, wavefields, snapshot arena, image) fits the
// budget, then advances every running request together: `tick_steps` time
// steps, as temporally blocked acoustic launches on one stream and per-step
// elastic launches on a second stream, so the two kinds overlap on the GPU.
// A request that finishes releases its wavefields and snapshots at once and
// keeps its image and illumination until the caller fetches them.
//
// Every request runs through the same device functions as the single-shot
// paths, with no cross-request arithmetic, so its result is bit-identical
// alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "arena.h"
#include "elabatch.h"
#include "post.h"
#include "tb.h"
#include "tswrtm.h"

namespace tsw {

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3, kCancelled = 4 };

struct Req {
  long long id = 0;
  int kind = TSW_SHOT_ACOUSTIC;
  int nx = 0, nz = 0, nt = 0, six = 0, siz = 0, riz = 0, nrec = 0;
  std::vector<float> model[8];  // acoustic: [0] fwd vel, [1] bwd vel; elastic: 4 fwd + 4 bwd
  std::vector<float> src, data;
  std::vector<int> rix;
  int state = kQueued;
  bool cancel = false;
  double t_submit = 0, t_admit = 0, t_done = 0;
  long long need = 0;     // device bytes while running
  long long keep = 0;     // device bytes kept after completion (image + illumination)
  int slot = -1;
  bool in_queue = false;  // listed in Server::queue
  bool resident = false;  // models are the window [z0, z0+nz) x [x0, x0+nx) of the server's
  int x0 = 0, z0 = 0;     //   resident models, copied on the device at admission
  int cursor = 0;         // acoustic: next block; elastic: unused (global step based)
  int e_start = 0;        // elastic: global elastic step of the first forward step
  char* heavy = nullptr;       // arena block: state, models, data, snapshots
  size_t heavy_bytes = 0;      //   (released at completion)
  char* keepb = nullptr;       // arena block: image + illumination (released when fetched)
  float* image = nullptr;
  float* illum = nullptr;
  std::string err;
  long long cells() const { return (long long)nx * nz; }
};

struct Server {
  // constants
  double dt = 0;
  int k = 2, kernel = 0, T = 1, W = 0, H = 0, tick_steps = 100, max_active = 1;
  long long budget = 0;
  BatchConst bc{};
  EConst ec{};
  hipStream_t sa = nullptr, se = nullptr;
  DevBuf d_table, d_tbshots, d_eshots, d_tbent, d_tbtile, d_elist;
  Arena arena;
  void* arena_mem = nullptr;
  // resident models shared by requests (0/1 acoustic fwd/bwd velocity,
  // 2..5 / 6..9 elastic fwd / bwd lambda, mu, buoyancy_x, buoyancy_z)
  DevBuf res[10];
  int res_nx = 0, res_nz = 0;
  bool res_acoustic = false, res_elastic = false;
  std::vector<TbShot> tbs;
  std::vector<EShot> es;
  // state
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req*> queue;
  std::map<long long, std::unique_ptr<Req>> reqs;
  std::vector<Req*> slots;
  long long mem_used = 0, next_id = 1;
  int g_e = 0;
  bool stop = false;
  std::thread worker;
  std::chrono::steady_clock::time_point t0;
  // statistics
  long long completed = 0, failed = 0, cancelled = 0, cell_updates = 0;
  double busy_s = 0;
  std::string fatal;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

// Bytes of a request's heavy block (the layout admit() carves) and of its
// kept block (image + illumination).
static long long req_heavy(const Req& r, int k) {
  auto up = [](long long b) { return (long long)Arena::up((size_t)(b ? b : 4)); };
  const long long n = r.cells();
  const long long nb = up(4 * n);
  const long long nsnap = (r.nt - 1) / k + 1;
  long long b = up(4LL * r.nt) + up(4LL * r.nt * r.nrec) + up(4LL * (r.nx + 1)) +
                up(4LL * std::max(r.nrec, 1)) + up(4 * nsnap * n);
  if (r.kind == TSW_SHOT_ACOUSTIC) b += 2 * nb + nb /* staging */ + 6 * nb;
  else b += 8 * up(4 * n) + 20 * nb;  // models, two sets of 5 + 5 fields
  return b;
}
static long long req_keep(const Req& r) { return 2 * (long long)Arena::up((size_t)(4 * r.cells())); }

static int alloc_up(DevBuf& d, const void* h, size_t bytes, hipStream_t st) {
  TSW_TRY(d.ensure(bytes ? bytes : 4));
  if (h && bytes) TSW_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, st));
  return kOk;
}

// Uploads a request's inputs into its arena blocks (already reserved) and
// writes its descriptor into slot `s`.  Called by the worker without the lock.
static int admit(Server* S, Req* r, int s) {
  TSW_TRY(fail_injected());
  hipStream_t st = r->kind == TSW_SHOT_ACOUSTIC ? S->sa : S->se;
  const long long n = r->cells();
  const size_t nb = Arena::up((size_t)n * sizeof(float));
  const long long nsnap = (r->nt - 1) / S->k + 1;
  RecvMap m;
  TSW_TRY(m.build(r->rix.data(), r->nrec, r->nx));
  char* cur = r->heavy;
  auto take = [&](size_t bytes) {
    char* p = cur;
    cur += Arena::up(bytes);
    return p;
  };
  auto up = [&](const void* h, size_t bytes) -> char* {
    char* p = take(bytes ? bytes : 4);
    if (bytes) (void)hipMemcpyAsync(p, h, bytes, hipMemcpyHostToDevice, st);
    return p;
  };
  float* src = (float*)up(r->src.data(), r->src.size() * sizeof(float));
  float* data = (float*)up(r->data.data(), r->data.size() * sizeof(float));
  int* roff = (int*)up(m.off.data(), m.off.size() * sizeof(int));
  int* ridx = (int*)up(m.idx.data(), m.idx.size() * sizeof(int));
  r->image = (float*)r->keepb;
  r->illum = (float*)(r->keepb + nb);
  if (r->kind == TSW_SHOT_ACOUSTIC) {
    TbShot P;
    std::memset(&P, 0, sizeof P);
    const float dtf = (float)S->dt;
    float* coef[2];
    for (int j = 0; j < 2; ++j) {
      coef[j] = (float*)take(nb);
      if (r->resident) {
        float* vel = (float*)take(nb);
        TSW_HIP(hipMemcpy2DAsync(vel, (size_t)r->nx * sizeof(float),
                                 S->res[j].f() + (size_t)r->z0 * S->res_nx + r->x0,
                                 (size_t)S->res_nx * sizeof(float), (size_t)r->nx * sizeof(float),
                                 r->nz, hipMemcpyDeviceToDevice, st));
        launch_coef(vel, coef[j], n, dtf * dtf, st);
        cur = (char*)vel;
      } else {
        float* vel = (float*)up(r->model[j].data(), (size_t)n * sizeof(float));
        launch_coef(vel, coef[j], n, dtf * dtf, st);  // `vel` space is reused below
        cur = (char*)vel;
      }
    }
    for (int j = 0; j < 3; ++j) {
      P.set[0][j] = (float*)take(nb);
      P.set[1][j] = (float*)take(nb);
    }
    P.snaps = (float*)take((size_t)nsnap * n * sizeof(float));
    P.nx = r->nx;
    P.nz = r->nz;
    P.nt = r->nt;
    P.k = S->k;
    P.src_cell = r->siz * r->nx + r->six;
    P.riz = r->riz;
    P.nrec = r->nrec;
    P.coef_f = coef[0];
    P.coef_b = coef[1];
    P.src = src;
    P.data = data;
    P.roff = roff;
    P.ridx = ridx;
    P.illum = r->illum;
    P.image = r->image;
    S->tbs[s] = P;
    TSW_HIP(hipMemcpyAsync((TbShot*)S->d_tbshots.p + s, &S->tbs[s], sizeof(TbShot),
                           hipMemcpyHostToDevice, st));
    r->cursor = 0;
  } else {
    EShot E;
    std::memset(&E, 0, sizeof E);
    for (int j = 0; j < 4; ++j) {
      if (r->resident) {
        for (int b = 0; b < 2; ++b) {
          float* dst = (float*)take((size_t)n * sizeof(float));
          TSW_HIP(hipMemcpy2DAsync(dst, (size_t)r->nx * sizeof(float),
                                   S->res[2 + 4 * b + j].f() + (size_t)r->z0 * S->res_nx + r->x0,
                                   (size_t)S->res_nx * sizeof(float),
                                   (size_t)r->nx * sizeof(float), r->nz, hipMemcpyDeviceToDevice,
                                   st));
          (b == 0 ? E.mf : E.mb)[j] = dst;
        }
      } else {
        E.mf[j] = (float*)up(r->model[j].data(), (size_t)n * sizeof(float));
        E.mb[j] = (float*)up(r->model[4 + j].data(), (size_t)n * sizeof(float));
      }
    }
    for (int j = 0; j < 5; ++j) {
      E.f[j] = (float*)take(nb);
      E.fb[j] = (float*)take(nb);
      E.f2[j] = (float*)take(nb);
      E.fb2[j] = (float*)take(nb);
    }
    E.snaps = (float*)take((size_t)nsnap * n * sizeof(float));
    E.nx = r->nx;
    E.nz = r->nz;
    E.nt = r->nt;
    E.k = S->k;
    E.start = S->g_e + 1;  // initialisation at the tick's first elastic step
    E.src_cell = r->siz * r->nx + r->six;
    E.riz = r->riz;
    E.nrec = r->nrec;
    E.src = src;
    E.data = data;
    E.roff = roff;
    E.ridx = ridx;
    E.illum = r->illum;
    E.image = r->image;
    E.tiles = ela_tiles(r->nx, r->nz, &E.tiles_x);
    S->es[s] = E;
    TSW_HIP(hipMemcpyAsync((EShot*)S->d_eshots.p + s, &S->es[s], sizeof(EShot),
                           hipMemcpyHostToDevice, st));
    r->e_start = E.start;
  }
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipStreamSynchronize(st));
  if ((size_t)(cur - r->heavy) > r->heavy_bytes) {
    set_error("internal: request layout exceeds its reservation");
    return kErrState;
  }
  // the host copies are no longer needed once the request lives in HBM
  for (auto& v : r->model) std::vector<float>().swap(v);
  std::vector<float>().swap(r->src);
  std::vector<float>().swap(r->data);
  return kOk;
}

static void release(Server* S, Req* r) {
  S->arena.release(r->heavy, r->heavy_bytes);
  r->heavy = nullptr;
  if (r->slot >= 0) S->slots[r->slot] = nullptr;
  r->slot = -1;
}
static void release_keep(Server* S, Req* r) {
  S->arena.release(r->keepb, (size_t)r->keep);
  r->keepb = nullptr;
  r->image = r->illum = nullptr;
}

// One tick: tick_steps time steps of every running request.  Runs without
// the server lock (only the worker touches slots and cursors); requests that
// finished are returned in `done`, work counters in `cu` / `busy`.
static int tick(Server* S, std::vector<Req*>& done, long long& cu, double& busy) {
  cu = 0;
  busy = 0;
  std::vector<Req*> act_a, act_e;
  for (Req* r : S->slots)
    if (r) (r->kind == TSW_SHOT_ACOUSTIC ? act_a : act_e).push_back(r);
  // acoustic: blocked launches, each running request advances one block each
  const int nlaunch = std::max(1, S->tick_steps / S->T);
  std::vector<TbEntry> ents;
  std::vector<TbTile> tab;              // per launch: (entry, tile) per workgroup
  std::vector<std::pair<int, int>> lr;  // (entry offset, table offset) per launch
  std::vector<int> ltiles;
  for (int L = 0; L < nlaunch && !act_a.empty(); ++L) {
    const int eo = (int)ents.size(), to = (int)tab.size();
    std::vector<TbTile> lin;
    for (Req* r : act_a) {
      const int nb = (r->nt + S->T - 1) / S->T;
      if (r->cursor >= 2 * nb) continue;
      const int b = r->cursor < nb ? r->cursor : r->cursor - nb;
      const int ph = r->cursor < nb ? 0 : 1;
      TbEntry e{};
      e.shot = r->slot;
      e.phase = ph;
      e.it0 = ph == 0 ? b * S->T : r->nt - 1 - b * S->T;
      e.nsub = ph == 0 ? std::min(S->T, r->nt - e.it0) : std::min(S->T, e.it0 + 1);
      e.flags = (b == 0 ? kTbFirst : 0) | (b == nb - 1 ? kTbLast : 0);
      e.rd = b & 1;
      e.tiles_x = (r->nx + S->W - 1) / S->W;
      tb_schedule(e, r->nt, S->k);
      const int tiles = e.tiles_x * ((r->nz + S->H - 1) / S->H);
      const int ei = (int)ents.size() - eo;
      for (int t = 0; t < tiles; ++t) lin.push_back({ei, (t % e.tiles_x) | (t / e.tiles_x) << 16});
      ents.push_back(e);
      r->cursor++;
      cu += (long long)e.nsub * r->cells();
    }
    if (lin.empty()) break;
    const int nti = (int)lin.size();
    for (int b = 0; b < nti; ++b) tab.push_back(lin[xcd_tile_tb(b, nti)]);
    lr.push_back({eo, to});
    ltiles.push_back(nti);
  }
  // elastic: per-step lists of running requests (ids are slots)
  std::vector<int> elist;
  std::vector<std::pair<int, int>> el;  // (offset, nact) per step
  std::vector<int> etiles;
  for (int q = 0; q < S->tick_steps && !act_e.empty(); ++q) {
    const int g = S->g_e + q;
    std::vector<int> ids;
    std::vector<int> pre{0};
    for (Req* r : act_e) {
      const int m = g - r->e_start;
      if (m < -1 || m >= 2 * r->nt) continue;
      ids.push_back(r->slot);
      pre.push_back(pre.back() + S->es[r->slot].tiles);
      if (m >= 0) cu += r->cells();
    }
    el.push_back({(int)elist.size(), (int)ids.size()});
    etiles.push_back(pre.back());
    elist.insert(elist.end(), ids.begin(), ids.end());
    elist.insert(elist.end(), pre.begin(), pre.end());
  }
  TSW_TRY(S->d_tbent.ensure(std::max<size_t>(ents.size(), 1) * sizeof(TbEntry)));
  TSW_TRY(S->d_tbtile.ensure(std::max<size_t>(tab.size(), 1) * sizeof(TbTile)));
  TSW_TRY(S->d_elist.ensure(std::max<size_t>(elist.size(), 1) * sizeof(int)));
  if (!ents.empty()) {
    TSW_HIP(hipMemcpyAsync(S->d_tbent.p, ents.data(), ents.size() * sizeof(TbEntry),
                           hipMemcpyHostToDevice, S->sa));
    TSW_HIP(hipMemcpyAsync(S->d_tbtile.p, tab.data(), tab.size() * sizeof(TbTile),
                           hipMemcpyHostToDevice, S->sa));
  }
  if (!elist.empty())
    TSW_HIP(hipMemcpyAsync(S->d_elist.p, elist.data(), elist.size() * sizeof(int),
                           hipMemcpyHostToDevice, S->se));
  const double t_start = S->now();
  for (size_t q = 0; q < el.size(); ++q) {
    if (el[q].second == 0) continue;
    ela_batch_step(S->ec, (const EShot*)S->d_eshots.p, S->d_elist.i() + el[q].first,
                   el[q].second, etiles[q], S->g_e + (int)q, S->se);
  }
  for (size_t L = 0; L < lr.size(); ++L)
    tb_launch(S->kernel, S->bc, (const TbShot*)S->d_tbshots.p,
              (const TbEntry*)S->d_tbent.p + lr[L].first, (const TbTile*)S->d_tbtile.p + lr[L].second,
              ltiles[L], S->sa);
  TSW_HIP(hipGetLastError());
  // completion: acoustic after its last block, elastic after step 2*nt-1
  const int g_end = S->g_e + S->tick_steps;
  for (Req* r : act_e) {
    if (g_end - r->e_start >= 2 * r->nt)
      ela_scale_launch(r->illum, r->cells(), S->ec.w, S->se);
  }
  TSW_HIP(hipStreamSynchronize(S->sa));
  TSW_HIP(hipStreamSynchronize(S->se));
  busy = S->now() - t_start;
  if (!act_e.empty()) S->g_e = g_end;
  for (Req* r : act_a) {
    const int nb = (r->nt + S->T - 1) / S->T;
    if (r->cursor >= 2 * nb) done.push_back(r);
  }
  for (Req* r : act_e)
    if (S->g_e - r->e_start >= 2 * r->nt) done.push_back(r);
  // divergence sentinel on the finished requests' images and illuminations
  for (Req* r : done) {
    const float* arr[2] = {r->image, r->illum};
    const long long len[2] = {r->cells(), r->cells()};
    if (check_finite(arr, len, 2, 2, (int)r->id, "request", S->sa) == kErrDiverged)
      r->err = tsw_last_error();
    else
      r->err.clear();
  }
  return kOk;
}

// Host copies of a request's inputs (freed at admission, cancellation or
// failure: the request no longer needs them).
static void drop_host_inputs(Req* r) {
  for (auto& v : r->model) std::vector<float>().swap(v);
  std::vector<float>().swap(r->src);
  std::vector<float>().swap(r->data);
  std::vector<int>().swap(r->rix);
}

static void fail_request(Server* S, Req* r, int state, const char* err) {
  drop_host_inputs(r);
  r->state = state;
  if (err) r->err = err;
  r->t_done = S->now();
  if (state == kFailed) S->failed++;
  else S->cancelled++;
  S->mem_used -= r->heavy_bytes + r->keep;
  release(S, r);
  release_keep(S, r);
  S->cv.notify_all();  // waiters on this request (no tick may follow to wake them)
}

static void worker_loop(Server* S) {
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // cancellations of running requests take effect between ticks
    for (Req* r : S->slots)
      if (r && r->cancel) fail_request(S, r, kCancelled, nullptr);
    // admission, in arrival order, while a slot is free and the arena has room
    std::vector<Req*> admitted;
    while (!S->queue.empty()) {
      Req* r = S->queue.front();
      if (r->cancel) {
        S->queue.pop_front();
        r->in_queue = false;
        continue;
      }
      int s = -1;
      for (int j = 0; j < S->max_active; ++j)
        if (!S->slots[j]) {
          s = j;
          break;
        }
      if (s < 0) break;
      char* h = S->arena.alloc(r->heavy_bytes);
      char* kb = h ? S->arena.alloc_high((size_t)r->keep) : nullptr;
      if (!kb) {
        S->arena.release(h, r->heavy_bytes);
        break;
      }
      S->queue.pop_front();
      r->in_queue = false;
      r->heavy = h;
      r->keepb = kb;
      r->slot = s;
      S->slots[s] = r;
      r->state = kRunning;
      r->t_admit = S->now();
      S->mem_used += r->heavy_bytes + r->keep;
      admitted.push_back(r);
    }
    if (!admitted.empty()) {
      lk.unlock();
      std::vector<int> rcs;
      std::vector<std::string> errs;
      for (Req* r : admitted) {
        rcs.push_back(admit(S, r, r->slot));
        errs.push_back(rcs.back() == kOk ? std::string() : std::string(tsw_last_error()));
      }
      lk.lock();
      for (size_t j = 0; j < admitted.size(); ++j)  // a failed admission fails that request only
        if (rcs[j] != kOk) fail_request(S, admitted[j], kFailed, errs[j].c_str());
    }
    bool any = false;
    for (Req* r : S->slots) any = any || r;
    if (!any) {
      if (S->stop) break;
      S->cv.wait_for(lk, std::chrono::milliseconds(2));
      continue;
    }
    std::vector<Req*> done;
    long long cu = 0;
    double busy = 0;
    lk.unlock();
    const int rc = tick(S, done, cu, busy);
    lk.lock();
    S->cell_updates += cu;
    S->busy_s += busy;
    const double t = S->now();
    for (Req* r : done) {
      if (!r->err.empty()) {  // diverged: failed with the sentinel's reason
        const std::string why = r->err;
        fail_request(S, r, kFailed, why.c_str());
        continue;
      }
      r->state = kDone;
      r->t_done = t;
      S->completed++;
      S->mem_used -= r->heavy_bytes;
      release(S, r);
    }
    S->cv.notify_all();
    if (rc != kOk) {
      S->fatal = tsw_last_error();
      for (Req* r : S->slots)
        if (r) fail_request(S, r, kFailed, S->fatal.c_str());
      break;
    }
  }
  S->cv.notify_all();
}

}  // namespace tsw

using namespace tsw;

extern "C" {

int tsw_server_create(double dx, double dz, double dt, int n_sponge, double sponge_alpha,
                      int free_surface_top, int k, long long budget_bytes, int max_active,
                      int kernel, int tick_steps, tsw_server_t* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(out, "out is null");
  TSW_ARG(dx > 0 && dz > 0 && dt > 0 && k >= 1 && n_sponge >= 0, "bad constants");
  TSW_ARG(budget_bytes > 0 && max_active >= 1 && tick_steps >= 1, "bad server limits");
  int T, W, H;
  TSW_ARG(tb_config(kernel, &T, &W, &H) == kOk, "kernel %d is not a blocked configuration",
          kernel);
  TSW_ARG(n_sponge <= tb_max_sponge(kernel), "kernel %d holds a sponge of at most %d cells",
          kernel, tb_max_sponge(kernel));
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  auto S = std::make_unique<Server>();
  S->dt = dt;
  S->k = k;
  S->kernel = kernel;
  S->T = T;
  S->W = W;
  S->H = H;
  S->tick_steps = tick_steps;
  S->max_active = max_active;
  S->budget = budget_bytes;
  S->t0 = std::chrono::steady_clock::now();
  TSW_HIP(hipStreamCreateWithFlags(&S->sa, hipStreamNonBlocking));
  TSW_HIP(hipStreamCreateWithFlags(&S->se, hipStreamNonBlocking));
  std::vector<float> tab;
  sponge_table(tab, n_sponge, (float)sponge_alpha);
  TSW_TRY(alloc_up(S->d_table, tab.data(), tab.size() * sizeof(float), S->sa));
  TSW_HIP(hipStreamSynchronize(S->sa));
  const float dxf = (float)dx, dzf = (float)dz;
  S->bc.fs = free_surface_top ? 1 : 0;
  S->bc.n_sponge = n_sponge;
  S->bc.rdx2 = 1.0f / (dxf * dxf);
  S->bc.rdz2 = 1.0f / (dzf * dzf);
  S->bc.inj = (float)(dt * dt);
  S->bc.w = (float)(dt * k);
  S->bc.table = S->d_table.f();
  S->ec.fs = S->bc.fs;
  S->ec.n_sponge = n_sponge;
  S->ec.rdx = 1.0f / dxf;
  S->ec.rdz = 1.0f / dzf;
  S->ec.dt = (float)dt;
  S->ec.w = (float)(dt * k);
  S->ec.table = S->d_table.f();
  S->slots.assign(max_active, nullptr);
  S->tbs.resize(max_active);
  S->es.resize(max_active);
  TSW_TRY(S->d_tbshots.ensure(max_active * sizeof(TbShot)));
  TSW_TRY(S->d_eshots.ensure(max_active * sizeof(EShot)));
  const size_t asz = Arena::up((size_t)budget_bytes);
  if (hipMalloc(&S->arena_mem, asz) != hipSuccess) {
    (void)hipGetLastError();
    set_error("cannot reserve the %lld-byte HBM budget", budget_bytes);
    return kErrMem;
  }
  S->arena.base = static_cast<char*>(S->arena_mem);
  S->arena.size = asz;
  S->arena.free_[0] = asz;
  Server* raw = S.release();
  raw->worker = std::thread(worker_loop, raw);
  *out = reinterpret_cast<tsw_server_t>(raw);
  return kOk;
}

int tsw_server_set_models(tsw_server_t h, int gnx, int gnz, const float* vel_fwd,
                          const float* vel_bwd, const float* const* ela_fwd,
                          const float* const* ela_bwd) {
  Server* S = reinterpret_cast<Server*>(h);
  TSW_ARG(S && gnx >= 5 && gnz >= 5, "bad arguments");
  TSW_ARG((vel_fwd == nullptr) == (vel_bwd == nullptr), "give both acoustic models or neither");
  TSW_ARG((ela_fwd == nullptr) == (ela_bwd == nullptr), "give both elastic models or neither");
  if (ela_fwd)
    for (int j = 0; j < 4; ++j) TSW_ARG(ela_fwd[j] && ela_bwd[j], "elastic model %d missing", j);
  std::lock_guard<std::mutex> lk(S->mu);
  bool busy = !S->queue.empty();
  for (Req* r : S->slots) busy = busy || r;
  TSW_ARG(!busy, "resident models can only be replaced while the server is idle");
  const size_t n = (size_t)gnx * gnz, nb = n * sizeof(float);
  const float* src[10] = {vel_fwd, vel_bwd, nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr, nullptr, nullptr};
  for (int j = 0; j < 4 && ela_fwd; ++j) {
    src[2 + j] = ela_fwd[j];
    src[6 + j] = ela_bwd[j];
  }
  S->res_acoustic = S->res_elastic = false;
  for (int j = 0; j < 10; ++j) {
    if (!src[j]) {
      S->res[j].release();
      continue;
    }
    TSW_TRY(S->res[j].ensure(nb));
    TSW_HIP(hipMemcpyAsync(S->res[j].p, src[j], nb, hipMemcpyHostToDevice, S->sa));
  }
  TSW_HIP(hipStreamSynchronize(S->sa));
  S->res_nx = gnx;
  S->res_nz = gnz;
  S->res_acoustic = vel_fwd != nullptr;
  S->res_elastic = ela_fwd != nullptr;
  return kOk;
}

int tsw_server_submit(tsw_server_t h, const tsw_shot_desc* d, long long* id_out) {
  Server* S = reinterpret_cast<Server*>(h);
  TSW_ARG(S && d && id_out, "null argument");
  TSW_ARG(d->kind == TSW_SHOT_ACOUSTIC || d->kind == TSW_SHOT_ELASTIC, "bad kind");
  TSW_ARG(d->nx >= 5 && d->nz >= 5 && d->nt >= 1 && d->src, "bad request grid or source");
  // fields stay under 4 GiB (32-bit byte offsets in the blocked kernel) and
  // tile indices fit the kernel's 16-bit halves
  TSW_ARG((long long)d->nx * d->nz < (1LL << 30) && (d->nx + S->W - 1) / S->W < (1 << 16) &&
              (d->nz + S->H - 1) / S->H < (1 << 15),
          "request grid %d x %d too large", d->nx, d->nz);
  TSW_ARG(d->six >= 0 && d->six < d->nx && d->siz >= 0 && d->siz < d->nz, "source outside grid");
  TSW_ARG(d->riz >= 0 && d->riz < d->nz && d->nrec >= 0, "bad receiver row");
  TSW_ARG(d->nrec == 0 || (d->rix && d->data), "receivers need rix and data");
  for (int r = 0; r < d->nrec; ++r)
    TSW_ARG(d->rix[r] >= 0 && d->rix[r] < d->nx, "receiver %d outside grid", r);
  auto r = std::make_unique<Req>();
  r->kind = d->kind;
  r->nx = d->nx;
  r->nz = d->nz;
  r->nt = d->nt;
  r->six = d->six;
  r->siz = d->siz;
  r->riz = d->riz;
  r->nrec = d->nrec;
  const size_t n = (size_t)d->nx * d->nz;
  bool none = true;
  if (d->kind == TSW_SHOT_ACOUSTIC) {
    none = !d->vel_fwd && !d->vel_bwd;
  } else {
    for (int j = 0; j < 4; ++j) none = none && !d->ela_fwd[j] && !d->ela_bwd[j];
  }
  if (none) {  // window of the resident models
    std::lock_guard<std::mutex> lk(S->mu);
    const bool have = d->kind == TSW_SHOT_ACOUSTIC ? S->res_acoustic : S->res_elastic;
    TSW_ARG(have, "request without models, and no resident %s models are set",
            d->kind == TSW_SHOT_ACOUSTIC ? "acoustic" : "elastic");
    TSW_ARG(d->x0 >= 0 && d->z0 >= 0 && d->x0 + d->nx <= S->res_nx &&
                d->z0 + d->nz <= S->res_nz,
            "window [%d, %d) x [%d, %d) outside the %d x %d resident models", d->z0,
            d->z0 + d->nz, d->x0, d->x0 + d->nx, S->res_nz, S->res_nx);
    r->resident = true;
    r->x0 = d->x0;
    r->z0 = d->z0;
  } else if (d->kind == TSW_SHOT_ACOUSTIC) {
    TSW_ARG(d->vel_fwd && d->vel_bwd, "missing velocity models");
    r->model[0].assign(d->vel_fwd, d->vel_fwd + n);
    r->model[1].assign(d->vel_bwd, d->vel_bwd + n);
  } else {
    for (int j = 0; j < 4; ++j) {
      TSW_ARG(d->ela_fwd[j] && d->ela_bwd[j], "missing elastic models");
      r->model[j].assign(d->ela_fwd[j], d->ela_fwd[j] + n);
      r->model[4 + j].assign(d->ela_bwd[j], d->ela_bwd[j] + n);
    }
  }
  r->src.assign(d->src, d->src + d->nt);
  if (d->nrec) r->data.assign(d->data, d->data + (size_t)d->nt * d->nrec);
  r->rix.assign(d->rix, d->rix + d->nrec);
  std::lock_guard<std::mutex> lk(S->mu);
  TSW_ARG(!S->stop, "server is shutting down");
  r->heavy_bytes = (size_t)req_heavy(*r, S->k);
  r->keep = req_keep(*r);
  r->need = (long long)r->heavy_bytes + r->keep;
  if (r->need > S->budget) {
    set_error("request needs %lld bytes, more than the %lld-byte budget", r->need, S->budget);
    return kErrMem;
  }
  r->id = S->next_id++;
  r->t_submit = S->now();
  *id_out = r->id;
  S->queue.push_back(r.get());
  r->in_queue = true;
  S->reqs[r->id] = std::move(r);
  S->cv.notify_all();
  return kOk;
}

int tsw_server_poll(tsw_server_t h, long long id, int* state, double* t_submit, double* t_admit,
                    double* t_done) {
  Server* S = reinterpret_cast<Server*>(h);
  TSW_ARG(S && state, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  TSW_ARG(it != S->reqs.end(), "unknown request %lld", id);
  const Req& r = *it->second;
  *state = r.state;
  if (t_submit) *t_submit = r.t_submit;
  if (t_admit) *t_admit = r.t_admit;
  if (t_done) *t_done = r.t_done;
  if (r.state == kFailed) set_error("%s", r.err.c_str());
  return kOk;
}

int tsw_server_wait(tsw_server_t h, long long id, double timeout_s) {
  Server* S = reinterpret_cast<Server*>(h);
  TSW_ARG(S, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  for (;;) {
    auto it = S->reqs.find(id);
    TSW_ARG(it != S->reqs.end(), "unknown request %lld", id);
    if (it->second->state >= kDone) return kOk;
    if (!S->fatal.empty()) {
      set_error("server stopped: %s", S->fatal.c_str());
      return kErrState;
    }
    if (std::chrono::steady_clock::now() >= until) {
      set_error("request %lld not finished after %.1f s", id, timeout_s);
      return kErrState;
    }
    S->cv.wait_until(lk, until);  // the state is checked again on every wake
  }
}

int tsw_server_result(tsw_server_t h, long long id, float* image_out, float* illum_out) {
  Server* S = reinterpret_cast<Server*>(h);
  TSW_ARG(S, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  TSW_ARG(it != S->reqs.end(), "unknown request %lld", id);
  Req& r = *it->second;
  // a failed or cancelled request is forgotten once this call reports it
  // (it holds no device memory and is in neither the queue nor a slot)
  if (r.state == kFailed) {
    set_error("request %lld failed: %s", id, r.err.c_str());
    S->reqs.erase(it);
    return kErrState;
  }
  if (r.state == kCancelled && !r.in_queue && r.slot < 0) {
    set_error("request %lld was cancelled", id);
    S->reqs.erase(it);
    return kErrState;
  }
  TSW_ARG(r.state == kDone, "request %lld is not done (state %d)", id, r.state);
  const size_t nb = (size_t)r.cells() * sizeof(float);
  if (image_out) TSW_HIP(hipMemcpy(image_out, r.image, nb, hipMemcpyDeviceToHost));
  if (illum_out) TSW_HIP(hipMemcpy(illum_out, r.illum, nb, hipMemcpyDeviceToHost));
  release_keep(S, &r);
  S->mem_used -= r.keep;
  S->reqs.erase(it);
  return kOk;
}

int tsw_server_cancel(tsw_server_t h, long long id) {
  Server* S = reinterpret_cast<Server*>(h);
  TSW_ARG(S, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  TSW_ARG(it != S->reqs.end(), "unknown request %lld", id);
  Req& r = *it->second;
  if (r.state == kQueued) {
    r.cancel = true;
    r.state = kCancelled;
    S->cancelled++;
    // out of the queue and without its host copies at once, so a cancelled
    // request holds only its record until tsw_server_result reports it
    auto q = std::find(S->queue.begin(), S->queue.end(), &r);
    if (q != S->queue.end()) S->queue.erase(q);
    r.in_queue = false;
    drop_host_inputs(&r);
  } else if (r.state == kRunning) {
    r.cancel = true;  // takes effect at the next tick boundary
  }
  S->cv.notify_all();
  return kOk;
}

int tsw_server_stats(tsw_server_t h, long long* completed, long long* failed,
                     long long* cancelled, long long* queued, long long* running,
                     long long* mem_used, long long* cell_updates, double* busy_s) {
  Server* S = reinterpret_cast<Server*>(h);
  TSW_ARG(S, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  long long q = 0, run = 0;
  for (Req* r : S->queue) q += r->cancel ? 0 : 1;
  for (Req* r : S->slots) run += r ? 1 : 0;
  if (completed) *completed = S->completed;
  if (failed) *failed = S->failed;
  if (cancelled) *cancelled = S->cancelled;
  if (queued) *queued = q;
  if (running) *running = run;
  if (mem_used) *mem_used = S->mem_used;
  if (cell_updates) *cell_updates = S->cell_updates;
  if (busy_s) *busy_s = S->busy_s;
  return kOk;
}

int tsw_server_destroy(tsw_server_t h) {
  Server* S = reinterpret_cast<Server*>(h);
  if (!S) return kOk;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    for (Req* r : S->queue) r->cancel = true;
    for (Req* r : S->slots)
      if (r) r->cancel = true;
    S->cv.notify_all();
  }
  if (S->worker.joinable()) S->worker.join();
  (void)hipStreamSynchronize(S->sa);
  (void)hipStreamSynchronize(S->se);
  (void)hipStreamDestroy(S->sa);
  (void)hipStreamDestroy(S->se);
  if (S->arena_mem) (void)hipFree(S->arena_mem);
  delete S;
  return kOk;
}

}  // extern "C"
