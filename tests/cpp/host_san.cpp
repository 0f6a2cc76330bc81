// Drives the engine's C ABI (include/tswrtm.h) for a build whose host code is
// compiled with -fsanitize=address,undefined (das_amd/build.py --sanitize):
//
//   host_san args   refused calls: every check of arguments the entry points
//                   make before touching the device, and the host-only calls.
//                   Needs no GPU.
//   host_san gpu    small runs of every group of entry points on the GPU:
//                   grids, single and multiple shots, batches with both lanes
//                   and several packings, the serving loop with cancellation
//                   and an injected failure, misfit gradients, gather
//                   construction and preprocessing.  Batch and served results
//                   must equal the single-shot results bit for bit.
//
// Host arrays are allocated at exactly the sizes the interface states, so a
// host read or write past one is reported by the sanitizer.  Prints "ok" and
// exits 0 when every call behaved; otherwise names the failing call.
#include <hip/hip_runtime.h>

#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tswrtm.h"

namespace {

int g_fail = 0;

void fail(const char* what, int line, int rc) {
  std::printf("FAIL line %d: %s -> %d (%s)\n", line, what, rc, tsw_last_error());
  std::fflush(stdout);
  g_fail = 1;
}

#define OK(call)                               \
  do {                                         \
    const int rc_ = (call);                    \
    if (rc_ != 0) fail(#call, __LINE__, rc_);  \
  } while (0)

// refused with kErrArg (-1) and a message
#define BAD(call)                                                         \
  do {                                                                    \
    const int rc_ = (call);                                               \
    if (rc_ != -1 || tsw_last_error()[0] == 0) fail(#call, __LINE__, rc_); \
  } while (0)

#define EXPECT(cond)                                        \
  do {                                                      \
    if (!(cond)) fail("expectation: " #cond, __LINE__, 0); \
  } while (0)

unsigned g_seed = 2024u;
float urand() {
  g_seed = g_seed * 1664525u + 1013904223u;
  return (float)((g_seed >> 8) * (1.0 / 16777216.0));
}

std::vector<float> uni(size_t n, float lo, float hi) {
  std::vector<float> v(n);
  for (auto& x : v) x = lo + (hi - lo) * urand();
  return v;
}

bool same(const std::vector<float>& a, const std::vector<float>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

bool finite_nonzero(const std::vector<float>& a) {
  bool nz = false;
  for (float x : a) {
    if (!std::isfinite(x)) return false;
    nz = nz || x != 0.0f;
  }
  return nz;
}

const double DX = 5.0, DT = 4e-4;
const int NSP = 8, K = 2;

// one acoustic or elastic shot with its own host arrays
struct Shot {
  int nx, nz, nt, six, siz, riz, x0, z0, kind;
  std::vector<float> vf, vb, src, data;
  std::vector<int> rix;
  std::vector<float> ef[4], eb[4];
  tsw_shot_desc desc() const {
    tsw_shot_desc d;
    std::memset(&d, 0, sizeof d);
    d.nx = nx;
    d.nz = nz;
    d.nt = nt;
    d.vel_fwd = kind == TSW_SHOT_ACOUSTIC ? vf.data() : nullptr;
    d.vel_bwd = kind == TSW_SHOT_ACOUSTIC ? vb.data() : nullptr;
    d.src = src.data();
    d.six = six;
    d.siz = siz;
    d.data = data.data();
    d.rix = rix.data();
    d.nrec = (int)rix.size();
    d.riz = riz;
    d.x0 = x0;
    d.z0 = z0;
    d.kind = kind;
    for (int j = 0; j < 4 && kind == TSW_SHOT_ELASTIC; ++j) {
      d.ela_fwd[j] = ef[j].data();
      d.ela_bwd[j] = eb[j].data();
    }
    return d;
  }
};

Shot make_shot(int nx, int nz, int nt, int six, int siz, int riz, int x0, int kind) {
  Shot s{nx, nz, nt, six, siz, riz, x0, 0, kind, {}, {}, {}, {}, {}, {}, {}};
  const size_t n = (size_t)nx * nz;
  for (int r = 0; r < nx; r += 3) s.rix.push_back(r);
  s.rix.push_back(nx - 1);
  s.rix.push_back(0);  // duplicated receiver column
  s.src = uni(nt, -1, 1);
  s.data = uni((size_t)nt * s.rix.size(), -1, 1);
  if (kind == TSW_SHOT_ACOUSTIC) {
    s.vf = uni(n, 1500, 2500);
    s.vb = uni(n, 1500, 2500);
  } else {
    // lambda, mu, buoyancies of vp 1800..2600, vs = vp / 2, rho 2000
    for (int j = 0; j < 4; ++j) {
      s.ef[j].resize(n);
      s.eb[j].resize(n);
    }
    for (size_t i = 0; i < n; ++i) {
      for (int m = 0; m < 2; ++m) {
        const float vp = 1800.0f + 800.0f * urand(), vs = 0.5f * vp, rho = 2000.0f;
        std::vector<float>* e = m ? s.eb : s.ef;
        e[1][i] = rho * vs * vs;
        e[0][i] = rho * vp * vp - 2.0f * e[1][i];
        e[2][i] = 1.0f / rho;
        e[3][i] = 1.0f / rho;
      }
    }
  }
  return s;
}

// the shot migrated alone through tsw_migrate_shot(_elastic)
void single(const Shot& s, std::vector<float>& img, std::vector<float>& ill) {
  tsw_grid_t g = nullptr;
  OK(tsw_grid_create(s.nx, s.nz, DX, DX, DT, NSP, 0.0015, 1, &g));
  img.assign((size_t)s.nx * s.nz, 0.0f);
  ill.assign(img.size(), 0.0f);
  if (s.kind == TSW_SHOT_ACOUSTIC) {
    OK(tsw_migrate_shot(g, s.vf.data(), s.vb.data(), s.src.data(), s.nt, s.six, s.siz,
                        s.data.data(), s.nt, s.rix.data(), (int)s.rix.size(), s.riz, K,
                        img.data(), ill.data()));
  } else {
    const float* f[4] = {s.ef[0].data(), s.ef[1].data(), s.ef[2].data(), s.ef[3].data()};
    const float* b[4] = {s.eb[0].data(), s.eb[1].data(), s.eb[2].data(), s.eb[3].data()};
    OK(tsw_migrate_shot_elastic(g, f, b, s.src.data(), s.six, s.siz, s.data.data(), s.nt,
                                s.rix.data(), (int)s.rix.size(), s.riz, K, img.data(),
                                ill.data()));
  }
  OK(tsw_grid_destroy(g));
}

// ---------------------------------------------------------------- refusals

void args_mode() {
  float f[64] = {0};
  int idx[4] = {0, 1, 2, 3};
  tsw_grid_t g = nullptr;
  EXPECT(tsw_version() == TSW_ABI_VERSION);
  BAD(tsw_grid_create(0, 10, DX, DX, DT, NSP, 0.0015, 1, &g));
  BAD(tsw_grid_create(10, 10, -1.0, DX, DT, NSP, 0.0015, 1, &g));
  BAD(tsw_grid_create(10, 10, DX, DX, DT, NSP, 0.0015, 1, nullptr));
  OK(tsw_grid_destroy(nullptr));  // a null handle: nothing to do
  BAD(tsw_grid_sponge(nullptr, f));
  BAD(tsw_migrate_shot(nullptr, f, f, f, 4, 0, 0, f, 4, idx, 1, 0, K, f, f));
  BAD(tsw_laplacian(nullptr, f, 4, 4, DX, DX));
  BAD(tsw_laplacian(f, f, -1, 4, DX, DX));
  // batches: empty list, null list, bad shapes and indices, unknown kernel
  tsw_batch_t b = nullptr;
  Shot s = make_shot(24, 16, 9, 5, 2, 2, 0, TSW_SHOT_ACOUSTIC);
  tsw_shot_desc d = s.desc();
  BAD(tsw_batch_create(0, &d, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  BAD(tsw_batch_create(1, nullptr, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  BAD(tsw_batch_create(1, &d, DX, DX, DT, NSP, 0.0015, 1, 0, 0, 0, 0, &b));
  BAD(tsw_batch_create(1, &d, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 9999, &b));
  tsw_shot_desc e = d;
  e.six = e.nx;
  BAD(tsw_batch_create(1, &e, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  e = d;
  e.riz = -1;
  BAD(tsw_batch_create(1, &e, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  e = d;
  e.nt = 0;
  BAD(tsw_batch_create(1, &e, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  e = d;
  e.vel_fwd = nullptr;
  BAD(tsw_batch_create(1, &e, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  e = d;
  e.kind = 7;
  BAD(tsw_batch_create(1, &e, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  std::vector<int> badrix = s.rix;
  badrix[1] = s.nx;  // a receiver column past the grid
  e = d;
  e.rix = badrix.data();
  BAD(tsw_batch_create(1, &e, DX, DX, DT, NSP, 0.0015, 1, K, 0, 0, 0, &b));
  BAD(tsw_batch_run(nullptr, TSW_RUN_BATCHED));
  BAD(tsw_batch_shot_result(nullptr, 0, f, f));
  OK(tsw_batch_destroy(nullptr));
  int T = 0, W = 0, H = 0, thr = 0;
  EXPECT(tsw_kernel_configs() >= 16);
  for (int c = 1; c <= tsw_kernel_configs(); ++c) OK(tsw_kernel_config(c, &T, &W, &H, &thr));
  BAD(tsw_kernel_config(0, &T, &W, &H, &thr));
  BAD(tsw_kernel_config(tsw_kernel_configs() + 1, &T, &W, &H, &thr));
  // serving loop
  tsw_server_t srv = nullptr;
  BAD(tsw_server_create(DX, DX, DT, NSP, 0.0015, 1, K, 0, 2, 11, 20, &srv));
  BAD(tsw_server_create(DX, DX, DT, NSP, 0.0015, 1, K, 1LL << 30, 0, 11, 20, &srv));
  BAD(tsw_server_create(DX, DX, DT, NSP, 0.0015, 1, K, 1LL << 30, 2, 11, 0, &srv));
  BAD(tsw_server_submit(nullptr, &d, nullptr));
  BAD(tsw_server_poll(nullptr, 1, nullptr, nullptr, nullptr, nullptr));
  OK(tsw_server_destroy(nullptr));
  // misfit gradients
  tsw_fwi_t fw = nullptr;
  tsw_fwi_shot fs;
  std::memset(&fs, 0, sizeof fs);
  fs.nt = 8;
  fs.src = f;
  fs.six = 3;
  fs.siz = 2;
  fs.rix = idx;
  fs.nrec = 4;
  fs.riz = 2;
  fs.d_obs = f;
  BAD(tsw_fwi_create(16, 12, DX, DX, DT, NSP, 0.0015, 1, 2, 1, &fs, 0, &fw));
  BAD(tsw_fwi_create(16, 12, DX, DX, DT, NSP, 0.0015, 1, 4, 0, &fs, 0, &fw));
  fs.six = 16;
  BAD(tsw_fwi_create(16, 12, DX, DX, DT, NSP, 0.0015, 1, 4, 1, &fs, 0, &fw));
  BAD(tsw_fwi_forward(nullptr, f, nullptr, nullptr));
  OK(tsw_fwi_destroy(nullptr));
  // gathers and preprocessing
  int nout = 3;
  BAD(tsw_xcorr(nullptr, 8, 4, idx, 1, idx, 2, 2, 0, f));
  BAD(tsw_xcorr(f, 8, 4, idx, 1, idx, 2, -1, 0, f));
  int far_ch[1] = {4};
  BAD(tsw_xcorr(f, 8, 4, far_ch, 1, idx, 2, 2, 0, f));
  BAD(tsw_xcorr(f, 8, 4, idx, 1, far_ch, 1, 2, 0, f));
  EXPECT(tsw_xcorr_noalloc_scratch_bytes(-1, 1, 1, 1) == -1);
  EXPECT(tsw_xcorr_noalloc_scratch_bytes(100, 2, 3, 4) > 0);
  EXPECT(tsw_xcorr_noalloc_scratch_max() == 64LL << 20);
  BAD(tsw_time_derivative(f, 8, 2, DT, 5, f));
  BAD(tsw_green_branch(f, 8, 2, 8, 0, f, &nout));
  BAD(tsw_green_branch(f, 8, 2, 3, 0, f, &nout));  // causal branch has 5 rows
  BAD(tsw_green_branch(f, 8, 2, 3, 3, f, &nout));
  BAD(tsw_resample_linear(f, 0, 2, DT, DT, 4, f));
  int badmap[4] = {0, 1, 9, -1};
  BAD(tsw_flood_columns(f, 4, 4, 2, badmap, 0, 1.0, f));
  BAD(tsw_flood_columns(f, 3, 4, 2, idx, 0, 1.0, f));
  double sos[6] = {1, 0, 0, 1, 0, 0}, zi[2] = {0, 0};
  BAD(tsw_sosfiltfilt(f, 8, 2, sos, 1, zi, 100, 1, 0, f));  // padlen >= nt
  BAD(tsw_sosfiltfilt(f, 8, 2, sos, 1, zi, 3, 0, 0, f));
  BAD(tsw_sign(nullptr, 4, f));
  BAD(tsw_spectral_whitening(f, 8, 2, 100.0, 30.0, 10.0, 1e-10, f));
  BAD(tsw_fk_filter(f, 8, 2, DT, DX, 10.0, 5.0, f));
  BAD(tsw_sosfiltfilt(f, 8, 2, sos, 1, nullptr, 3, 1, 0, f));
  // host-only CSV writer: a real write, and refusals
  const double rows[2] = {0.5, 1.5}, cols[3] = {0.0, 2.5, 5.0};
  const float m[6] = {1.0f, -2.5f, 3e-8f, INFINITY, -0.0f, NAN};
  char path[64];
  std::snprintf(path, sizeof path, "/tmp/host_san_%d.csv", (int)getpid());
  OK(tsw_save_matrix_csv(path, m, 4, 2, 3, rows, cols, "depth", "offset"));
  std::remove(path);
  BAD(tsw_save_matrix_csv(path, m, 5, 2, 3, rows, cols, "depth", "offset"));
  BAD(tsw_save_matrix_csv(nullptr, m, 4, 2, 3, rows, cols, "depth", "offset"));
}

// ---------------------------------------------------------------- GPU runs

void gpu_shots_batches_server() {
  // uneven shots: sources and receivers on edge rows and columns, nt not a
  // multiple of any kernel's steps per launch, one elastic shot
  std::vector<Shot> shots;
  shots.push_back(make_shot(96, 40, 53, 5, 2, 2, 0, TSW_SHOT_ACOUSTIC));
  shots.push_back(make_shot(48, 24, 41, 47, 0, 0, 30, TSW_SHOT_ACOUSTIC));
  shots.push_back(make_shot(131, 37, 37, 0, 3, 3, 50, TSW_SHOT_ACOUSTIC));
  shots.push_back(make_shot(64, 30, 29, 32, 2, 1, 10, TSW_SHOT_ELASTIC));
  const int ns = (int)shots.size();
  std::vector<std::vector<float>> img(ns), ill(ns);
  for (int i = 0; i < ns; ++i) {
    single(shots[i], img[i], ill[i]);
    EXPECT(finite_nonzero(img[i]) && finite_nonzero(ill[i]));
  }
  std::vector<tsw_shot_desc> d;
  for (auto& s : shots) d.push_back(s.desc());
  int gnx = 0, gnz = 0;
  for (auto& s : shots) {
    gnx = std::max(gnx, s.x0 + s.nx);
    gnz = std::max(gnz, s.z0 + s.nz);
  }
  const int kernels[3] = {0, 11, 15};
  for (int kern : kernels) {
    for (int slots : {0, 2}) {
      tsw_batch_t b = nullptr;
      OK(tsw_batch_create(ns, d.data(), DX, DX, DT, NSP, 0.0015, 1, K, slots, slots ? 16 : 0,
                          kern, &b));
      if (!b) continue;
      for (int mode : {TSW_RUN_BATCHED, TSW_RUN_PLAIN}) {
        OK(tsw_batch_run(b, mode));
        for (int i = 0; i < ns; ++i) {
          std::vector<float> bi(img[i].size()), bl(ill[i].size());
          OK(tsw_batch_shot_result(b, i, bi.data(), bl.data()));
          if (!same(bi, img[i]) || !same(bl, ill[i])) {
            std::printf("batch kernel %d slots %d mode %d shot %d differs\n", kern, slots, mode,
                        i);
            g_fail = 1;
          }
        }
      }
      std::vector<float> si((size_t)gnx * gnz), sl(si.size());
      OK(tsw_batch_stack(b, gnx, gnz, si.data(), sl.data()));
      EXPECT(finite_nonzero(si));
      BAD(tsw_batch_stack(b, gnx - 1, gnz, si.data(), sl.data()));
      BAD(tsw_batch_shot_result(b, ns, si.data(), sl.data()));
      long long cu = 0, bytes = 0, ecu = 0;
      int steps = 0, graphs = 0;
      OK(tsw_batch_info(b, &cu, &bytes, &steps, &graphs, &ecu));
      EXPECT(cu > 0 && bytes > 0 && ecu > 0);
      double ms = 0;
      int launches = 0;
      OK(tsw_batch_time_kernels(b, 0, &ms, &launches));
      EXPECT(launches > 0);
      OK(tsw_batch_destroy(b));
    }
  }
  // multiple shots on one grid, stacked and post-processed
  {
    const Shot& s = shots[0];
    tsw_grid_t g = nullptr;
    OK(tsw_grid_create(s.nx, s.nz, DX, DX, DT, NSP, 0.0015, 1, &g));
    const float* srcs[2] = {s.src.data(), s.src.data()};
    const float* datas[2] = {s.data.data(), s.data.data()};
    const int nts[2] = {s.nt, s.nt}, sixs[2] = {s.six, s.nx - 1}, sizs[2] = {s.siz, 0};
    std::vector<float> raw(img[0].size()), lap(raw.size()), nrm(raw.size());
    OK(tsw_migrate_shots(g, s.vf.data(), s.vb.data(), 2, srcs, nts, sixs, sizs, datas, nts,
                         s.rix.data(), (int)s.rix.size(), s.riz, K, 1, raw.data(), lap.data(),
                         nrm.data()));
    EXPECT(finite_nonzero(raw) && finite_nonzero(nrm));
    // a caller-modified sponge goes through the per-shot path
    std::vector<float> sp(raw.size());
    OK(tsw_grid_sponge(g, sp.data()));
    for (auto& x : sp) x *= 0.999f;
    OK(tsw_grid_set_sponge(g, sp.data()));
    OK(tsw_migrate_shots(g, s.vf.data(), s.vb.data(), 2, srcs, nts, sixs, sizs, datas, nts,
                         s.rix.data(), (int)s.rix.size(), s.riz, K, 0, raw.data(), lap.data(),
                         nrm.data()));
    // forward with snapshots, then back-propagation into an image
    const long long nsnap = (s.nt - 1) / K + 1;
    std::vector<float> rec((size_t)s.nt * s.rix.size()), snaps((size_t)nsnap * raw.size());
    std::vector<float> im(raw.size(), 0.0f);
    OK(tsw_forward_acoustic(g, s.vf.data(), s.src.data(), s.nt, s.six, s.siz, s.rix.data(),
                            (int)s.rix.size(), s.riz, K, rec.data(), snaps.data()));
    OK(tsw_backward_correlate(g, s.vb.data(), s.data.data(), s.nt, s.rix.data(),
                              (int)s.rix.size(), s.riz, snaps.data(), nsnap, im.data(), K));
    EXPECT(finite_nonzero(im));
    BAD(tsw_backward_correlate(g, s.vb.data(), s.data.data(), s.nt, s.rix.data(),
                               (int)s.rix.size(), s.nz, snaps.data(), nsnap, im.data(), K));
    OK(tsw_grid_destroy(g));
  }
  // serving loop: uneven requests, a tight slot limit, a cancelled request
  // and an injected admission failure
  {
    tsw_server_t srv = nullptr;
    OK(tsw_server_create(DX, DX, DT, NSP, 0.0015, 1, K, 1LL << 30, 2, 11, 13, &srv));
    if (srv) {
      std::vector<long long> ids(ns, 0);
      for (int i = 0; i < ns; ++i) OK(tsw_server_submit(srv, &d[i], &ids[i]));
      long long extra = 0, failing = 0;
      OK(tsw_server_submit(srv, &d[1], &extra));
      OK(tsw_server_cancel(srv, extra));
      for (int i = 0; i < ns; ++i) {
        OK(tsw_server_wait(srv, ids[i], 120.0));
        std::vector<float> ri(img[i].size()), rl(ill[i].size());
        OK(tsw_server_result(srv, ids[i], ri.data(), rl.data()));
        if (!same(ri, img[i]) || !same(rl, ill[i])) {
          std::printf("served request %d differs\n", i);
          g_fail = 1;
        }
      }
      tsw_debug_fail_next(1);
      OK(tsw_server_submit(srv, &d[2], &failing));
      OK(tsw_server_wait(srv, failing, 120.0));
      int state = -1;
      double ts = 0, ta = 0, td = 0;
      OK(tsw_server_poll(srv, failing, &state, &ts, &ta, &td));
      EXPECT(state == 3);
      EXPECT(tsw_server_result(srv, failing, nullptr, nullptr) == -4);
      EXPECT(tsw_server_result(srv, extra, nullptr, nullptr) == -4);
      BAD(tsw_server_poll(srv, extra, &state, &ts, &ta, &td));
      long long c[7];
      double busy = 0;
      OK(tsw_server_stats(srv, &c[0], &c[1], &c[2], &c[3], &c[4], &c[5], &c[6], &busy));
      EXPECT(c[0] == ns && c[1] == 1 && c[2] == 1 && c[5] == 0);
      // requests still queued when the server is destroyed
      long long left = 0;
      OK(tsw_server_submit(srv, &d[0], &left));
      OK(tsw_server_destroy(srv));
    }
  }
}

void gpu_fwi_gathers() {
  // misfit gradient of two shots, FP32 and FP64
  const int nx = 48, nz = 30, nt = 40;
  std::vector<int> rix = {0, 5, 17, 47};
  std::vector<float> src = uni(nt, -1, 1), obs = uni((size_t)nt * rix.size(), -1, 1);
  std::vector<double> src64(src.begin(), src.end()), obs64(obs.begin(), obs.end());
  for (int prec : {4, 8}) {
    tsw_fwi_shot sh[2];
    for (int i = 0; i < 2; ++i) {
      std::memset(&sh[i], 0, sizeof sh[i]);
      sh[i].nt = nt;
      sh[i].src = prec == 4 ? (const void*)src.data() : (const void*)src64.data();
      sh[i].six = i ? nx - 1 : 3;
      sh[i].siz = i ? 0 : 2;
      sh[i].rix = rix.data();
      sh[i].nrec = (int)rix.size();
      sh[i].riz = 2;
      sh[i].d_obs = prec == 4 ? (const void*)obs.data() : (const void*)obs64.data();
    }
    tsw_fwi_t fw = nullptr;
    OK(tsw_fwi_create(nx, nz, DX, DX, DT, NSP, 0.0015, 1, prec, 2, sh, 0, &fw));
    if (!fw) continue;
    const size_t n = (size_t)nx * nz;
    void *dv = nullptr, *dg = nullptr;
    std::vector<float> v32 = uni(n, 1600, 2400);
    std::vector<double> v64(v32.begin(), v32.end());
    if (hipMalloc(&dv, n * prec) != hipSuccess || hipMalloc(&dg, n * prec) != hipSuccess) {
      fail("hipMalloc", __LINE__, -2);
      return;
    }
    (void)hipMemcpy(dv, prec == 4 ? (const void*)v32.data() : (const void*)v64.data(), n * prec,
                    hipMemcpyHostToDevice);
    for (int cfg : {0, 2}) {
      if (prec == 4) {
        OK(tsw_fwi_set_blocking(fw, cfg));
        OK(tsw_fwi_set_adjoint_blocking(fw, cfg));
      }
      double loss = 0;
      OK(tsw_fwi_forward(fw, dv, nullptr, &loss));
      EXPECT(loss > 0 && std::isfinite(loss));
      OK(tsw_fwi_backward(fw, dv, dg, nullptr));
    }
    std::vector<char> pred((size_t)nt * rix.size() * prec), probe((size_t)nt * prec);
    OK(tsw_fwi_predicted(fw, 1, pred.data()));
    OK(tsw_fwi_set_residual(fw, 0, pred.data()));
    OK(tsw_fwi_backward(fw, dv, dg, nullptr));
    OK(tsw_fwi_probe(fw, 0, probe.data()));
    BAD(tsw_fwi_predicted(fw, 2, pred.data()));
    OK(tsw_fwi_destroy(fw));
    (void)hipFree(dv);
    (void)hipFree(dg);
  }
  // virtual-shot gathers: correlations on the host and device paths
  const int cnt = 300, nch = 24, nlag = 20, nl = 2 * nlag + 1;
  std::vector<float> x = uni((size_t)cnt * nch, -1, 1);
  std::vector<int> srcch = {0, 7, 23}, recch = {1, 2, 3, 10, 23, 0};
  const size_t no = srcch.size() * nl * recch.size();
  std::vector<float> xc(no), xd(no), xn(no);
  OK(tsw_xcorr(x.data(), cnt, nch, srcch.data(), 3, recch.data(), 6, nlag, 0, xc.data()));
  OK(tsw_xcorr(x.data(), cnt, nch, srcch.data(), 3, recch.data(), 6, nlag, 1, xc.data()));
  float *dx = nullptr, *dout = nullptr;
  if (hipMalloc(&dx, x.size() * 4) != hipSuccess || hipMalloc(&dout, no * 4) != hipSuccess) {
    fail("hipMalloc", __LINE__, -2);
    return;
  }
  (void)hipMemcpy(dx, x.data(), x.size() * 4, hipMemcpyHostToDevice);
  OK(tsw_xcorr_device(dx, cnt, nch, srcch.data(), 3, recch.data(), 6, nlag, 0, dout, nullptr));
  (void)hipMemcpy(xd.data(), dout, no * 4, hipMemcpyDeviceToHost);
  OK(tsw_xcorr_device_noalloc(dx, cnt, nch, srcch.data(), 3, recch.data(), 6, nlag, 0, dout,
                              nullptr));
  (void)hipMemcpy(xn.data(), dout, no * 4, hipMemcpyDeviceToHost);
  EXPECT(same(xd, xn) && finite_nonzero(xd));
  // derivative, branches, resampling of one correlogram block (nl, 6)
  std::vector<float> blk(xd.begin(), xd.begin() + (size_t)nl * 6), der(blk.size());
  OK(tsw_time_derivative(blk.data(), nl, 6, DT, 0, der.data()));
  OK(tsw_time_derivative(blk.data(), nl, 6, DT, 1, der.data()));
  for (int br = 0; br < 3; ++br) {
    int m = br == 0 ? nl - nlag : br == 1 ? nlag + 1 : nlag + 1;
    std::vector<float> o((size_t)m * 6);
    OK(tsw_green_branch(der.data(), nl, 6, nlag, br, o.data(), &m));
  }
  std::vector<float> rs((size_t)77 * 6);
  OK(tsw_resample_linear(blk.data(), nl, 6, 2e-3, DT, 77, rs.data()));
  // flood columns, Laplacian, preprocessing
  std::vector<float> base = uni(5 * 8, 1500, 3000), flood(5 * 8);
  std::vector<double> base64(base.begin(), base.end());
  std::vector<int> cmap = {0, 0, 1, 2, -1, 7, 7, 3};
  OK(tsw_flood_columns(base.data(), 4, 8, 5, cmap.data(), 1, 1.1, flood.data()));
  OK(tsw_flood_columns(base64.data(), 8, 8, 5, cmap.data(), 0, 1.0, flood.data()));
  std::vector<float> lap(base.size());
  OK(tsw_laplacian(base.data(), lap.data(), 8, 5, DX, DX));
  double sos[12] = {0.2, 0.4, 0.2, 1.0, -0.3, 0.1, 1.0, -2.0, 1.0, 1.0, -1.6, 0.65};
  double zi[4] = {0.0, 0.0, 0.0, 0.0};
  std::vector<float> filt((size_t)((cnt + 1) / 2) * nch), sg(x.size()), wh(x.size());
  OK(tsw_sosfiltfilt(x.data(), cnt, nch, sos, 2, zi, 9, 2, 0, filt.data()));
  OK(tsw_sosfiltfilt(x.data(), cnt, nch, sos, 2, zi, 9, 1, 1, sg.data()));
  OK(tsw_sign(x.data(), (long long)x.size(), sg.data()));
  OK(tsw_spectral_whitening(x.data(), cnt, nch, 500.0, 5.0, 80.0, 1e-10, wh.data()));
  OK(tsw_fk_filter(x.data(), cnt, nch, 2e-3, DX, 100.0, 4000.0, wh.data()));
  float* dpre[2] = {nullptr, nullptr};
  const int nq = (cnt + 1) / 2;
  for (auto& p : dpre)
    if (hipMalloc(&p, (size_t)nq * nch * 4) != hipSuccess) {
      fail("hipMalloc", __LINE__, -2);
      return;
    }
  OK(tsw_preprocess_device(x.data(), cnt, nch, sos, 2, zi, 9, 2, sos, 2, zi, 9, 250.0, 5.0,
                           80.0, 1e-10, dpre[0], nullptr));
  const float* xs[2] = {x.data(), x.data()};
  OK(tsw_preprocess_device_blocks(2, xs, cnt, nch, sos, 2, zi, 9, 2, sos, 2, zi, 9, 250.0, 5.0,
                                  80.0, 1e-10, dpre, nullptr));
  std::vector<float> p0((size_t)nq * nch), p1(p0.size());
  (void)hipMemcpy(p0.data(), dpre[0], p0.size() * 4, hipMemcpyDeviceToHost);
  (void)hipMemcpy(p1.data(), dpre[1], p1.size() * 4, hipMemcpyDeviceToHost);
  EXPECT(same(p0, p1));
  for (auto& p : dpre) (void)hipFree(p);
  (void)hipFree(dx);
  (void)hipFree(dout);
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "args";
  if (mode == "args") {
    args_mode();
  } else if (mode == "gpu") {
    OK(tsw_init(0));
    args_mode();
    gpu_shots_batches_server();
    gpu_fwi_gathers();
  } else {
    std::printf("usage: host_san args|gpu\n");
    return 2;
  }
  if (g_fail) return 1;
  std::printf("ok\n");
  return 0;
}
