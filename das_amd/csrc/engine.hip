// C-ABI of the tsw_rtm HIP engine: reference-compatible kernels with host
// pointers (tsw_rtm/c_engine/wave2d.c) and the device-resident time loops
// behind tsw_rtm/wave_solver.py and tsw_rtm/rtm_engine.py.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <string>

#include "engine.h"
#include "batch.h"
#include "post.h"
#include "tswrtm.h"

namespace tsw {

static thread_local std::string g_err;
static std::mutex g_fail_mu;
static int g_fail_pending = 0;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

int fail_injected() {
  std::lock_guard<std::mutex> lk(g_fail_mu);
  if (g_fail_pending > 0) {
    --g_fail_pending;
    set_error("injected failure (tsw_debug_fail_next)");
    return kErrInjected;
  }
  return 0;
}

int DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes && p) return kOk;
  release();
  if (nbytes == 0) nbytes = 4;
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    set_error("hipMalloc(%zu bytes) failed: %s", nbytes, hipGetErrorString(e));
    return kErrMem;
  }
  bytes = nbytes;
  return kOk;
}

static Ctx g_ctx;
static std::mutex g_ctx_mu;
// Serialises the host-pointer entry points, which share scratch device buffers.
static std::recursive_mutex g_api_mu;
std::recursive_mutex& api_mutex() { return g_api_mu; }

int ctx_get(Ctx** out) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (g_ctx.stream == nullptr) {
    int dev = 0;
    TSW_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    TSW_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      set_error("device %d is %s; this engine is built for gfx950 only", dev, prop.gcnArchName);
      return kErrState;
    }
    TSW_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    g_ctx.device = dev;
  }
  *out = &g_ctx;
  return kOk;
}

int RecvMap::build(const int* rix, int nrec, int nx) {
  off.assign(nx + 1, 0);
  idx.assign(nrec > 0 ? nrec : 1, 0);
  for (int r = 0; r < nrec; ++r) {
    TSW_ARG(rix[r] >= 0 && rix[r] < nx, "receiver %d column %d outside [0, %d)", r, rix[r], nx);
    off[rix[r] + 1]++;
  }
  for (int x = 0; x < nx; ++x) off[x + 1] += off[x];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int r = 0; r < nrec; ++r) idx[fill[rix[r]]++] = r;  // stable: receiver order per column
  return kOk;
}

int check_finite(const float* const* arrays, const long long* n, int narr, int per_item,
                 int index_base, const char* what, hipStream_t st) {
  // counters kept across calls; the callers hold the library lock or the
  // serving loop's worker thread, and each call drains its stream
  static std::mutex mu;
  static DevBuf* counts = new DevBuf();
  std::lock_guard<std::mutex> lk(mu);
  if (narr <= 0) return kOk;
  TSW_TRY(counts->ensure((size_t)narr * sizeof(unsigned)));
  launch_fill(counts->f(), narr, 0.0f, st);  // all-zero bits
  for (int i = 0; i < narr; ++i)
    launch_count_nonfinite(arrays[i], n[i], counts->u() + i, st);
  std::vector<unsigned> h(narr);
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(h.data(), counts->p, narr * sizeof(unsigned), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < narr; ++i) {
    if (h[i]) {
      set_error("%s %d: %u non-finite image / illumination values: the wavefield diverged "
                "(dt may exceed the CFL limit; see WaveSolver2D.check_stability)",
                what, index_base + i / per_item, h[i]);
      return kErrDiverged;
    }
  }
  return kOk;
}

void sponge_table(std::vector<float>& t, int n, float alpha) {
  t.assign(n + 1, 1.0f);
  for (int d = 1; d <= n; ++d) {
    const float df = (float)d;
    t[d] = expf(-alpha * df * df);
  }
}

AcuStep Grid::base() const {
  AcuStep s;
  std::memset(&s, 0, sizeof s);
  s.nx = nx;
  s.nz = nz;
  s.fs = fs;
  s.rdx2 = rdx2;
  s.rdz2 = rdz2;
  s.inj = inj;
  s.src_cell = -1;
  s.src_it = -1;
  s.rdata_it = -1;
  s.rec_it = -1;
  s.riz = -1;
  return s;
}

int grid_init(Grid* g, int nx, int nz, double dx, double dz, double dt, int n_sponge,
              double alpha, int fs) {
  TSW_ARG(nx >= 5 && nz >= 5, "grid %d x %d: nx and nz must be >= 5", nx, nz);
  TSW_ARG((long long)nx * nz < (1LL << 31), "grid %d x %d exceeds 2^31 cells", nx, nz);
  TSW_ARG(dx > 0 && dz > 0 && dt > 0, "dx, dz, dt must be positive");
  TSW_ARG(n_sponge >= 0, "n_sponge must be >= 0");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  g->nx = nx;
  g->nz = nz;
  g->dx = dx;
  g->dz = dz;
  g->dt = dt;
  g->alpha = alpha;
  g->n_sponge = n_sponge;
  g->fs = fs ? 1 : 0;
  g->dxf = (float)dx;
  g->dzf = (float)dz;
  g->dtf = (float)dt;
  g->rdx2 = 1.0f / (g->dxf * g->dxf);
  g->rdz2 = 1.0f / (g->dzf * g->dzf);
  g->dt2 = g->dtf * g->dtf;
  g->inj = (float)(dt * dt);
  const size_t nb = (size_t)g->cells() * sizeof(float);
  TSW_TRY(g->sponge.ensure(nb));
  std::vector<float> tab;
  sponge_table(tab, n_sponge, (float)alpha);
  TSW_TRY(g->table.ensure(tab.size() * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(g->table.p, tab.data(), tab.size() * sizeof(float),
                         hipMemcpyHostToDevice, ctx->stream));
  launch_sponge(g->sponge.f(), nx, nz, n_sponge, g->fs, g->table.f(), ctx->stream);
  TSW_HIP(hipGetLastError());
  for (auto& b : g->buf) TSW_TRY(b.ensure(nb));
  TSW_HIP(hipStreamSynchronize(ctx->stream));
  return kOk;
}

// ---------------------------------------------------------------- loops
namespace {

struct H2D {
  hipStream_t st;
  int operator()(DevBuf& d, const void* h, size_t bytes) {
    TSW_TRY(d.ensure(bytes));
    if (bytes) TSW_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, st));
    return kOk;
  }
};

int upload_receivers(Grid* g, const int* rix, int nrec, hipStream_t st) {
  RecvMap m;
  TSW_TRY(m.build(rix, nrec, g->nx));
  H2D up{st};
  TSW_TRY(up(g->roff, m.off.data(), m.off.size() * sizeof(int)));
  TSW_TRY(up(g->ridx, m.idx.data(), m.idx.size() * sizeof(int)));
  // host vectors die at return: make the copies complete first
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int upload_model(Grid* g, const float* vel, int slot, hipStream_t st) {
  const size_t nb = (size_t)g->cells() * sizeof(float);
  H2D up{st};
  TSW_TRY(up(g->vel, vel, nb));
  TSW_TRY(g->coef[slot].ensure(nb));
  launch_coef(g->vel.f(), g->coef[slot].f(), g->cells(), g->dt2, st);
  TSW_HIP(hipGetLastError());
  return kOk;
}

// Forward loop on device-resident inputs (wave_solver.py:146-186).
// snaps: nullptr or ((nt-1)/k+1) levels; illum: nullptr or accumulated sum of squares.
int run_forward(Grid* g, int slot, const float* d_src, int nt, int six, int siz, int nrec,
                int riz, int k, float* d_rec, float* d_snaps, float* d_illum, hipStream_t st) {
  const long long n = g->cells();
  for (auto& b : g->buf) launch_fill(b.f(), n, 0.0f, st);
  AcuStep s = g->base();
  s.coef = g->coef[slot].f();
  s.sponge = g->sponge.f();
  s.src = d_src;
  s.src_cell = siz * g->nx + six;
  s.riz = riz;
  s.nrec = nrec;
  s.roff = g->roff.i();
  s.ridx = g->ridx.i();
  s.rec = d_rec;
  // iteration 0 injection into the buffer that is p_curr at it = 0
  AcuStep s0 = s;
  s0.nxt = g->buf[1].f();
  s0.src_it = 0;
  launch_acoustic_inject(s0, st);
  for (int it = 0; it < nt; ++it) {
    s.prv = g->buf[it % 3].f();
    s.cur = g->buf[(it + 1) % 3].f();
    s.nxt = g->buf[(it + 2) % 3].f();
    s.src_it = it + 1 < nt ? it + 1 : -1;
    s.rec_it = d_rec ? it : -1;
    s.snap = nullptr;
    s.illum = nullptr;
    if (d_snaps && it % k == 0) {
      s.snap = d_snaps + (long long)(it / k) * n;
      s.illum = d_illum;
    }
    launch_acoustic_step(s, st);
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}

// Back-propagation + imaging loop (wave_solver.py:201-242), snapshot cursor
// rule included.  d_snaps holds nsnap levels.
int run_backward(Grid* g, int slot, const float* d_data, int nt, int nrec, int riz,
                 const float* d_snaps, long long nsnap, float* d_image, int k, hipStream_t st) {
  const long long n = g->cells();
  for (auto& b : g->buf) launch_fill(b.f(), n, 0.0f, st);
  AcuStep s = g->base();
  s.coef = g->coef[slot].f();
  s.sponge = g->sponge.f();
  s.riz = riz;
  s.nrec = nrec;
  s.roff = g->roff.i();
  s.ridx = g->ridx.i();
  s.rdata = d_data;
  s.w = (float)(g->dt * k);
  AcuStep s0 = s;
  s0.nxt = g->buf[1].f();
  s0.rdata_it = nt - 1;
  launch_acoustic_inject(s0, st);
  long long cursor = (nt - 1) / k;
  for (int m = 0; m < nt; ++m) {
    const int it = nt - 1 - m;
    s.prv = g->buf[m % 3].f();
    s.cur = g->buf[(m + 1) % 3].f();
    s.nxt = g->buf[(m + 2) % 3].f();
    s.rdata_it = it >= 1 ? it - 1 : -1;
    s.image = nullptr;
    s.isnap = nullptr;
    if (it % k == 0 && cursor >= 0 && cursor < nsnap) {
      s.image = d_image;
      s.isnap = d_snaps + cursor * n;
      --cursor;
    }
    launch_acoustic_step(s, st);
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}

int check_shot_args(Grid* g, int nt, int nrec, int riz, int k) {
  TSW_ARG(nt >= 1, "nt must be >= 1 (got %d)", nt);
  TSW_ARG(nrec >= 0, "n_rec must be >= 0");
  TSW_ARG(riz >= 0 && riz < g->nz, "receiver_iz %d outside [0, %d)", riz, g->nz);
  TSW_ARG(k >= 1, "snapshot_interval must be >= 1 (got %d)", k);
  return kOk;
}

}  // namespace

int migrate_shot_dev(Grid* g, const float* d_src, int nt_src, int six, int siz,
                     const float* d_data, int nt, int nrec, int riz, int k, float* d_image,
                     float* d_illum, hipStream_t st) {
  const long long n = g->cells();
  const long long nsnap = (nt_src - 1) / k + 1;
  TSW_TRY(g->snaps.ensure((size_t)nsnap * n * sizeof(float)));
  launch_fill(d_illum, n, 0.0f, st);
  TSW_TRY(run_forward(g, 0, d_src, nt_src, six, siz, nrec, riz, k, nullptr, g->snaps.f(), d_illum,
                      st));
  launch_scale(d_illum, n, (float)(g->dt * k), st);
  launch_fill(d_image, n, 0.0f, st);
  TSW_TRY(run_backward(g, 1, d_data, nt, nrec, riz, g->snaps.f(), nsnap, d_image, k, st));
  return kOk;
}

// rtm_engine.py:145-158 on device: mute stacked/illum in place, then
// normalized -> norm_out, Laplacian (muted) -> lap_out.
int postprocess_dev(Grid* g, float* stacked, float* illum, int mute, float* norm_out,
                    float* lap_out, unsigned* maxbits, hipStream_t st) {
  const long long n = g->cells();
  if (mute) launch_mute(stacked, illum, g->sponge.f(), n, st);
  launch_fill(reinterpret_cast<float*>(maxbits), 1, 0.0f, st);  // bit pattern 0
  launch_max_nonneg(illum, n, maxbits, st);
  launch_normalize(stacked, illum, norm_out, n, maxbits, st);
  launch_laplacian(stacked, lap_out, g->nx, g->nz, g->dx, g->dz, mute ? g->sponge.f() : nullptr,
                   st);
  TSW_HIP(hipGetLastError());
  return kOk;
}

}  // namespace tsw

using namespace tsw;

extern "C" {

const char* tsw_last_error(void) { return g_err.c_str(); }

int tsw_version(void) { return TSW_ABI_VERSION; }

int tsw_init(int device) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  if (device >= 0) TSW_HIP(hipSetDevice(device));
  Ctx* c;
  return ctx_get(&c);
}

int tsw_mem_info(long long* free_bytes, long long* total_bytes) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  size_t f = 0, t = 0;
  TSW_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = (long long)f;
  if (total_bytes) *total_bytes = (long long)t;
  return kOk;
}

int tsw_device_name(char* out, int len) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  TSW_ARG(out && len > 0, "bad buffer");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipDeviceProp_t prop;
  TSW_HIP(hipGetDeviceProperties(&prop, ctx->device));
  snprintf(out, (size_t)len, "%s:%d", prop.gcnArchName, prop.multiProcessorCount);
  return kOk;
}

void tsw_debug_fail_next(int count) {
  std::lock_guard<std::mutex> lk(g_fail_mu);
  g_fail_pending = count;
}

int tsw_grid_create(int nx, int nz, double dx, double dz, double dt, int n_sponge,
                    double sponge_alpha, int free_surface_top, tsw_grid_t* out) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  TSW_ARG(out != nullptr, "out is null");
  *out = nullptr;
  Grid* g = new Grid();
  int rc = grid_init(g, nx, nz, dx, dz, dt, n_sponge, sponge_alpha, free_surface_top);
  if (rc != kOk) {
    delete g;
    return rc;
  }
  *out = reinterpret_cast<tsw_grid_t>(g);
  return kOk;
}

int tsw_grid_destroy(tsw_grid_t h) {
  if (!h) return kOk;
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  TSW_HIP(hipStreamSynchronize(ctx->stream));
  delete reinterpret_cast<Grid*>(h);
  return kOk;
}

int tsw_grid_sponge(tsw_grid_t h, float* out) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && out, "null argument");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  TSW_HIP(hipMemcpyAsync(out, g->sponge.p, g->cells() * sizeof(float), hipMemcpyDeviceToHost,
                         ctx->stream));
  TSW_HIP(hipStreamSynchronize(ctx->stream));
  return kOk;
}

int tsw_grid_set_sponge(tsw_grid_t h, const float* sponge) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && sponge, "null argument");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  TSW_HIP(hipMemcpyAsync(g->sponge.p, sponge, g->cells() * sizeof(float), hipMemcpyHostToDevice,
                         ctx->stream));
  TSW_HIP(hipStreamSynchronize(ctx->stream));
  g->custom_sponge = true;
  return kOk;
}

int tsw_forward_acoustic(tsw_grid_t h, const float* vel, const float* src, int nt, int six,
                         int siz, const int* rix, int nrec, int riz, int k, float* rec_out,
                         float* snaps_out) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  TSW_TRY(fail_injected());
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && vel && src, "null argument");
  TSW_TRY(check_shot_args(g, nt, nrec, riz, k));
  TSW_ARG(six >= 0 && six < g->nx && siz >= 0 && siz < g->nz, "source (%d, %d) outside grid",
          six, siz);
  TSW_ARG(nrec == 0 || (rix && rec_out), "receivers need rix and rec_out");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = g->cells();
  H2D up{st};
  TSW_TRY(upload_model(g, vel, 0, st));
  TSW_TRY(up(g->src, src, (size_t)nt * sizeof(float)));
  TSW_TRY(upload_receivers(g, rix, nrec, st));
  TSW_TRY(g->rec.ensure((size_t)nt * (nrec > 0 ? nrec : 1) * sizeof(float)));
  const long long nsnap = (nt - 1) / k + 1;
  if (snaps_out) {
    TSW_TRY(g->snaps.ensure((size_t)nsnap * n * sizeof(float)));
    TSW_TRY(g->illum.ensure((size_t)n * sizeof(float)));
    launch_fill(g->illum.f(), n, 0.0f, st);
  }
  TSW_TRY(run_forward(g, 0, g->src.f(), nt, six, siz, nrec, riz, k, nrec ? g->rec.f() : nullptr,
                      snaps_out ? g->snaps.f() : nullptr, snaps_out ? g->illum.f() : nullptr, st));
  if (nrec)
    TSW_HIP(hipMemcpyAsync(rec_out, g->rec.p, (size_t)nt * nrec * sizeof(float),
                           hipMemcpyDeviceToHost, st));
  if (snaps_out)
    TSW_HIP(hipMemcpyAsync(snaps_out, g->snaps.p, (size_t)nsnap * n * sizeof(float),
                           hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_backward_correlate(tsw_grid_t h, const float* vel, const float* data, int nt,
                           const int* rix, int nrec, int riz, const float* snaps,
                           long long nsnap, float* image_inout, int k) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  TSW_TRY(fail_injected());
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && vel && image_inout, "null argument");
  TSW_TRY(check_shot_args(g, nt, nrec, riz, k));
  TSW_ARG(nrec == 0 || (rix && data), "receivers need rix and data");
  TSW_ARG(nsnap >= 0 && (nsnap == 0 || snaps), "snapshots missing");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = g->cells();
  H2D up{st};
  TSW_TRY(upload_model(g, vel, 1, st));
  TSW_TRY(up(g->data, data, (size_t)nt * nrec * sizeof(float)));
  TSW_TRY(upload_receivers(g, rix, nrec, st));
  // only the snapshots the cursor can reach are needed on the device
  long long reach = (nt - 1) / k + 1;
  long long used = nsnap < reach ? nsnap : reach;
  if ((nt - 1) / k >= nsnap) used = 0;  // cursor starts past the end: never images
  TSW_TRY(up(g->isnap, snaps, (size_t)used * n * sizeof(float)));
  TSW_TRY(up(g->image, image_inout, (size_t)n * sizeof(float)));
  TSW_TRY(run_backward(g, 1, g->data.f(), nt, nrec, riz, g->isnap.f(), used, g->image.f(), k, st));
  TSW_HIP(hipMemcpyAsync(image_inout, g->image.p, (size_t)n * sizeof(float),
                         hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_migrate_shot(tsw_grid_t h, const float* vel_fwd, const float* vel_bwd, const float* src,
                     int nt_src, int six, int siz, const float* data, int nt, const int* rix,
                     int nrec, int riz, int k, float* image_out, float* illum_out) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  TSW_TRY(fail_injected());
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && vel_fwd && vel_bwd && src && image_out && illum_out, "null argument");
  TSW_TRY(check_shot_args(g, nt, nrec, riz, k));
  TSW_ARG(nt_src >= 1, "source length must be >= 1");
  // The reference runs the forward over len(source_ts) and the back-propagation
  // over receiver_data.shape[0] (rtm_engine.py:76-104, wave_solver.py:146,201);
  // unequal lengths silently pair shifted snapshots, so they are rejected.
  TSW_ARG(nt_src == nt, "source length %d differs from receiver data length %d", nt_src, nt);
  TSW_ARG(six >= 0 && six < g->nx && siz >= 0 && siz < g->nz, "source (%d, %d) outside grid",
          six, siz);
  TSW_ARG(nrec == 0 || (rix && data), "receivers need rix and data");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = g->cells();
  if (!g->custom_sponge) {
    // one-shot batch: the temporally blocked kernel, bit-identical to the
    // per-step path below
    tsw_shot_desc d;
    std::memset(&d, 0, sizeof d);
    d.nx = g->nx;
    d.nz = g->nz;
    d.nt = nt;
    d.vel_fwd = vel_fwd;
    d.vel_bwd = vel_bwd;
    d.src = src;
    d.six = six;
    d.siz = siz;
    d.data = data;
    d.rix = rix;
    d.nrec = nrec;
    d.riz = riz;
    d.kind = TSW_SHOT_ACOUSTIC;
    tsw_batch_t b = nullptr;
    TSW_TRY(tsw_batch_create(1, &d, g->dx, g->dz, g->dt, g->n_sponge, g->alpha, g->fs, k, 1, 0,
                             kDefaultBlockedKernel, &b));
    int rc = tsw_batch_run(b, TSW_RUN_BATCHED);
    if (rc == kOk || rc == kErrDiverged) {  // a diverged shot's result is still copied out
      const std::string err = rc ? std::string(tsw_last_error()) : std::string();
      const int r2 = tsw_batch_shot_result(b, 0, image_out, illum_out);
      if (r2 != kOk) rc = r2;
      else if (rc) set_error("%s", err.c_str());
    }
    tsw_batch_destroy(b);
    return rc;
  }
  H2D up{st};
  TSW_TRY(upload_model(g, vel_fwd, 0, st));
  TSW_HIP(hipStreamSynchronize(st));  // g->vel is reused for the second model
  TSW_TRY(upload_model(g, vel_bwd, 1, st));
  TSW_TRY(up(g->src, src, (size_t)nt_src * sizeof(float)));
  TSW_TRY(up(g->data, data, (size_t)nt * nrec * sizeof(float)));
  TSW_TRY(upload_receivers(g, rix, nrec, st));
  TSW_TRY(g->image.ensure((size_t)n * sizeof(float)));
  TSW_TRY(g->illum.ensure((size_t)n * sizeof(float)));
  // Backward pairs iteration it with forward snapshot it/k while the cursor,
  // started at (nt-1)/k, stays below the forward's snapshot count.
  TSW_TRY(migrate_shot_dev(g, g->src.f(), nt_src, six, siz, g->data.f(), nt, nrec, riz, k,
                           g->image.f(), g->illum.f(), st));
  TSW_HIP(hipMemcpyAsync(image_out, g->image.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                         st));
  TSW_HIP(hipMemcpyAsync(illum_out, g->illum.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                         st));
  TSW_HIP(hipStreamSynchronize(st));
  const float* arr[2] = {g->image.f(), g->illum.f()};
  const long long len[2] = {n, n};
  return check_finite(arr, len, 2, 2, 0, "shot", st);
}

int tsw_migrate_shots(tsw_grid_t h, const float* vel_fwd, const float* vel_bwd, int nshot,
                      const float* const* srcs, const int* nt_srcs, const int* sixs,
                      const int* sizs, const float* const* datas, const int* nts, const int* rix,
                      int nrec, int riz, int k, int mute, float* raw_out, float* lap_out,
                      float* norm_out) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  TSW_TRY(fail_injected());
  Grid* g = reinterpret_cast<Grid*>(h);
  TSW_ARG(g && vel_fwd && vel_bwd && raw_out && lap_out && norm_out, "null argument");
  TSW_ARG(nshot >= 0, "nshot must be >= 0");
  TSW_ARG(nshot == 0 || (srcs && nt_srcs && sixs && sizs && datas && nts), "null shot arrays");
  for (int i = 0; i < nshot; ++i) {
    TSW_TRY(check_shot_args(g, nts[i], nrec, riz, k));
    TSW_ARG(nt_srcs[i] >= 1 && srcs[i], "shot %d: empty source", i);
    TSW_ARG(nt_srcs[i] == nts[i], "shot %d: source length %d differs from data length %d", i,
            nt_srcs[i], nts[i]);
    TSW_ARG(sixs[i] >= 0 && sixs[i] < g->nx && sizs[i] >= 0 && sizs[i] < g->nz,
            "shot %d: source (%d, %d) outside grid", i, sixs[i], sizs[i]);
    TSW_ARG(nrec == 0 || datas[i], "shot %d: data missing", i);
  }
  TSW_ARG(nrec == 0 || rix, "rix missing");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = g->cells();
  const size_t nb = (size_t)n * sizeof(float);
  H2D up{st};
  TSW_TRY(upload_model(g, vel_fwd, 0, st));
  TSW_HIP(hipStreamSynchronize(st));
  TSW_TRY(upload_model(g, vel_bwd, 1, st));
  TSW_TRY(upload_receivers(g, rix, nrec, st));
  DevBuf stacked, total, shot_img, shot_ill, out, maxbits;
  TSW_TRY(stacked.ensure(nb));
  TSW_TRY(total.ensure(nb));
  TSW_TRY(shot_img.ensure(nb));
  TSW_TRY(shot_ill.ensure(nb));
  TSW_TRY(out.ensure(nb));
  TSW_TRY(maxbits.ensure(sizeof(unsigned)));
  std::string diverged;
  if (!g->custom_sponge && nshot >= 1) {
    // All shots through the batch engine (temporally blocked steps, every
    // shot resident at once as memory allows): each shot's image and
    // illumination are bit-identical to migrate_shot_dev's, and they are
    // stacked in shot order below, as the loop of the per-shot path does.
    std::vector<tsw_shot_desc> d(nshot);
    long long per = 0;
    for (int i = 0; i < nshot; ++i) {
      tsw_shot_desc& e = d[i];
      std::memset(&e, 0, sizeof e);
      e.nx = g->nx;
      e.nz = g->nz;
      e.nt = nts[i];
      e.vel_fwd = vel_fwd;
      e.vel_bwd = vel_bwd;
      e.src = srcs[i];
      e.six = sixs[i];
      e.siz = sizs[i];
      e.data = datas[i];
      e.rix = rix;
      e.nrec = nrec;
      e.riz = riz;
      e.kind = TSW_SHOT_ACOUSTIC;
      per = std::max(per, ((long long)(nts[i] - 1) / k + 1 + 12) * n * (long long)sizeof(float));
    }
    size_t free_b = 0, total_b = 0;
    TSW_HIP(hipMemGetInfo(&free_b, &total_b));
    const long long fit = std::max<long long>(1, (long long)(0.7 * (double)free_b) / per);
    const int slots = (int)std::min<long long>(std::min(nshot, 64), fit);
    tsw_batch_t b = nullptr;
    TSW_TRY(tsw_batch_create(nshot, d.data(), g->dx, g->dz, g->dt, g->n_sponge, g->alpha, g->fs, k,
                             slots, 100, kDefaultBlockedKernel, &b));
    int rc = tsw_batch_run(b, TSW_RUN_BATCHED);
    // a diverged shot still stacks (its NaN/Inf reach the outputs); the
    // batch's message naming that shot is reported at the end
    if (rc == kErrDiverged) {
      diverged = tsw_last_error();
      rc = kOk;
    }
    if (rc == kOk) rc = batch_stack_device(b, g->nx, g->nz, stacked.f(), total.f(), st);
    if (rc == kOk && hipStreamSynchronize(st) != hipSuccess) rc = kErrHip;
    tsw_batch_destroy(b);
    TSW_TRY(rc);
    // each shot's illumination is already scaled by dt*k inside the batch
  } else {
  launch_fill(stacked.f(), n, 0.0f, st);
  launch_fill(total.f(), n, 0.0f, st);
  for (int i = 0; i < nshot; ++i) {
    TSW_TRY(up(g->src, srcs[i], (size_t)nt_srcs[i] * sizeof(float)));
    TSW_TRY(up(g->data, datas[i], (size_t)nts[i] * nrec * sizeof(float)));
    TSW_TRY(migrate_shot_dev(g, g->src.f(), nt_srcs[i], sixs[i], sizs[i], g->data.f(), nts[i],
                             nrec, riz, k, shot_img.f(), shot_ill.f(), st));
    launch_stack_add(stacked.f(), shot_img.f(), n, st);
    launch_stack_add(total.f(), shot_ill.f(), n, st);
    // the staging buffers g->src / g->data are overwritten by the next shot
    TSW_HIP(hipStreamSynchronize(st));
  }
  }
  TSW_TRY(postprocess_dev(g, stacked.f(), total.f(), mute, out.f(), shot_img.f(),
                          static_cast<unsigned*>(maxbits.p), st));
  TSW_HIP(hipMemcpyAsync(raw_out, stacked.p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipMemcpyAsync(lap_out, shot_img.p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipMemcpyAsync(norm_out, out.p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  if (!diverged.empty()) {
    set_error("%s", diverged.c_str());
    return kErrDiverged;
  }
  const float* arr[2] = {stacked.f(), total.f()};
  const long long len[2] = {n, n};
  return check_finite(arr, len, 2, 2, 0, "stack of shots", st);
}

int tsw_laplacian(const float* image, float* out, int nx, int nz, double dx, double dz) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  TSW_ARG(image && out && nx >= 1 && nz >= 1, "bad argument");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const size_t nb = (size_t)nx * nz * sizeof(float);
  DevBuf a, b;
  TSW_TRY(a.ensure(nb));
  TSW_TRY(b.ensure(nb));
  TSW_HIP(hipMemcpyAsync(a.p, image, nb, hipMemcpyHostToDevice, st));
  launch_laplacian(a.f(), b.f(), nx, nz, dx, dz, nullptr, st);
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(out, b.p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

// ------------------------------------------- reference-compatible kernels
// Same signatures and host-pointer, synchronous semantics as wave2d.c.  They
// return void like the reference; a failure is reported by tsw_last_error().
static Grid g_scratch;  // device buffers reused across calls (guarded by g_api_mu)

static int ref_acoustic(const float* p_curr, const float* p_prev, float* p_next, const float* vel,
                        const float* sponge, int nx, int nz, float dx, float dz, float dt,
                        int fs) {
  TSW_ARG(p_curr && p_prev && p_next && vel, "null argument");
  TSW_ARG(nx >= 1 && nz >= 1, "bad grid");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const size_t nb = (size_t)nx * nz * sizeof(float);
  Grid& g = g_scratch;
  H2D up{st};
  TSW_TRY(up(g.buf[0], p_curr, nb));
  TSW_TRY(up(g.buf[1], p_prev, nb));
  TSW_TRY(up(g.buf[2], p_next, nb));
  TSW_TRY(up(g.vel, vel, nb));
  if (sponge) TSW_TRY(up(g.sponge, sponge, nb));
  TSW_TRY(g.coef[0].ensure(nb));
  launch_coef(g.vel.f(), g.coef[0].f(), (long long)nx * nz, dt * dt, st);
  AcuStep s;
  std::memset(&s, 0, sizeof s);
  s.nx = nx;
  s.nz = nz;
  s.fs = fs ? 1 : 0;
  s.rdx2 = 1.0f / (dx * dx);
  s.rdz2 = 1.0f / (dz * dz);
  s.coef = g.coef[0].f();
  s.sponge = sponge ? g.sponge.f() : nullptr;
  s.cur = g.buf[0].f();
  s.prv = g.buf[1].f();
  s.nxt = g.buf[2].f();
  s.src_cell = -1;
  s.src_it = -1;
  s.riz = -1;
  s.rdata_it = -1;
  s.rec_it = -1;
  // nx < 5 leaves no interior column, but the free-surface rows are still
  // written for every ix (wave2d.c:75-81): the step kernel's row-2 cells do
  // that.  With nz < 3 the reference's row 1 = -row 2 reads past the grid;
  // nothing is written then.
  if (nz >= 3) launch_acoustic_step(s, st);
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(p_next, g.buf[2].p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

void acoustic_2d_step(const float* p_curr, const float* p_prev, float* p_next, const float* vel,
                      const float* sponge, int nx, int nz, float dx, float dz, float dt,
                      int free_surface_top) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  g_err.clear();
  (void)ref_acoustic(p_curr, p_prev, p_next, vel, sponge, nx, nz, dx, dz, dt, free_surface_top);
}

static int ref_correlate(const float* src, const float* rec, float* image, int nx, int nz,
                         float dt) {
  TSW_ARG(src && rec && image, "null argument");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = (long long)nx * nz;
  if (n <= 0) return kOk;
  const size_t nb = (size_t)n * sizeof(float);
  Grid& g = g_scratch;
  H2D up{st};
  TSW_TRY(up(g.buf[0], src, nb));
  TSW_TRY(up(g.buf[1], rec, nb));
  TSW_TRY(up(g.image, image, nb));
  launch_correlate(g.buf[0].f(), g.buf[1].f(), g.image.f(), n, dt, st);
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(image, g.image.p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

void correlate_imaging_condition(const float* src_wavefield, const float* rec_wavefield,
                                 float* image, int nx, int nz, float dt) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  g_err.clear();
  (void)ref_correlate(src_wavefield, rec_wavefield, image, nx, nz, dt);
}

static int ref_sponge(float* sponge, int nx, int nz, int n_sponge, float alpha, int fs) {
  TSW_ARG(sponge, "null argument");
  if (nx <= 0 || nz <= 0) return kOk;
  TSW_ARG(n_sponge >= 0, "n_sponge must be >= 0");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  std::vector<float> tab;
  sponge_table(tab, n_sponge, alpha);
  Grid& g = g_scratch;
  H2D up{st};
  TSW_TRY(up(g.table, tab.data(), tab.size() * sizeof(float)));
  const size_t nb = (size_t)nx * nz * sizeof(float);
  TSW_TRY(g.sponge.ensure(nb));
  launch_sponge(g.sponge.f(), nx, nz, n_sponge, fs ? 1 : 0, g.table.f(), st);
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(sponge, g.sponge.p, nb, hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

void build_cerjan_sponge(float* sponge, int nx, int nz, int n_sponge, float alpha,
                         int free_surface_top) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  g_err.clear();
  (void)ref_sponge(sponge, nx, nz, n_sponge, alpha, free_surface_top);
}

}  // extern "C"
