// Differentiable acoustic modelling: waveform misfit and its exact discrete
// adjoint (gradient with respect to velocity), FP32 or FP64.
//
// Forward model (the reference loop, tsw_rtm/wave_solver.py:157-183 with
// tsw_rtm/c_engine/wave2d.c:19-82): with u^n the post-injection level at
// iteration n,
//   P^{n+1} = Q(Z^{n+1}),  Z on I = S.(2u^n - u^{n-1} + C.(L u^n)),
//                          Z on E = u^{n-2}     (cells never written: the
//                                                buffer three levels back)
//   u^{n+1} = P^{n+1} + b_{n+1},  d_n = R P^{n+1},
// where I = [2,nz-2) x [2,nx-2), E = the other cells except the rows a free
// surface rewrites, Q sets row 0 to 0 and row 1 to -row 2 (free surface),
// C = (v*v)*dt^2, S the sponge and L the 4th-order Laplacian.  The misfit is
// J = 1/2 sum_n |d_n - d_obs_n|^2.
//
// Adjoint (exact transpose, backward in time): with mu^m = dJ/du^m,
//   zeta = Q^T mu^{n+1},  w = S.zeta on I (0 elsewhere),
//   mu^n     += 2 w + L^T (C.w) + R^T r_{n-1}
//   mu^{n-1} -= w
//   mu^{n-2}  = zeta on E (first contribution, assigned)
//   g_C      += w . (L u^n) on I,      dJ/dv = g_C . 2 v dt^2.
// L is symmetric (central stencil), so L^T is the same stencil applied to the
// zero-extended field.  Four adjoint buffers rotate with period 4.
//
// The forward levels u^n are not stored: every ceil(sqrt(nt)) steps the three
// levels (u^{n-2}, u^{n-1}, u^n) are checkpointed, and the backward sweep
// recomputes each segment from its checkpoint (bit-identical to the first
// pass) before running that segment's adjoint steps.
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "tswrtm.h"

namespace tsw {

template <typename T> struct FwiK {
  static constexpr T c0 = T(-2.5);
};
template <> struct FwiK<float> {
  static constexpr float c0 = -2.5f;
};
template <typename T> __host__ __device__ inline T fwi_c1() { return T(4) / T(3); }
template <typename T> __host__ __device__ inline T fwi_c2() { return T(-1) / T(12); }
template <> __host__ __device__ inline float fwi_c1<float>() { return 4.0f / 3.0f; }
template <> __host__ __device__ inline float fwi_c2<float>() { return -1.0f / 12.0f; }

template <typename T>
struct FwdArgs {
  int nx, nz, fs;
  T rdx2, rdz2, inj;
  const T* C;
  const T* S;
  const T* cur;
  const T* prv;
  T* nxt;
  int src_cell, src_it;
  const T* src;
  int riz, nrec;
  const int* roff;
  const int* ridx;
  T* rec;
  int rec_it;
  T* save;  // non-null: save = cur (u^n)
};

template <typename T>
__device__ __forceinline__ T lap4(const T* p, int c, int nx, T rdx2, T rdz2) {
  const T c0 = FwiK<T>::c0, c1 = fwi_c1<T>(), c2 = fwi_c2<T>();
  const T pc = p[c];
  T lx = c2 * (p[c - 2] + p[c + 2]);
  lx = lx + c1 * (p[c - 1] + p[c + 1]);
  lx = (lx + c0 * pc) * rdx2;
  T lz = c2 * (p[c - 2 * nx] + p[c + 2 * nx]);
  lz = lz + c1 * (p[c - nx] + p[c + nx]);
  lz = (lz + c0 * pc) * rdz2;
  return lx + lz;
}

template <typename T>
__device__ __forceinline__ void fwd_io(const FwdArgs<T>& s, int iz, int ix, T& v, bool& touched) {
  if (iz == s.riz && s.roff) {
    const int a = s.roff[ix], b = s.roff[ix + 1];
    if (s.rec_it >= 0)
      for (int r = a; r < b; ++r) s.rec[(long long)s.rec_it * s.nrec + s.ridx[r]] = v;
  }
  if (s.src_it >= 0 && iz * s.nx + ix == s.src_cell) {
    v = v + s.src[s.src_it] * s.inj;
    touched = true;
  }
}

// One forward step (same operation order as the FP32 production kernel).
template <typename T>
__global__ __launch_bounds__(256) void k_fwi_fwd(FwdArgs<T> s) {
  const int ix = blockIdx.x * 64 + threadIdx.x;
  const int iz = blockIdx.y * 4 + threadIdx.y;
  if (ix >= s.nx || iz >= s.nz) return;
  const int nx = s.nx, c = iz * nx + ix;
  if (s.save) s.save[c] = s.cur[c];
  if (s.fs && iz < 2) return;
  const bool interior = iz >= 2 && iz < s.nz - 2 && ix >= 2 && ix < nx - 2;
  T v;
  if (interior) {
    const T pc = s.cur[c];
    v = (T(2) * pc - s.prv[c]) + s.C[c] * lap4(s.cur, c, nx, s.rdx2, s.rdz2);
    v = v * s.S[c];
  } else {
    v = s.nxt[c];
  }
  bool t0 = false, t1 = false, t2 = false;
  if (s.fs && iz == 2) {
    T v0 = T(0);
    fwd_io(s, 0, ix, v0, t0);
    s.nxt[ix] = v0;
    T v1 = -v;
    fwd_io(s, 1, ix, v1, t1);
    s.nxt[nx + ix] = v1;
  }
  fwd_io(s, iz, ix, v, t2);
  if (interior || t2) s.nxt[c] = v;
}

// iteration-0 injection: buf[cell] += src[0] * inj
template <typename T>
__global__ void k_fwi_inject0(T* buf, int cell, const T* src, T inj) {
  if (threadIdx.x == 0 && blockIdx.x == 0) buf[cell] = buf[cell] + src[0] * inj;
}

template <typename T>
struct AdjArgs {
  int nx, nz, fs;
  T rdx2, rdz2;
  const T* C;
  const T* S;
  const T* mu1;  // mu^{n+1}, complete
  T* mu0;        // mu^n
  T* mum1;       // mu^{n-1}
  T* mum2;       // mu^{n-2} (assigned)
  const T* un;   // forward u^n
  T* grad;       // g_C
  int riz, nrec;
  const int* roff;
  const int* ridx;
  const T* resid;  // (nt, nrec) residual d - d_obs
  int res_it;      // residual row added to mu^n (n-1), -1: none
  T* probe;        // non-null: probe[probe_it] = dJ/du^n at probe_cell
  int probe_cell, probe_it;
};

template <typename T>
__device__ __forceinline__ bool fwi_interior(int iz, int ix, int nz, int nx) {
  return iz >= 2 && iz < nz - 2 && ix >= 2 && ix < nx - 2;
}

// zeta = Q^T mu^{n+1} at (iz, ix)
template <typename T>
__device__ __forceinline__ T fwi_zeta(const AdjArgs<T>& a, int iz, int ix) {
  const int c = iz * a.nx + ix;
  if (!a.fs) return a.mu1[c];
  if (iz < 2) return T(0);
  if (iz == 2) return a.mu1[c] - a.mu1[c - a.nx];
  return a.mu1[c];
}

// C.w at (iz, ix), zero outside the grid and outside I
template <typename T>
__device__ __forceinline__ T fwi_cw(const AdjArgs<T>& a, int iz, int ix) {
  if (iz < 0 || iz >= a.nz || ix < 0 || ix >= a.nx) return T(0);
  if (!fwi_interior<T>(iz, ix, a.nz, a.nx)) return T(0);
  const int c = iz * a.nx + ix;
  return a.C[c] * (a.S[c] * fwi_zeta(a, iz, ix));
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwi_adj(AdjArgs<T> a) {
  const int ix = blockIdx.x * 64 + threadIdx.x;
  const int iz = blockIdx.y * 4 + threadIdx.y;
  if (ix >= a.nx || iz >= a.nz) return;
  const int nx = a.nx, c = iz * nx + ix;
  const T c0 = FwiK<T>::c0, c1 = fwi_c1<T>(), c2 = fwi_c2<T>();
  const bool in_i = fwi_interior<T>(iz, ix, a.nz, nx);
  const T zeta = fwi_zeta(a, iz, ix);
  const T w = in_i ? a.S[c] * zeta : T(0);
  // L^T (C.w) at c
  T lx = c2 * (fwi_cw(a, iz, ix - 2) + fwi_cw(a, iz, ix + 2));
  lx = lx + c1 * (fwi_cw(a, iz, ix - 1) + fwi_cw(a, iz, ix + 1));
  const T cwc = fwi_cw(a, iz, ix);
  lx = (lx + c0 * cwc) * a.rdx2;
  T lz = c2 * (fwi_cw(a, iz - 2, ix) + fwi_cw(a, iz + 2, ix));
  lz = lz + c1 * (fwi_cw(a, iz - 1, ix) + fwi_cw(a, iz + 1, ix));
  lz = (lz + c0 * cwc) * a.rdz2;
  T m0 = a.mu0[c] + (T(2) * w + (lx + lz));
  // probe: dJ/du^n at the source cell (complete here; the receiver term
  // below belongs to dJ/dP^n, not to the injected u^n)
  if (a.probe && c == a.probe_cell) a.probe[a.probe_it] = m0;
  if (a.res_it >= 0 && iz == a.riz) {
    const int r0 = a.roff[ix], r1 = a.roff[ix + 1];
    const T* row = a.resid + (long long)a.res_it * a.nrec;
    for (int r = r0; r < r1; ++r) m0 = m0 + row[a.ridx[r]];
  }
  a.mu0[c] = m0;
  a.mum1[c] = a.mum1[c] - w;
  const bool kept = !in_i && !(a.fs && iz < 2);
  a.mum2[c] = kept ? zeta : T(0);
  if (in_i) a.grad[c] = a.grad[c] + w * lap4(a.un, c, nx, a.rdx2, a.rdz2);
}

template <typename T>
__global__ void k_fwi_fill(T* p, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// mu^{nt} = R^T r_{nt-1}
template <typename T>
__global__ void k_fwi_seed(T* mu, int nx, int riz, const int* roff, const int* ridx,
                           const T* row) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= nx) return;
  const int r0 = roff[ix], r1 = roff[ix + 1];
  if (r0 == r1) return;
  T m = mu[riz * nx + ix];
  for (int r = r0; r < r1; ++r) m = m + row[ridx[r]];
  mu[riz * nx + ix] = m;
}

template <typename T>
__global__ void k_fwi_coef(const T* v, T* C, long long n, T dt2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    C[i] = (v[i] * v[i]) * dt2;
}

// dJ/dv = g_C . 2 v dt^2, written to out (accumulated when acc != 0)
template <typename T>
__global__ void k_fwi_chain(const T* g, const T* v, T* out, long long n, T dt2, int acc) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const T d = g[i] * ((T(2) * v[i]) * dt2);
    out[i] = acc ? out[i] + d : d;
  }
}

template <typename T>
__global__ void k_fwi_resid(const T* pred, const T* obs, T* res, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    res[i] = pred[i] - obs[i];
}

struct FwiShotHost {
  int nt, six, siz, riz, nrec;
  DevBuf src, obs, pred, res, roff, ridx, ckpt, probe;
  std::vector<double> host_res;
};

struct Fwi {
  int nx = 0, nz = 0, fs = 1, prec = 4, kint = 1, n_sponge = 0;
  double dx = 0, dz = 0, dt = 0, alpha = 0;
  std::vector<FwiShotHost> shots;
  DevBuf S, C, buf[3], mu[4], seg, gC, tmp;
  double loss = 0;
  bool have_forward = false;
  size_t esz() const { return prec == 8 ? 8 : 4; }
  long long cells() const { return (long long)nx * nz; }
};

static unsigned g1(long long n) {
  long long b = (n + 255) / 256;
  return (unsigned)std::max<long long>(1, std::min<long long>(b, 65536));
}

template <typename T>
static int fwi_forward_t(Fwi* F, const T* d_vel, hipStream_t st, bool store_all) {
  const long long n = F->cells();
  const T dtT = (T)F->dt;
  const T dt2 = dtT * dtT;
  hipLaunchKernelGGL(k_fwi_coef<T>, dim3(g1(n)), dim3(256), 0, st, d_vel, (T*)F->C.p, n, dt2);
  const T dxT = (T)F->dx, dzT = (T)F->dz;
  F->loss = 0;
  for (auto& sh : F->shots) {
    const int nt = sh.nt;
    for (auto& b : F->buf)
      hipLaunchKernelGGL(k_fwi_fill<T>, dim3(g1(n)), dim3(256), 0, st, (T*)b.p, n, T(0));
    FwdArgs<T> s;
    std::memset(&s, 0, sizeof s);
    s.nx = F->nx;
    s.nz = F->nz;
    s.fs = F->fs;
    s.rdx2 = T(1) / (dxT * dxT);
    s.rdz2 = T(1) / (dzT * dzT);
    s.inj = (T)(F->dt * F->dt);
    s.C = (const T*)F->C.p;
    s.S = (const T*)F->S.p;
    s.src_cell = sh.siz * F->nx + sh.six;
    s.src = (const T*)sh.src.p;
    s.riz = sh.riz;
    s.nrec = sh.nrec;
    s.roff = sh.roff.i();
    s.ridx = sh.ridx.i();
    s.rec = (T*)sh.pred.p;
    hipLaunchKernelGGL(k_fwi_inject0<T>, dim3(1), dim3(64), 0, st, (T*)F->buf[1].p, s.src_cell,
                       (const T*)sh.src.p, s.inj);
    const int K = F->kint;
    for (int it = 0; it < nt; ++it) {
      T* prv = (T*)F->buf[it % 3].p;
      T* cur = (T*)F->buf[(it + 1) % 3].p;
      T* nxt = (T*)F->buf[(it + 2) % 3].p;
      s.save = nullptr;
      if (!store_all && it % K == 0) {  // checkpoint (u^{n-2}, u^{n-1}, u^n)
        T* ck = (T*)sh.ckpt.p + (long long)(it / K) * 3 * n;
        TSW_HIP(hipMemcpyAsync(ck, nxt, n * sizeof(T), hipMemcpyDeviceToDevice, st));
        TSW_HIP(hipMemcpyAsync(ck + n, prv, n * sizeof(T), hipMemcpyDeviceToDevice, st));
        TSW_HIP(hipMemcpyAsync(ck + 2 * n, cur, n * sizeof(T), hipMemcpyDeviceToDevice, st));
      }
      s.prv = prv;
      s.cur = cur;
      s.nxt = nxt;
      s.src_it = it + 1 < nt ? it + 1 : -1;
      s.rec_it = it;
      hipLaunchKernelGGL(k_fwi_fwd<T>, dim3((F->nx + 63) / 64, (F->nz + 3) / 4), dim3(64, 4), 0,
                         st, s);
    }
    const long long nr = (long long)nt * sh.nrec;
    hipLaunchKernelGGL(k_fwi_resid<T>, dim3(g1(nr)), dim3(256), 0, st, (const T*)sh.pred.p,
                       (const T*)sh.obs.p, (T*)sh.res.p, nr);
    TSW_HIP(hipGetLastError());
    std::vector<T> h(nr);
    TSW_HIP(hipMemcpyAsync(h.data(), sh.res.p, nr * sizeof(T), hipMemcpyDeviceToHost, st));
    TSW_HIP(hipStreamSynchronize(st));
    double acc = 0;
    for (const T& x : h) acc += (double)x * (double)x;
    F->loss += 0.5 * acc;
  }
  F->have_forward = true;
  return kOk;
}

template <typename T>
static int fwi_backward_t(Fwi* F, const T* d_vel, T* d_grad, hipStream_t st, bool store_all) {
  const long long n = F->cells();
  const T dtT = (T)F->dt;
  const T dt2 = dtT * dtT;
  const T dxT = (T)F->dx, dzT = (T)F->dz;
  hipLaunchKernelGGL(k_fwi_fill<T>, dim3(g1(n)), dim3(256), 0, st, (T*)F->gC.p, n, T(0));
  for (auto& sh : F->shots) {
    const int nt = sh.nt;
    for (auto& m : F->mu)
      hipLaunchKernelGGL(k_fwi_fill<T>, dim3(g1(n)), dim3(256), 0, st, (T*)m.p, n, T(0));
    // mu^{nt} lives in buffer nt % 4
    hipLaunchKernelGGL(k_fwi_seed<T>, dim3((F->nx + 255) / 256), dim3(256), 0, st,
                       (T*)F->mu[nt % 4].p, F->nx, sh.riz, sh.roff.i(), sh.ridx.i(),
                       (const T*)sh.res.p + (long long)(nt - 1) * sh.nrec);
    AdjArgs<T> a;
    std::memset(&a, 0, sizeof a);
    a.nx = F->nx;
    a.nz = F->nz;
    a.fs = F->fs;
    a.rdx2 = T(1) / (dxT * dxT);
    a.rdz2 = T(1) / (dzT * dzT);
    a.C = (const T*)F->C.p;
    a.S = (const T*)F->S.p;
    a.grad = (T*)F->gC.p;
    a.riz = sh.riz;
    a.nrec = sh.nrec;
    a.roff = sh.roff.i();
    a.ridx = sh.ridx.i();
    a.resid = (const T*)sh.res.p;
    a.probe = sh.probe.p ? (T*)sh.probe.p : nullptr;
    a.probe_cell = sh.siz * F->nx + sh.six;
    // forward recompute machinery
    FwdArgs<T> s;
    std::memset(&s, 0, sizeof s);
    s.nx = F->nx;
    s.nz = F->nz;
    s.fs = F->fs;
    s.rdx2 = a.rdx2;
    s.rdz2 = a.rdz2;
    s.inj = (T)(F->dt * F->dt);
    s.C = a.C;
    s.S = a.S;
    s.src_cell = a.probe_cell;
    s.src = (const T*)sh.src.p;
    s.riz = -1;
    s.rec_it = -1;
    const int K = store_all ? nt : F->kint;
    const int nseg = (nt + K - 1) / K;
    for (int sg = nseg - 1; sg >= 0; --sg) {
      const int n0 = sg * K, n1 = std::min(nt, n0 + K);
      {
        if (store_all) {
          // one segment: the whole forward again from the zero initial state
          for (auto& b : F->buf)
            hipLaunchKernelGGL(k_fwi_fill<T>, dim3(g1(n)), dim3(256), 0, st, (T*)b.p, n, T(0));
          hipLaunchKernelGGL(k_fwi_inject0<T>, dim3(1), dim3(64), 0, st, (T*)F->buf[1].p,
                             s.src_cell, (const T*)sh.src.p, s.inj);
        } else {
          // restore (u^{n0-2}, u^{n0-1}, u^{n0}) into the buffers the step at n0 uses
          const T* ck = (const T*)sh.ckpt.p + (long long)sg * 3 * n;
          TSW_HIP(hipMemcpyAsync(F->buf[(n0 + 2) % 3].p, ck, n * sizeof(T),
                                 hipMemcpyDeviceToDevice, st));
          TSW_HIP(hipMemcpyAsync(F->buf[n0 % 3].p, ck + n, n * sizeof(T),
                                 hipMemcpyDeviceToDevice, st));
          TSW_HIP(hipMemcpyAsync(F->buf[(n0 + 1) % 3].p, ck + 2 * n, n * sizeof(T),
                                 hipMemcpyDeviceToDevice, st));
        }
        for (int it = n0; it < n1; ++it) {
          s.prv = (T*)F->buf[it % 3].p;
          s.cur = (T*)F->buf[(it + 1) % 3].p;
          s.nxt = (T*)F->buf[(it + 2) % 3].p;
          s.src_it = it + 1 < nt ? it + 1 : -1;
          s.save = (T*)F->seg.p + (long long)(it - n0) * n;
          hipLaunchKernelGGL(k_fwi_fwd<T>, dim3((F->nx + 63) / 64, (F->nz + 3) / 4),
                             dim3(64, 4), 0, st, s);
        }
      }
      for (int it = n1 - 1; it >= n0; --it) {
        a.mu1 = (const T*)F->mu[(it + 1) % 4].p;
        a.mu0 = (T*)F->mu[it % 4].p;
        a.mum1 = (T*)F->mu[(it + 3) % 4].p;
        a.mum2 = (T*)F->mu[(it + 2) % 4].p;
        a.un = (const T*)F->seg.p + (long long)(it - n0) * n;
        a.res_it = it >= 1 ? it - 1 : -1;
        a.probe_it = it;
        hipLaunchKernelGGL(k_fwi_adj<T>, dim3((F->nx + 63) / 64, (F->nz + 3) / 4), dim3(64, 4),
                           0, st, a);
      }
    }
    TSW_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fwi_chain<T>, dim3(g1(n)), dim3(256), 0, st, (const T*)F->gC.p, d_vel,
                     d_grad, n, dt2, 0);
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

}  // namespace tsw

using namespace tsw;

extern "C" {

int tsw_fwi_create(int nx, int nz, double dx, double dz, double dt, int n_sponge,
                   double sponge_alpha, int free_surface_top, int precision, int nshot,
                   const tsw_fwi_shot* shots, int ckpt_interval, tsw_fwi_t* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(out && shots && nshot >= 1, "need at least one shot");
  TSW_ARG(precision == 4 || precision == 8, "precision must be 4 (FP32) or 8 (FP64)");
  TSW_ARG(nx >= 5 && nz >= 5 && dx > 0 && dz > 0 && dt > 0 && n_sponge >= 0, "bad grid");
  *out = nullptr;
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  auto* F = new Fwi();
  auto fail = [&](int rc) {
    delete F;
    return rc;
  };
  F->nx = nx;
  F->nz = nz;
  F->dx = dx;
  F->dz = dz;
  F->dt = dt;
  F->fs = free_surface_top ? 1 : 0;
  F->prec = precision;
  F->n_sponge = n_sponge;
  F->alpha = sponge_alpha;
  const long long n = F->cells();
  const size_t es = F->esz();
  int ntmax = 1;
  for (int i = 0; i < nshot; ++i) {
    const tsw_fwi_shot& d = shots[i];
    if (!(d.nt >= 1 && d.src && d.six >= 0 && d.six < nx && d.siz >= 0 && d.siz < nz &&
          d.riz >= 0 && d.riz < nz && d.nrec >= 1 && d.rix && d.d_obs))
      return fail((set_error("shot %d: bad arguments", i), kErrArg));
    ntmax = std::max(ntmax, d.nt);
  }
  const bool store_all = ckpt_interval < 0;
  F->kint = ckpt_interval > 0 ? ckpt_interval
                              : std::max(1, (int)std::ceil(std::sqrt((double)ntmax)));
  int rc;
  // sponge in the working precision: expf table (FP32) or exp (FP64)
  {
    std::vector<float> tf;
    sponge_table(tf, n_sponge, (float)sponge_alpha);
    std::vector<unsigned char> host(n * es);
    for (int iz = 0; iz < nz; ++iz)
      for (int ix = 0; ix < nx; ++ix) {
        const int d = sponge_depth(ix, iz, nx, nz, n_sponge, F->fs);
        const size_t c = (size_t)iz * nx + ix;
        if (es == 4) ((float*)host.data())[c] = d > 0 ? tf[d] : 1.0f;
        else ((double*)host.data())[c] = d > 0 ? std::exp(-sponge_alpha * d * d) : 1.0;
      }
    if ((rc = F->S.ensure(n * es))) return fail(rc);
    if (hipMemcpy(F->S.p, host.data(), n * es, hipMemcpyHostToDevice))
      return fail((set_error("sponge upload failed"), kErrHip));
  }
  if ((rc = F->C.ensure(n * es)) || (rc = F->gC.ensure(n * es))) return fail(rc);
  for (auto& b : F->buf)
    if ((rc = b.ensure(n * es))) return fail(rc);
  for (auto& m : F->mu)
    if ((rc = m.ensure(n * es))) return fail(rc);
  if ((rc = F->seg.ensure((size_t)(store_all ? ntmax : F->kint) * n * es))) return fail(rc);
  F->shots.resize(nshot);
  for (int i = 0; i < nshot; ++i) {
    const tsw_fwi_shot& d = shots[i];
    FwiShotHost& h = F->shots[i];
    h.nt = d.nt;
    h.six = d.six;
    h.siz = d.siz;
    h.riz = d.riz;
    h.nrec = d.nrec;
    RecvMap m;
    if ((rc = m.build(d.rix, d.nrec, nx))) return fail(rc);
    const size_t nr = (size_t)d.nt * d.nrec * es;
    if ((rc = h.src.ensure(d.nt * es)) || (rc = h.obs.ensure(nr)) || (rc = h.pred.ensure(nr)) ||
        (rc = h.res.ensure(nr)) || (rc = h.roff.ensure(m.off.size() * sizeof(int))) ||
        (rc = h.ridx.ensure(m.idx.size() * sizeof(int))) || (rc = h.probe.ensure(d.nt * es)))
      return fail(rc);
    if (!store_all) {
      const int nck = (d.nt + F->kint - 1) / F->kint;
      if ((rc = h.ckpt.ensure((size_t)nck * 3 * n * es))) return fail(rc);
    }
    if (hipMemcpy(h.src.p, d.src, d.nt * es, hipMemcpyHostToDevice) ||
        hipMemcpy(h.obs.p, d.d_obs, nr, hipMemcpyHostToDevice) ||
        hipMemcpy(h.roff.p, m.off.data(), m.off.size() * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(h.ridx.p, m.idx.data(), m.idx.size() * sizeof(int), hipMemcpyHostToDevice))
      return fail((set_error("shot %d upload failed", i), kErrHip));
  }
  F->kint = store_all ? -1 : F->kint;
  *out = reinterpret_cast<tsw_fwi_t>(F);
  return kOk;
}

int tsw_fwi_forward(tsw_fwi_t h, const void* d_vel, void* stream, double* loss_out) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F && d_vel, "null argument");
  TSW_TRY(fail_injected());
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool store_all = F->kint < 0;
  const int rc = F->prec == 8 ? fwi_forward_t<double>(F, (const double*)d_vel, st, store_all)
                              : fwi_forward_t<float>(F, (const float*)d_vel, st, store_all);
  if (rc == kOk && loss_out) *loss_out = F->loss;
  return rc;
}

int tsw_fwi_backward(tsw_fwi_t h, const void* d_vel, void* d_grad, void* stream) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F && d_vel && d_grad, "null argument");
  TSW_ARG(F->have_forward, "tsw_fwi_forward must run before tsw_fwi_backward");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool store_all = F->kint < 0;
  return F->prec == 8 ? fwi_backward_t<double>(F, (const double*)d_vel, (double*)d_grad, st,
                                               store_all)
                      : fwi_backward_t<float>(F, (const float*)d_vel, (float*)d_grad, st,
                                              store_all);
}

int tsw_fwi_predicted(tsw_fwi_t h, int shot, void* host_out) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F && host_out && shot >= 0 && shot < (int)F->shots.size(), "bad argument");
  const auto& s = F->shots[shot];
  TSW_HIP(hipDeviceSynchronize());
  TSW_HIP(hipMemcpy(host_out, s.pred.p, (size_t)s.nt * s.nrec * F->esz(), hipMemcpyDeviceToHost));
  return kOk;
}

int tsw_fwi_set_residual(tsw_fwi_t h, int shot, const void* host_res) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F && host_res && shot >= 0 && shot < (int)F->shots.size(), "bad argument");
  const auto& s = F->shots[shot];
  TSW_HIP(hipDeviceSynchronize());
  TSW_HIP(hipMemcpy(s.res.p, host_res, (size_t)s.nt * s.nrec * F->esz(), hipMemcpyHostToDevice));
  return kOk;
}

int tsw_fwi_probe(tsw_fwi_t h, int shot, void* host_out) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F && host_out && shot >= 0 && shot < (int)F->shots.size(), "bad argument");
  const auto& s = F->shots[shot];
  TSW_HIP(hipDeviceSynchronize());
  TSW_HIP(hipMemcpy(host_out, s.probe.p, (size_t)s.nt * F->esz(), hipMemcpyDeviceToHost));
  return kOk;
}

int tsw_fwi_destroy(tsw_fwi_t h) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  (void)hipDeviceSynchronize();
  delete reinterpret_cast<Fwi*>(h);
  return kOk;
}

}  // extern "C"
