// CPU oracle for the tsw_rtm reverse-time-migration path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine in
// das_amd/ is compared against, and the timed CPU baseline.  The product never
// links, loads or calls it.
//
// It restates, in FP32 and FP64 (one template), the reference algorithm:
//   - acoustic_2d_step        tsw_rtm/c_engine/wave2d.c:19-82
//   - elastic_2d_step         tsw_rtm/c_engine/wave2d.c:95-212
//   - correlate_imaging_cond. tsw_rtm/c_engine/wave2d.c:218-231
//   - build_cerjan_sponge     tsw_rtm/c_engine/wave2d.c:237-265
//   - forward time loop       tsw_rtm/wave_solver.py:128-186
//   - backward time loop      tsw_rtm/wave_solver.py:188-242
//   - per-shot illumination   tsw_rtm/rtm_engine.py:89
//   - Laplacian post-filter   tsw_rtm/rtm_engine.py:27-38 (scipy convolve2d,
//                             mode='same', boundary='symm')
//
// Floating-point association follows the reference expression by expression
// (C evaluates a+b+c as (a+b)+c); the file is built with -ffp-contract=off so
// that the FP32 instantiation is bit-identical to wave2d.c compiled by gcc for
// x86-64 (which has no FMA contraction without -mfma).
//
// Deviation, documented: the elastic step with a free surface reads one row
// above the grid (wave2d.c:135 and :170 read index idx-2*nx at iz=1), which is
// an out-of-bounds read in the reference.  Here that row is defined as 0.
//
// Pinning: tests/test_oracle.py checks every function below against golden
// vectors produced by wave2d.c compiled from source (oracle/Makefile) and by
// numpy/scipy expressions of the reference's Python layer, plus invariants.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <typename T> struct K4 {  // 4th-order second-derivative weights (wave2d.c:37-39)
  static constexpr T c0 = T(-2.5);
};
template <> struct K4<float> {
  static constexpr float c0 = -2.5f;
};

template <typename T> inline T lap_c1();
template <> inline float lap_c1<float>() { return 4.0f / 3.0f; }
template <> inline double lap_c1<double>() { return 4.0 / 3.0; }
template <typename T> inline T lap_c2();
template <> inline float lap_c2<float>() { return -1.0f / 12.0f; }
template <> inline double lap_c2<double>() { return -1.0 / 12.0; }
// staggered first-derivative weights (wave2d.c:117-118)
template <typename T> inline T stg_c1();
template <> inline float stg_c1<float>() { return 9.0f / 8.0f; }
template <> inline double stg_c1<double>() { return 9.0 / 8.0; }
template <typename T> inline T stg_c2();
template <> inline float stg_c2<float>() { return -1.0f / 24.0f; }
template <> inline double stg_c2<double>() { return -1.0 / 24.0; }

// One leapfrog step of the 4th-order acoustic scheme on rows [2, nz-2) x
// cols [2, nx-2); other cells of `nxt` are left untouched, except rows 0/1
// under a free surface (zero / antisymmetric mirror of row 2).
template <typename T>
void acoustic_step(const T* cur, const T* prv, T* nxt, const T* vel, const T* sponge,
                   int nx, int nz, T dx, T dz, T dt, int free_surface) {
  const T inv_dx2 = T(1) / (dx * dx);
  const T inv_dz2 = T(1) / (dz * dz);
  const T dt2 = dt * dt;
  const T c0 = K4<T>::c0, c1 = lap_c1<T>(), c2 = lap_c2<T>();
  const long long s = nx;
#pragma omp parallel for schedule(static)
  for (int iz = 2; iz < nz - 2; ++iz) {
    for (int ix = 2; ix < nx - 2; ++ix) {
      const long long i = (long long)iz * s + ix;
      const T pc = cur[i];
      T lx = c2 * (cur[i - 2] + cur[i + 2]);
      lx = lx + c1 * (cur[i - 1] + cur[i + 1]);
      lx = (lx + c0 * pc) * inv_dx2;
      T lz = c2 * (cur[i - 2 * s] + cur[i + 2 * s]);
      lz = lz + c1 * (cur[i - s] + cur[i + s]);
      lz = (lz + c0 * pc) * inv_dz2;
      const T vv = vel[i];
      const T coef = (vv * vv) * dt2;
      T out = (T(2) * pc - prv[i]) + coef * (lx + lz);
      if (sponge) out = out * sponge[i];
      nxt[i] = out;
    }
  }
  if (free_surface) {
    for (int ix = 0; ix < nx; ++ix) {
      nxt[ix] = T(0);
      nxt[s + ix] = -nxt[2 * s + ix];
    }
  }
}

// Staggered-grid P-SV velocity-stress step (Levander).  Row -1 reads as 0.
template <typename T>
void elastic_step(T* vx, T* vz, T* sxx, T* szz, T* sxz, const T* lam, const T* mu,
                  const T* bx, const T* bz, const T* sponge, int nx, int nz, T dx, T dz,
                  T dt, int free_surface) {
  const T rdx = T(1) / dx, rdz = T(1) / dz;
  const T a = stg_c1<T>(), b = stg_c2<T>();
  const long long s = nx;
  const int z0 = free_surface ? 1 : 2;
  // value of f at flat index i, with rows above the grid reading 0
  auto at = [](const T* f, long long i) -> T { return i < 0 ? T(0) : f[i]; };
#pragma omp parallel for schedule(static)
  for (int iz = z0; iz < nz - 2; ++iz) {
    for (int ix = 2; ix < nx - 2; ++ix) {
      const long long i = (long long)iz * s + ix;
      const T dsxx = (a * (sxx[i + 1] - sxx[i]) + b * (sxx[i + 2] - sxx[i - 1])) * rdx;
      const T dsxz_z = (a * (sxz[i] - sxz[i - s]) + b * (sxz[i + s] - at(sxz, i - 2 * s))) * rdz;
      vx[i] = vx[i] + (dt * bx[i]) * (dsxx + dsxz_z);
      const T dsxz_x = (a * (sxz[i] - sxz[i - 1]) + b * (sxz[i + 1] - sxz[i - 2])) * rdx;
      const T dszz = (a * (szz[i + s] - szz[i]) + b * (szz[i + 2 * s] - szz[i - s])) * rdz;
      vz[i] = vz[i] + (dt * bz[i]) * (dsxz_x + dszz);
      if (sponge) {
        const T g = sponge[i];
        vx[i] = vx[i] * g;
        vz[i] = vz[i] * g;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int iz = z0; iz < nz - 2; ++iz) {
    for (int ix = 2; ix < nx - 2; ++ix) {
      const long long i = (long long)iz * s + ix;
      const T dvx_x = (a * (vx[i] - vx[i - 1]) + b * (vx[i + 1] - vx[i - 2])) * rdx;
      const T dvz_z = (a * (vz[i] - vz[i - s]) + b * (vz[i + s] - at(vz, i - 2 * s))) * rdz;
      const T l = lam[i], m = mu[i];
      const T l2m = l + T(2) * m;
      sxx[i] = sxx[i] + dt * (l2m * dvx_x + l * dvz_z);
      szz[i] = szz[i] + dt * (l * dvx_x + l2m * dvz_z);
      const T dvz_x = (a * (vz[i + 1] - vz[i]) + b * (vz[i + 2] - vz[i - 1])) * rdx;
      const T dvx_z = (a * (vx[i + s] - vx[i]) + b * (vx[i + 2 * s] - vx[i - s])) * rdz;
      sxz[i] = sxz[i] + (dt * m) * (dvz_x + dvx_z);
      if (sponge) {
        const T g = sponge[i];
        sxx[i] = sxx[i] * g;
        szz[i] = szz[i] * g;
        sxz[i] = sxz[i] * g;
      }
    }
  }
  if (free_surface) {
    for (int ix = 0; ix < nx; ++ix) {
      szz[ix] = T(0);
      sxz[ix] = T(0);
      if (ix >= 2 && ix < nx - 2) {
        const T l = lam[ix], m = mu[ix];
        const T d = (a * (vx[ix] - vx[ix - 1]) + b * (vx[ix + 1] - vx[ix - 2])) * rdx;
        // assignment without dt: reproduces wave2d.c:208
        sxx[ix] = (((T(4) * m) * (l + m)) / ((l + T(2) * m) + T(1e-8f))) * d;
      }
    }
  }
}

template <typename T>
void correlate(const T* src, const T* rec, T* image, long long n, T w) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < n; ++i) image[i] = image[i] + (src[i] * rec[i]) * w;
}

// Distance into the absorbing strip, as an integer count of cells
// (wave2d.c:247-256).  0 inside the interior.
inline int sponge_depth(int ix, int iz, int nx, int nz, int n, int fs) {
  int d = 0;
  d = std::max(d, n - ix);
  d = std::max(d, ix - (nx - 1 - n));
  d = std::max(d, iz - (nz - 1 - n));
  if (!fs) d = std::max(d, n - iz);
  return d;
}

// Function-pointer types of the reference's own kernels (wave2d.c signatures),
// so the loops below can also drive wave2d.c compiled from source.
typedef void (*ref_acoustic_fn)(const float*, const float*, float*, const float*, const float*,
                                int, int, float, float, float, int);
typedef void (*ref_correlate_fn)(const float*, const float*, float*, int, int, float);

template <typename T> struct Stepper {
  ref_acoustic_fn ext_step = nullptr;
  ref_correlate_fn ext_corr = nullptr;
  void step(const T* c, const T* p, T* n, const T* v, const T* sp, int nx, int nz, T dx, T dz,
            T dt, int fs) const {
    acoustic_step<T>(c, p, n, v, sp, nx, nz, dx, dz, dt, fs);
  }
  void corr(const T* a, const T* b, T* img, int nx, int nz, T w) const {
    correlate<T>(a, b, img, (long long)nx * nz, w);
  }
};
template <> void Stepper<float>::step(const float* c, const float* p, float* n, const float* v,
                                      const float* sp, int nx, int nz, float dx, float dz,
                                      float dt, int fs) const {
  if (ext_step) ext_step(c, p, n, v, sp, nx, nz, dx, dz, dt, fs);
  else acoustic_step<float>(c, p, n, v, sp, nx, nz, dx, dz, dt, fs);
}
template <> void Stepper<float>::corr(const float* a, const float* b, float* img, int nx, int nz,
                                      float w) const {
  if (ext_corr) ext_corr(a, b, img, nx, nz, w);
  else correlate<float>(a, b, img, (long long)nx * nz, w);
}

// Forward modelling loop of wave_solver.py:146-186.  Injection scale is
// T(dt_host*dt_host) (numpy weak-scalar promotion of `source_ts[it] * dt**2`).
// snaps (optional) receives p_curr (post-injection, pre-step) every k steps.
template <typename T>
void forward_loop(const Stepper<T>& st, const T* vel, const T* sponge, int nx, int nz, T dx, T dz,
                  double dt_host, int fs, const T* src, int nt, int six, int siz, const int* rix,
                  int nrec, int riz, int k, T* rec_out, T* snaps_out) {
  const size_t n = (size_t)nx * nz;
  std::vector<T> b0(n, T(0)), b1(n, T(0)), b2(n, T(0));
  T* prv = b0.data();
  T* cur = b1.data();
  T* nxt = b2.data();
  const T inj = T(dt_host * dt_host);
  const T dt = T(dt_host);
  size_t snap = 0;
  for (int it = 0; it < nt; ++it) {
    cur[(size_t)siz * nx + six] += src[it] * inj;
    st.step(cur, prv, nxt, vel, sponge, nx, nz, dx, dz, dt, fs);
    if (rec_out)
      for (int r = 0; r < nrec; ++r) rec_out[(size_t)it * nrec + r] = nxt[(size_t)riz * nx + rix[r]];
    if (snaps_out && it % k == 0) {
      std::memcpy(snaps_out + snap * n, cur, n * sizeof(T));
      ++snap;
    }
    T* t = prv;
    prv = cur;
    cur = nxt;
    nxt = t;
  }
}

// Back-propagation + imaging loop of wave_solver.py:201-242, including its
// snapshot cursor rule (the cursor only moves when it is in range).
template <typename T>
void backward_loop(const Stepper<T>& st, const T* vel, const T* sponge, int nx, int nz, T dx,
                   T dz, double dt_host, int fs, const T* data, int nt, const int* rix, int nrec,
                   int riz, const T* snaps, long long nsnap, T* image, int k) {
  const size_t n = (size_t)nx * nz;
  std::vector<T> b0(n, T(0)), b1(n, T(0)), b2(n, T(0));
  T* prv = b0.data();
  T* cur = b1.data();
  T* nxt = b2.data();
  const T inj = T(dt_host * dt_host);
  const T dt = T(dt_host);
  const T w = T(dt_host * k);
  long long cursor = (nt - 1) / k;
  for (int it = nt - 1; it >= 0; --it) {
    for (int r = 0; r < nrec; ++r) cur[(size_t)riz * nx + rix[r]] += data[(size_t)it * nrec + r] * inj;
    st.step(cur, prv, nxt, vel, sponge, nx, nz, dx, dz, dt, fs);
    if (it % k == 0 && cursor >= 0 && cursor < nsnap) {
      st.corr(snaps + (size_t)cursor * n, cur, image, nx, nz, w);
      --cursor;
    }
    T* t = prv;
    prv = cur;
    cur = nxt;
    nxt = t;
  }
}

// Elastic driver conventions (no reference caller exists for
// elastic_2d_step, wave2d.c:95-212; wave_solver.py:36-53 only declares it):
//   iteration it: vz[src] += (dt*bz[src]) * s[it]  (vertical force), side work
//   on the post-injection vz (snapshot / illumination, or imaging), one
//   elastic step, then the receivers record the vertical strain rate
//   D-z vz = (c1*(vz[i]-vz[i-nx]) + c2*(vz[i+nx]-vz[i-2nx]))/dz at (riz, rix)
//   with rows outside the grid reading 0.  Back-propagation injects the data
//   like the source, receiver by receiver, and images vz against the forward
//   vz snapshots with the acoustic loop's snapshot cursor.
template <typename T>
T strain_rate(const T* vz, int nx, int nz, int iz, int ix, T rdz) {
  auto at = [&](int z) -> T { return (z < 0 || z >= nz) ? T(0) : vz[(size_t)z * nx + ix]; };
  const T a = stg_c1<T>(), b = stg_c2<T>();
  return (a * (at(iz) - at(iz - 1)) + b * (at(iz + 1) - at(iz - 2))) * rdz;
}

template <typename T>
struct ElaModel {
  const T *lam, *mu, *bx, *bz, *sponge;
};

template <typename T>
void elastic_forward_loop(const ElaModel<T>& m, int nx, int nz, T dx, T dz, T dt, int fs,
                          const T* src, int nt, int six, int siz, const int* rix, int nrec,
                          int riz, int k, T* rec_out, T* snaps_out, T* illum) {
  const size_t n = (size_t)nx * nz;
  std::vector<T> f[5];
  for (auto& v : f) v.assign(n, T(0));
  const size_t sc = (size_t)siz * nx + six;
  const T rdz = T(1) / dz;
  size_t snap = 0;
  for (int it = 0; it < nt; ++it) {
    f[1][sc] = f[1][sc] + (dt * m.bz[sc]) * src[it];
    if (it % k == 0) {
      if (snaps_out) std::memcpy(snaps_out + snap * n, f[1].data(), n * sizeof(T));
      if (illum)
        for (size_t i = 0; i < n; ++i) illum[i] = illum[i] + f[1][i] * f[1][i];
      ++snap;
    }
    elastic_step<T>(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), m.lam, m.mu,
                    m.bx, m.bz, m.sponge, nx, nz, dx, dz, dt, fs);
    if (rec_out)
      for (int r = 0; r < nrec; ++r)
        rec_out[(size_t)it * nrec + r] = strain_rate<T>(f[1].data(), nx, nz, riz, rix[r], rdz);
  }
}

template <typename T>
void elastic_backward_loop(const ElaModel<T>& m, int nx, int nz, T dx, T dz, T dt, int fs,
                           const T* data, int nt, const int* rix, int nrec, int riz,
                           const T* snaps, long long nsnap, T* image, int k, T w) {
  const size_t n = (size_t)nx * nz;
  std::vector<T> f[5];
  for (auto& v : f) v.assign(n, T(0));
  long long cursor = (nt - 1) / k;
  for (int it = nt - 1; it >= 0; --it) {
    for (int r = 0; r < nrec; ++r) {
      const size_t c = (size_t)riz * nx + rix[r];
      f[1][c] = f[1][c] + (dt * m.bz[c]) * data[(size_t)it * nrec + r];
    }
    if (it % k == 0 && cursor >= 0 && cursor < nsnap) {
      const T* s = snaps + (size_t)cursor * n;
      for (size_t i = 0; i < n; ++i) image[i] = image[i] + (s[i] * f[1][i]) * w;
      --cursor;
    }
    elastic_step<T>(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), m.lam, m.mu,
                    m.bx, m.bz, m.sponge, nx, nz, dx, dz, dt, fs);
  }
}

}  // namespace

extern "C" {

#define ORA_INST(T, SFX)                                                                          \
  void ora_acoustic_step_##SFX(const T* c, const T* p, T* n, const T* v, const T* sp, int nx,     \
                               int nz, T dx, T dz, T dt, int fs) {                                \
    acoustic_step<T>(c, p, n, v, sp, nx, nz, dx, dz, dt, fs);                                     \
  }                                                                                               \
  void ora_elastic_step_##SFX(T* vx, T* vz, T* sxx, T* szz, T* sxz, const T* lam, const T* mu,    \
                              const T* bx, const T* bz, const T* sp, int nx, int nz, T dx, T dz,  \
                              T dt, int fs) {                                                     \
    elastic_step<T>(vx, vz, sxx, szz, sxz, lam, mu, bx, bz, sp, nx, nz, dx, dz, dt, fs);          \
  }                                                                                               \
  void ora_correlate_##SFX(const T* a, const T* b, T* img, int nx, int nz, T w) {                 \
    correlate<T>(a, b, img, (long long)nx * nz, w);                                               \
  }                                                                                               \
  void ora_forward_##SFX(const T* vel, const T* sp, int nx, int nz, double dx, double dz,         \
                         double dt, int fs, const T* src, int nt, int six, int siz,               \
                         const int* rix, int nrec, int riz, int k, T* rec_out, T* snaps_out) {    \
    Stepper<T> st;                                                                                \
    forward_loop<T>(st, vel, sp, nx, nz, T(dx), T(dz), dt, fs, src, nt, six, siz, rix, nrec, riz, \
                    k, rec_out, snaps_out);                                                       \
  }                                                                                               \
  void ora_backward_##SFX(const T* vel, const T* sp, int nx, int nz, double dx, double dz,        \
                          double dt, int fs, const T* data, int nt, const int* rix, int nrec,     \
                          int riz, const T* snaps, long long nsnap, T* image, int k) {            \
    Stepper<T> st;                                                                                \
    backward_loop<T>(st, vel, sp, nx, nz, T(dx), T(dz), dt, fs, data, nt, rix, nrec, riz, snaps,  \
                     nsnap, image, k);                                                            \
  }

ORA_INST(float, f32)
ORA_INST(double, f64)

// Cerjan sponge, FP32 with libm expf exactly as wave2d.c:237-265.
void ora_build_sponge_f32(float* out, int nx, int nz, int n, float alpha, int fs) {
  for (int iz = 0; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      const int d = sponge_depth(ix, iz, nx, nz, n, fs);
      const float df = (float)d;
      out[(size_t)iz * nx + ix] = d > 0 ? expf(-alpha * df * df) : 1.0f;
    }
}
void ora_build_sponge_f64(double* out, int nx, int nz, int n, double alpha, int fs) {
  for (int iz = 0; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      const int d = sponge_depth(ix, iz, nx, nz, n, fs);
      out[(size_t)iz * nx + ix] = d > 0 ? std::exp(-alpha * d * d) : 1.0;
    }
}

// migrate_shot (rtm_engine.py:76-104): forward in vel_fwd with snapshots held
// in RAM, illumination sum_j snap_j^2 accumulated in snapshot order then
// scaled by T(dt*k) (np.sum(axis=0) adds slices one after another), then the
// back-propagation in vel_bwd into a zeroed image.  `ext_step`/`ext_corr`, when
// non-null, replace the restated kernels by the reference's own (FP32 only).
static int migrate_common_f32(const Stepper<float>& st, const float* vel_fwd, const float* vel_bwd,
                              const float* sp, int nx, int nz, double dx, double dz, double dt,
                              int fs, const float* src, int six, int siz, const float* data,
                              int nt, const int* rix, int nrec, int riz, int k, float* image,
                              float* illum) {
  const size_t n = (size_t)nx * nz;
  const long long nsnap = (nt - 1) / k + 1;
  std::vector<float> snaps;
  try {
    snaps.resize((size_t)nsnap * n);
  } catch (...) {
    return -1;
  }
  forward_loop<float>(st, vel_fwd, sp, nx, nz, (float)dx, (float)dz, dt, fs, src, nt, six, siz,
                      rix, nrec, riz, k, nullptr, snaps.data());
  std::memset(illum, 0, n * sizeof(float));
  for (long long j = 0; j < nsnap; ++j) {
    const float* s = snaps.data() + (size_t)j * n;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) illum[i] = illum[i] + s[i] * s[i];
  }
  const float w = (float)(dt * k);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) illum[i] = illum[i] * w;
  std::memset(image, 0, n * sizeof(float));
  backward_loop<float>(st, vel_bwd, sp, nx, nz, (float)dx, (float)dz, dt, fs, data, nt, rix, nrec,
                       riz, snaps.data(), nsnap, image, k);
  return 0;
}

int ora_migrate_shot_f32(const float* vel_fwd, const float* vel_bwd, const float* sp, int nx,
                         int nz, double dx, double dz, double dt, int fs, const float* src,
                         int six, int siz, const float* data, int nt, const int* rix, int nrec,
                         int riz, int k, float* image, float* illum, void* ext_step,
                         void* ext_corr) {
  Stepper<float> st;
  st.ext_step = (ref_acoustic_fn)ext_step;
  st.ext_corr = (ref_correlate_fn)ext_corr;
  return migrate_common_f32(st, vel_fwd, vel_bwd, sp, nx, nz, dx, dz, dt, fs, src, six, siz, data,
                            nt, rix, nrec, riz, k, image, illum);
}

int ora_migrate_shot_f64(const double* vel_fwd, const double* vel_bwd, const double* sp, int nx,
                         int nz, double dx, double dz, double dt, int fs, const double* src,
                         int six, int siz, const double* data, int nt, const int* rix, int nrec,
                         int riz, int k, double* image, double* illum) {
  const size_t n = (size_t)nx * nz;
  const long long nsnap = (nt - 1) / k + 1;
  std::vector<double> snaps((size_t)nsnap * n);
  Stepper<double> st;
  forward_loop<double>(st, vel_fwd, sp, nx, nz, dx, dz, dt, fs, src, nt, six, siz, rix, nrec, riz,
                       k, nullptr, snaps.data());
  std::memset(illum, 0, n * sizeof(double));
  for (long long j = 0; j < nsnap; ++j)
    for (size_t i = 0; i < n; ++i) illum[i] += snaps[(size_t)j * n + i] * snaps[(size_t)j * n + i];
  for (size_t i = 0; i < n; ++i) illum[i] *= dt * k;
  std::memset(image, 0, n * sizeof(double));
  backward_loop<double>(st, vel_bwd, sp, nx, nz, dx, dz, dt, fs, data, nt, rix, nrec, riz,
                        snaps.data(), nsnap, image, k);
  return 0;
}

#define ORA_ELA(T, SFX)                                                                           \
  void ora_elastic_forward_##SFX(const T* lam, const T* mu, const T* bx, const T* bz,             \
                                 const T* sp, int nx, int nz, double dx, double dz, double dt,    \
                                 int fs, const T* src, int nt, int six, int siz, const int* rix,  \
                                 int nrec, int riz, int k, T* rec_out, T* snaps_out, T* illum) {  \
    ElaModel<T> m{lam, mu, bx, bz, sp};                                                           \
    elastic_forward_loop<T>(m, nx, nz, T(dx), T(dz), T(dt), fs, src, nt, six, siz, rix, nrec,     \
                            riz, k, rec_out, snaps_out, illum);                                   \
  }                                                                                               \
  void ora_elastic_backward_##SFX(const T* lam, const T* mu, const T* bx, const T* bz,            \
                                  const T* sp, int nx, int nz, double dx, double dz, double dt,   \
                                  int fs, const T* data, int nt, const int* rix, int nrec,        \
                                  int riz, const T* snaps, long long nsnap, T* image, int k) {    \
    ElaModel<T> m{lam, mu, bx, bz, sp};                                                           \
    elastic_backward_loop<T>(m, nx, nz, T(dx), T(dz), T(dt), fs, data, nt, rix, nrec, riz,       \
                             snaps, nsnap, image, k, T(dt * k));                                  \
  }
ORA_ELA(float, f32)
ORA_ELA(double, f64)

// 5-point Laplacian with symmetric (edge-duplicating) boundary, the action of
// scipy.signal.convolve2d(img, K, 'same', 'symm') for the kernel of
// rtm_engine.py:31-35.  Weights: wz = T(1/dz^2), wx = T(1/dx^2),
// wc = T(-2/dx^2 - 2/dz^2), each formed in double then rounded.
void ora_laplacian_f32(const float* img, float* out, int nx, int nz, double dx, double dz) {
  const float wz = (float)(1.0 / (dz * dz));
  const float wx = (float)(1.0 / (dx * dx));
  const float wc = (float)(-2.0 / (dx * dx) - 2.0 / (dz * dz));
  auto cl = [](int i, int n) { return i < 0 ? -i - 1 : (i >= n ? 2 * n - 1 - i : i); };
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      const float up = img[(size_t)cl(iz - 1, nz) * nx + ix];
      const float dn = img[(size_t)cl(iz + 1, nz) * nx + ix];
      const float lf = img[(size_t)iz * nx + cl(ix - 1, nx)];
      const float rt = img[(size_t)iz * nx + cl(ix + 1, nx)];
      const float c = img[(size_t)iz * nx + ix];
      // FP32 sum in scipy.signal.convolve2d's window order (bottom, right,
      // centre, left, top), checked bit-exact against scipy in the tests
      float acc = dn * wz;
      acc = acc + rt * wx;
      acc = acc + c * wc;
      acc = acc + lf * wx;
      acc = acc + up * wz;
      out[(size_t)iz * nx + ix] = acc;
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Waveform misfit J = 1/2 sum |R P - d_obs|^2 of the forward loop above and
// its gradient by the discrete adjoint, written out level by level with every
// forward level kept in memory (small grids only).  There is no reference
// counterpart: the reference's back-propagation (wave_solver.py:188-242) is
// not an adjoint.  With U^m the post-injection level at iteration m and
// P^{m+1} the level the step writes (d_m = R P^{m+1}):
//   Z^{m+1} = S.(2U^m - U^{m-1} + C.(L U^m)) on I = [2,nz-2) x [2,nx-2),
//   Z^{m+1} = U^{m-2} on the other cells E (never written: three buffers
//             back), P^{m+1} = Q Z^{m+1} (free surface: row 0 = 0, row 1 =
//             -row 2), U^{m+1} = P^{m+1} + source.
// A^m = dJ/dU^m is accumulated scatter-wise from Pbar = dJ/dP^{m+1} =
// A^{m+1} + R^T r_m, descending in m.  probe[m] = A^m at the source cell
// (dJ/ds_m = inj * probe[m]).
namespace {
template <typename T>
double misfit_gradient(const T* vel, const T* sponge, int nx, int nz, T dx, T dz, double dt_host,
                       int fs, const T* src, int nt, int six, int siz, const int* rix, int nrec,
                       int riz, const T* d_obs, const T* res_in, T* grad_v, T* pred_out,
                       T* probe_out) {
  const size_t n = (size_t)nx * nz;
  const T dt = T(dt_host);
  const T inj = T(dt_host * dt_host);
  std::vector<T> U((size_t)nt * n), pred((size_t)nt * nrec), res((size_t)nt * nrec);
  {
    std::vector<T> b0(n, T(0)), b1(n, T(0)), b2(n, T(0));
    T* prv = b0.data();
    T* cur = b1.data();
    T* nxt = b2.data();
    for (int it = 0; it < nt; ++it) {
      cur[(size_t)siz * nx + six] += src[it] * inj;
      std::memcpy(&U[(size_t)it * n], cur, n * sizeof(T));
      acoustic_step<T>(cur, prv, nxt, vel, sponge, nx, nz, dx, dz, dt, fs);
      for (int r = 0; r < nrec; ++r) pred[(size_t)it * nrec + r] = nxt[(size_t)riz * nx + rix[r]];
      T* t = prv;
      prv = cur;
      cur = nxt;
      nxt = t;
    }
  }
  double loss = 0;
  for (size_t i = 0; i < res.size(); ++i) {
    res[i] = res_in ? res_in[i] : pred[i] - d_obs[i];
    loss += 0.5 * (double)res[i] * (double)res[i];
  }
  const T rdx2 = T(1) / (dx * dx), rdz2 = T(1) / (dz * dz);
  const T c0 = K4<T>::c0, c1 = lap_c1<T>(), c2 = lap_c2<T>();
  const long long off[9] = {0, -1, 1, -2, 2, -(long long)nx, nx, -2LL * nx, 2LL * nx};
  const T wgt[9] = {c0 * rdx2 + c0 * rdz2, c1 * rdx2, c1 * rdx2, c2 * rdx2, c2 * rdx2,
                    c1 * rdz2, c1 * rdz2, c2 * rdz2, c2 * rdz2};
  auto interior = [&](int iz, int ix) { return iz >= 2 && iz < nz - 2 && ix >= 2 && ix < nx - 2; };
  // A[m + 2] = dJ/dU^m for m = -2 .. nt
  std::vector<std::vector<T>> A(nt + 3, std::vector<T>(n, T(0)));
  std::vector<T> gC(n, T(0)), Pbar(n), zeta(n);
  for (int it = nt - 1; it >= 0; --it) {
    Pbar = A[it + 3];
    for (int r = 0; r < nrec; ++r) Pbar[(size_t)riz * nx + rix[r]] += res[(size_t)it * nrec + r];
    zeta = Pbar;
    if (fs)
      for (int ix = 0; ix < nx; ++ix) {
        zeta[2 * (size_t)nx + ix] -= Pbar[(size_t)nx + ix];
        zeta[ix] = T(0);
        zeta[(size_t)nx + ix] = T(0);
      }
    std::vector<T>& a0 = A[it + 2];
    std::vector<T>& a1 = A[it + 1];
    std::vector<T>& a2 = A[it];
    const T* u = &U[(size_t)it * n];
    for (int iz = 0; iz < nz; ++iz)
      for (int ix = 0; ix < nx; ++ix) {
        const size_t c = (size_t)iz * nx + ix;
        if (interior(iz, ix)) {
          const T w = sponge[c] * zeta[c];
          a0[c] += T(2) * w;
          a1[c] -= w;
          const T cw = ((vel[c] * vel[c]) * (dt * dt)) * w;
          T lu = T(0);
          for (int k = 0; k < 9; ++k) {
            a0[c + off[k]] += wgt[k] * cw;
            lu += wgt[k] * u[c + off[k]];
          }
          gC[c] += w * lu;
        } else if (!(fs && iz < 2)) {
          a2[c] += zeta[c];
        }
      }
    if (probe_out) probe_out[it] = a0[(size_t)siz * nx + six];
  }
  for (size_t c = 0; c < n; ++c) grad_v[c] = gC[c] * ((T(2) * vel[c]) * (dt * dt));
  if (pred_out) std::memcpy(pred_out, pred.data(), pred.size() * sizeof(T));
  return loss;
}
}  // namespace

extern "C" {
double ora_misfit_gradient_f32(const float* vel, const float* sponge, int nx, int nz, double dx,
                               double dz, double dt, int fs, const float* src, int nt, int six,
                               int siz, const int* rix, int nrec, int riz, const float* d_obs,
                               const float* res_in, float* grad_v, float* pred, float* probe) {
  return misfit_gradient<float>(vel, sponge, nx, nz, (float)dx, (float)dz, dt, fs, src, nt, six,
                                siz, rix, nrec, riz, d_obs, res_in, grad_v, pred, probe);
}
double ora_misfit_gradient_f64(const double* vel, const double* sponge, int nx, int nz,
                               double dx, double dz, double dt, int fs, const double* src, int nt,
                               int six, int siz, const int* rix, int nrec, int riz,
                               const double* d_obs, const double* res_in, double* grad_v,
                               double* pred, double* probe) {
  return misfit_gradient<double>(vel, sponge, nx, nz, dx, dz, dt, fs, src, nt, six, siz, rix,
                                 nrec, riz, d_obs, res_in, grad_v, pred, probe);
}
}  // extern "C"
