// Drives every entry point of the CPU oracle (oracle/rtm_oracle.cpp) at the
// edges of its domain, for building with -fsanitize=address,undefined: the
// smallest grids each kernel accepts, sources and receivers on the first and
// last rows and columns, sponges wider than the grid, nt of 1, snapshot
// intervals that do not divide nt, with and without the free surface.  Every
// array is allocated at exactly the size the interface states, so any read or
// write past it is reported.  Prints "ok" and exits 0 when nothing is.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
void ora_acoustic_step_f32(const float*, const float*, float*, const float*, const float*, int,
                           int, float, float, float, int);
void ora_acoustic_step_f64(const double*, const double*, double*, const double*, const double*,
                           int, int, double, double, double, int);
void ora_elastic_step_f32(float*, float*, float*, float*, float*, const float*, const float*,
                          const float*, const float*, const float*, int, int, float, float, float,
                          int);
void ora_elastic_step_f64(double*, double*, double*, double*, double*, const double*,
                          const double*, const double*, const double*, const double*, int, int,
                          double, double, double, int);
void ora_correlate_f32(const float*, const float*, float*, int, int, float);
void ora_correlate_f64(const double*, const double*, double*, int, int, double);
void ora_forward_f32(const float*, const float*, int, int, double, double, double, int,
                     const float*, int, int, int, const int*, int, int, int, float*, float*);
void ora_forward_f64(const double*, const double*, int, int, double, double, double, int,
                     const double*, int, int, int, const int*, int, int, int, double*, double*);
void ora_backward_f32(const float*, const float*, int, int, double, double, double, int,
                      const float*, int, const int*, int, int, const float*, long long, float*,
                      int);
void ora_backward_f64(const double*, const double*, int, int, double, double, double, int,
                      const double*, int, const int*, int, int, const double*, long long,
                      double*, int);
void ora_build_sponge_f32(float*, int, int, int, float, int);
void ora_build_sponge_f64(double*, int, int, int, double, int);
int ora_migrate_shot_f32(const float*, const float*, const float*, int, int, double, double,
                         double, int, const float*, int, int, const float*, int, const int*, int,
                         int, int, float*, float*, void*, void*);
int ora_migrate_shot_f64(const double*, const double*, const double*, int, int, double, double,
                         double, int, const double*, int, int, const double*, int, const int*,
                         int, int, int, double*, double*);
void ora_elastic_forward_f32(const float*, const float*, const float*, const float*,
                             const float*, int, int, double, double, double, int, const float*,
                             int, int, int, const int*, int, int, int, float*, float*, float*);
void ora_elastic_forward_f64(const double*, const double*, const double*, const double*,
                             const double*, int, int, double, double, double, int, const double*,
                             int, int, int, const int*, int, int, int, double*, double*, double*);
void ora_elastic_backward_f32(const float*, const float*, const float*, const float*,
                              const float*, int, int, double, double, double, int, const float*,
                              int, const int*, int, int, const float*, long long, float*, int);
void ora_elastic_backward_f64(const double*, const double*, const double*, const double*,
                              const double*, int, int, double, double, double, int,
                              const double*, int, const int*, int, int, const double*, long long,
                              double*, int);
void ora_laplacian_f32(const float*, float*, int, int, double, double);
double ora_misfit_gradient_f32(const float*, const float*, int, int, double, double, double, int,
                               const float*, int, int, int, const int*, int, int, const float*,
                               const float*, float*, float*, float*);
double ora_misfit_gradient_f64(const double*, const double*, int, int, double, double, double,
                               int, const double*, int, int, int, const int*, int, int,
                               const double*, const double*, double*, double*, double*);
}

namespace {

unsigned g_seed = 12345u;
double urand() {  // deterministic, no <random> state to instrument
  g_seed = g_seed * 1664525u + 1013904223u;
  return (g_seed >> 8) * (1.0 / 16777216.0);
}

// a heap array of exactly n elements, so the sanitizer sees its true bounds
template <typename T> std::vector<T> arr(size_t n, double lo, double hi) {
  std::vector<T> v(n);
  for (auto& x : v) x = T(lo + (hi - lo) * urand());
  return v;
}

int g_bad = 0;
template <typename T> void finite(const std::vector<T>& v, const char* what, int nx, int nz) {
  for (const T& x : v)
    if (!std::isfinite((double)x)) {
      std::printf("non-finite %s at %dx%d\n", what, nx, nz);
      g_bad = 1;
      return;
    }
}

template <typename T> struct Api;
template <> struct Api<float> {
  static constexpr auto step = ora_acoustic_step_f32;
  static constexpr auto ela = ora_elastic_step_f32;
  static constexpr auto corr = ora_correlate_f32;
  static constexpr auto fwd = ora_forward_f32;
  static constexpr auto bwd = ora_backward_f32;
  static constexpr auto sponge = ora_build_sponge_f32;
  static constexpr auto efwd = ora_elastic_forward_f32;
  static constexpr auto ebwd = ora_elastic_backward_f32;
  static constexpr auto misfit = ora_misfit_gradient_f32;
  static int migrate(const float* vf, const float* vb, const float* sp, int nx, int nz, double dx,
                     double dt, int fs, const float* src, int six, int siz, const float* data,
                     int nt, const int* rix, int nrec, int riz, int k, float* img, float* ill) {
    return ora_migrate_shot_f32(vf, vb, sp, nx, nz, dx, dx, dt, fs, src, six, siz, data, nt, rix,
                                nrec, riz, k, img, ill, nullptr, nullptr);
  }
};
template <> struct Api<double> {
  static constexpr auto step = ora_acoustic_step_f64;
  static constexpr auto ela = ora_elastic_step_f64;
  static constexpr auto corr = ora_correlate_f64;
  static constexpr auto fwd = ora_forward_f64;
  static constexpr auto bwd = ora_backward_f64;
  static constexpr auto sponge = ora_build_sponge_f64;
  static constexpr auto efwd = ora_elastic_forward_f64;
  static constexpr auto ebwd = ora_elastic_backward_f64;
  static constexpr auto misfit = ora_misfit_gradient_f64;
  static int migrate(const double* vf, const double* vb, const double* sp, int nx, int nz,
                     double dx, double dt, int fs, const double* src, int six, int siz,
                     const double* data, int nt, const int* rix, int nrec, int riz, int k,
                     double* img, double* ill) {
    return ora_migrate_shot_f64(vf, vb, sp, nx, nz, dx, dx, dt, fs, src, six, siz, data, nt, rix,
                                nrec, riz, k, img, ill);
  }
};

template <typename T> void grid_case(int nx, int nz, int fs, int nsp, int nt, int k) {
  using A = Api<T>;
  const size_t n = (size_t)nx * nz;
  const double dx = 5.0, dt = 4e-4;
  std::vector<T> sp(n);
  A::sponge(sp.data(), nx, nz, nsp, T(0.0015), fs);
  // single steps
  auto c = arr<T>(n, -1, 1), p = arr<T>(n, -1, 1), nx_ = arr<T>(n, 0, 0);
  auto v = arr<T>(n, 1500, 3000);
  A::step(c.data(), p.data(), nx_.data(), v.data(), sp.data(), nx, nz, T(dx), T(dx), T(dt), fs);
  finite(nx_, "acoustic step", nx, nz);
  auto vx = arr<T>(n, -1, 1), vz = arr<T>(n, -1, 1), sxx = arr<T>(n, -1, 1),
       szz = arr<T>(n, -1, 1), sxz = arr<T>(n, -1, 1);
  auto lam = arr<T>(n, 1e9, 4e9), mu = arr<T>(n, 1e9, 3e9), bx = arr<T>(n, 3e-4, 5e-4),
       bz = arr<T>(n, 3e-4, 5e-4);
  A::ela(vx.data(), vz.data(), sxx.data(), szz.data(), sxz.data(), lam.data(), mu.data(),
         bx.data(), bz.data(), sp.data(), nx, nz, T(dx), T(dx), T(1e-5), fs);
  A::ela(vx.data(), vz.data(), sxx.data(), szz.data(), sxz.data(), lam.data(), mu.data(),
         bx.data(), bz.data(), nullptr, nx, nz, T(dx), T(dx), T(1e-5), fs);
  finite(sxx, "elastic step", nx, nz);
  auto img = arr<T>(n, 0, 0);
  A::corr(c.data(), p.data(), img.data(), nx, nz, T(dt));
  // loops: source in a corner, receivers on the first and last columns
  const int corners[4][2] = {{0, 0}, {nx - 1, nz - 1}, {nx - 1, 0}, {0, nz - 1}};
  std::vector<int> rix = {0, nx - 1, nx / 2, 0};
  const int nrec = (int)rix.size();
  const long long nsnap = (nt - 1) / k + 1;
  auto src = arr<T>(nt, -1, 1), data = arr<T>((size_t)nt * nrec, -1, 1);
  for (const auto& sc : corners) {
    for (int riz : {0, nz - 1}) {
      std::vector<T> rec((size_t)nt * nrec), snaps((size_t)nsnap * n), im(n), il(n);
      A::fwd(v.data(), sp.data(), nx, nz, dx, dx, dt, fs, src.data(), nt, sc[0], sc[1],
             rix.data(), nrec, riz, k, rec.data(), snaps.data());
      A::bwd(v.data(), sp.data(), nx, nz, dx, dx, dt, fs, data.data(), nt, rix.data(), nrec, riz,
             snaps.data(), nsnap, im.data(), k);
      finite(im, "backward image", nx, nz);
      if (A::migrate(v.data(), v.data(), sp.data(), nx, nz, dx, dt, fs, src.data(), sc[0], sc[1],
                     data.data(), nt, rix.data(), nrec, riz, k, im.data(), il.data()) != 0) {
        std::printf("migrate refused %dx%d\n", nx, nz);
        g_bad = 1;
      }
      finite(il, "illumination", nx, nz);
      std::vector<T> erec((size_t)nt * nrec), esnap((size_t)nsnap * n), eill(n), eim(n);
      A::efwd(lam.data(), mu.data(), bx.data(), bz.data(), sp.data(), nx, nz, dx, dx, 1e-5, fs,
              src.data(), nt, sc[0], sc[1], rix.data(), nrec, riz, k, erec.data(), esnap.data(),
              eill.data());
      A::ebwd(lam.data(), mu.data(), bx.data(), bz.data(), sp.data(), nx, nz, dx, dx, 1e-5, fs,
              data.data(), nt, rix.data(), nrec, riz, esnap.data(), nsnap, eim.data(), k);
      std::vector<T> grad(n), pred((size_t)nt * nrec), probe(nt);
      A::misfit(v.data(), sp.data(), nx, nz, dx, dx, dt, fs, src.data(), nt, sc[0], sc[1],
                rix.data(), nrec, riz, data.data(), nullptr, grad.data(), pred.data(),
                probe.data());
      A::misfit(v.data(), sp.data(), nx, nz, dx, dx, dt, fs, src.data(), nt, sc[0], sc[1],
                rix.data(), nrec, riz, data.data(), data.data(), grad.data(), nullptr, nullptr);
      finite(grad, "misfit gradient", nx, nz);
    }
  }
  // forward without snapshots or receivers' output buffer use
  std::vector<T> rec((size_t)nt * nrec);
  A::fwd(v.data(), sp.data(), nx, nz, dx, dx, dt, fs, src.data(), nt, 0, 0, rix.data(), nrec, 0,
         k, rec.data(), nullptr);
}

}  // namespace

int main() {
  // grids from the smallest the stencils touch (5 x 3) to a few tiles; sponge
  // widths 0, 1, 3 and wider than the grid
  const int shapes[][2] = {{5, 3}, {5, 5}, {6, 4}, {7, 6}, {16, 9}, {33, 17}};
  for (const auto& s : shapes)
    for (int fs = 0; fs < 2; ++fs)
      for (int nsp : {0, 1, 3, 40}) {
        grid_case<float>(s[0], s[1], fs, nsp, 1, 1);
        grid_case<float>(s[0], s[1], fs, nsp, 7, 3);
        grid_case<double>(s[0], s[1], fs, nsp, 6, 2);
      }
  for (int nx : {1, 2, 5, 9})
    for (int nz : {1, 2, 4}) {
      const size_t n = (size_t)nx * nz;
      auto img = arr<float>(n, -1, 1);
      std::vector<float> out(n);
      ora_laplacian_f32(img.data(), out.data(), nx, nz, 5.0, 2.5);
      std::vector<float> spf(n);
      std::vector<double> spd(n);
      ora_build_sponge_f32(spf.data(), nx, nz, 3, 0.0015f, 1);
      ora_build_sponge_f64(spd.data(), nx, nz, 3, 0.0015, 0);
    }
  if (g_bad) return 1;
  std::printf("ok\n");
  return 0;
}
