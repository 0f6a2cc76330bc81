// One free-surface elastic step on a small grid, every field allocated at
// exactly nz*nx floats, for a build with -fsanitize=address,undefined.
//
//   built with -DREFERENCE_ENGINE and linked against the reference engine
//   (elastic_2d_step, wave2d.c:95-212): the sanitizer reports the stencil's
//   read of row -1 on the first updated row, the defect that the oracle and
//   the HIP kernels replace by reading zeros there;
//   built against oracle/rtm_oracle.cpp (ora_elastic_step_f32): no report.
//
// Prints "ok" and exits 0 when the step returns.
#include <cstdio>
#include <vector>

extern "C" {
#ifdef REFERENCE_ENGINE
void elastic_2d_step(float* vx, float* vz, float* sxx, float* szz, float* sxz, const float* lam,
                     const float* mu, const float* bx, const float* bz, const float* sponge, int nx,
                     int nz, float dx, float dz, float dt, int free_surface_top);
#else
void ora_elastic_step_f32(float* vx, float* vz, float* sxx, float* szz, float* sxz,
                          const float* lam, const float* mu, const float* bx, const float* bz,
                          const float* sponge, int nx, int nz, float dx, float dz, float dt,
                          int free_surface);
#endif
}

int main() {
  const int nx = 12, nz = 8;
  const size_t n = (size_t)nx * nz;
  // separate heap blocks, so a read before one of them lands in its redzone
  auto* vx = new std::vector<float>(n, 0.1f);
  auto* vz = new std::vector<float>(n, 0.2f);
  auto* sxx = new std::vector<float>(n, 0.3f);
  auto* szz = new std::vector<float>(n, 0.4f);
  auto* sxz = new std::vector<float>(n, 0.5f);
  std::vector<float> lam(n, 4e9f), mu(n, 2e9f), b(n, 5e-4f), sp(n, 1.0f);
#ifdef REFERENCE_ENGINE
  elastic_2d_step(vx->data(), vz->data(), sxx->data(), szz->data(), sxz->data(), lam.data(),
                  mu.data(), b.data(), b.data(), sp.data(), nx, nz, 5.0f, 5.0f, 1e-4f, 1);
#else
  ora_elastic_step_f32(vx->data(), vz->data(), sxx->data(), szz->data(), sxz->data(), lam.data(),
                       mu.data(), b.data(), b.data(), sp.data(), nx, nz, 5.0f, 5.0f, 1e-4f, 1);
#endif
  std::printf("ok %g\n", (double)(*vx)[(size_t)nx + 2]);
  delete vx;
  delete vz;
  delete sxx;
  delete szz;
  delete sxz;
  return 0;
}
