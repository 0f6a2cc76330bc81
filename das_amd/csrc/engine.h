// Host-side engine objects shared by the C-ABI translation units.
#pragma once
#include <mutex>
#include <vector>

#include "acoustic.h"

namespace tsw {

// Launches between the two events of one pair in the kernel-timing runs.
constexpr int kTimedRun = 32;

// Owning device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // (Re)allocates when the current allocation is smaller than `nbytes`.
  int ensure(size_t nbytes);
  float* f() const { return static_cast<float*>(p); }
  int* i() const { return static_cast<int*>(p); }
  unsigned* u() const { return static_cast<unsigned*>(p); }
};

// Process-wide HIP state: the device and stream the engine runs on.
struct Ctx {
  int device = -1;
  hipStream_t stream = nullptr;
};
int ctx_get(Ctx** out);  // lazily binds to the current device and creates a stream

// Receivers on one row, as a CSR map column -> receiver indices (in the
// caller's order, so duplicate columns accumulate in that order).
struct RecvMap {
  std::vector<int> off, idx;
  int build(const int* rix, int nrec, int nx);
};

// One propagation grid (tsw_rtm/wave_solver.py:89-117): constants, the
// sponge, three rotating wavefield buffers and per-model coefficient arrays.
struct Grid {
  int nx = 0, nz = 0, n_sponge = 0, fs = 1;
  double dx = 0, dz = 0, dt = 0, alpha = 0;
  float dxf = 0, dzf = 0, dtf = 0, rdx2 = 0, rdz2 = 0, dt2 = 0, inj = 0;
  DevBuf sponge, table, buf[3], coef[2], vel, roff, ridx, src, data, rec, snaps, illum, image,
      isnap;
  DevBuf ef[5], epar[2][4];  // elastic state (vx, vz, sxx, szz, sxz) and models
  bool custom_sponge = false;  // the caller replaced the sponge (tsw_grid_set_sponge)
  long long cells() const { return (long long)nx * nz; }
  AcuStep base() const;  // geometry + constants, no buffers bound
};

int migrate_shot_dev(Grid* g, const float* d_src, int nt_src, int six, int siz,
                     const float* d_data, int nt, int nrec, int riz, int k, float* d_image,
                     float* d_illum, hipStream_t st);
int postprocess_dev(Grid* g, float* stacked, float* illum, int mute, float* norm_out,
                    float* lap_out, unsigned* maxbits, hipStream_t st);
std::recursive_mutex& api_mutex();

int grid_init(Grid* g, int nx, int nz, double dx, double dz, double dt, int n_sponge,
              double alpha, int fs);
// Host table expf(-alpha*d*d), d = 0..n, as wave2d.c:258-262 evaluates it.
void sponge_table(std::vector<float>& t, int n, float alpha);

// Divergence sentinel: counts NaN / +-inf values in each of `narr` device
// arrays (lengths n[i]) on stream `st` and waits for the counts.  The first
// array with any returns kErrDiverged with a message naming `what` and the
// index (index_base + i / per_item), otherwise kOk.
int check_finite(const float* const* arrays, const long long* n, int narr, int per_item,
                 int index_base, const char* what, hipStream_t st);

}  // namespace tsw
