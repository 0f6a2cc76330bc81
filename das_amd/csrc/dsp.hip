// Preprocessing of continuous DAS data before correlation
// (run_das_pipeline.py:165-168, :280).  The reference's preprocessing module
// is absent from its source tree; the steps follow its README (README.md:27-52)
// and the scipy routines named there:
//   - zero-phase IIR filtering, scipy.signal.sosfiltfilt semantics (odd
//     extension of padlen samples, steady-state initial conditions scaled by
//     the first sample, forward pass, reversed pass), in FP64 with scipy's
//     direct-form-II-transposed recurrence, optionally keeping every q-th
//     output sample (decimation) and/or its sign (1-bit normalisation);
//   - spectral whitening X(f) / (|X(f)| + eps) inside [fmin, fmax], 0 outside;
//   - an F-K fan filter keeping apparent velocities |f / k| in [vmin, vmax].
// The FFTs are hipFFT library transforms (plain, batched).
#include <hipfft/hipfft.h>

#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "engine.h"
#include "tswrtm.h"

namespace tsw {

constexpr int kMaxSections = 16;
constexpr int kSosChunk = 32;  // samples per chunk read ahead of the recursion

struct SosArgs {
  const float* x;   // nblk blocks (nt, nch) row-major, xs elements apart
  int nt, nch, nsec, padlen, step, sign, nblk;
  long long xs, os;  // elements between consecutive blocks of x and of out
  double sos[kMaxSections][6];
  double zi[kMaxSections][2];
  double* tmp;      // (nt + 2 padlen, nblk * nch) forward pass
  float* out;       // nblk blocks (ceil(nt / step), nch), os elements apart
};

// Extended sample e of channel c (odd extension, float32 as numpy forms it),
// in two halves so that a chunk's loads can all be issued before any value is
// used: sos_ext_at is the raw sample the value is formed from (any e, clamped
// into the block, no branch around the load) and sos_ext_val forms the value.
__device__ __forceinline__ float sos_ext_at(const SosArgs& a, const float* x, int c, int e) {
  const int n = a.nt, p = a.padlen;
  int r = e < p ? p - e : (e < p + n ? e - p : n - 2 - (e - p - n));
  r = min(max(r, 0), n - 1);  // samples past the end of the last chunk
  return x[(long long)r * a.nch + c];
}

__device__ __forceinline__ float sos_ext_val(const SosArgs& a, int e, float raw, float first,
                                             float last) {
  if (e < a.padlen) return 2.0f * first - raw;
  if (e < a.padlen + a.nt) return raw;
  return 2.0f * last - raw;
}

// One thread per channel: consecutive channels are consecutive addresses.
// The channels of nblk blocks (independent recordings of one shape) are
// filtered side by side: thread g takes channel g % nch of block g / nch.
// NS > 0: exactly NS sections, unrolled, with the state and coefficients in
// registers; NS == 0: any count up to kMaxSections (state in private memory).
// The recursion is sequential in time; its input is read in chunks of
// kSosChunk samples, two chunks in flight (A and B): a chunk's loads are
// issued before the previous chunk's recursion, so the memory latency is
// hidden behind it.  A 64-thread block is one wave and there are few of them
// (one per 64 channels), so the kernel may use the whole register file.
template <int NS>
__global__ __launch_bounds__(64, 1) void k_sosfiltfilt(SosArgs a) {
  const int cg = blockIdx.x * blockDim.x + threadIdx.x;
  if (cg >= a.nch * a.nblk) return;
  const int blk = cg / a.nch, c = cg - blk * a.nch;
  const float* x = a.x + blk * a.xs;
  float* out = a.out + blk * a.os;
  const long long tstride = (long long)a.nch * a.nblk;  // tmp row: every block's channels
  constexpr int MS = NS > 0 ? NS : kMaxSections;
  constexpr int K = kSosChunk;
  const int nsec = NS > 0 ? NS : a.nsec;
  const int ne = a.nt + 2 * a.padlen;
  double z[MS][2], b0[MS], b1[MS], b2[MS], a1[MS], a2[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    b0[s] = a.sos[s][0];
    b1[s] = a.sos[s][1];
    b2[s] = a.sos[s][2];
    a1[s] = a.sos[s][4];
    a2[s] = a.sos[s][5];
  }
  const float first = x[c], last = x[(long long)(a.nt - 1) * a.nch + c];
  const double x0 = (double)sos_ext_val(a, 0, sos_ext_at(a, x, c, 0), first, last);
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    z[s][0] = a.zi[s][0] * x0;
    z[s][1] = a.zi[s][1] * x0;
  }
  // one sample through the cascade, in scipy's (direct form II transposed) order
  auto cascade = [&](double xn) -> double {
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      if (NS == 0 && s >= nsec) break;
      const double xc = xn;
      xn = b0[s] * xc + z[s][0];
      z[s][0] = b1[s] * xc - a1[s] * xn + z[s][1];
      z[s][1] = b2[s] * xc - a2[s] * xn;
    }
    return xn;
  };
  // forward pass over the extended samples, into tmp
  double y = 0.0;
  {
    float xa[K], xb[K];
    auto load = [&](float* buf, int e0) {
#pragma unroll
      for (int u = 0; u < K; ++u) buf[u] = sos_ext_at(a, x, c, e0 + u);
    };
    auto run = [&](const float* buf, int e0) {
#pragma unroll
      for (int u = 0; u < K; ++u) {
        const int e = e0 + u;
        if (e >= ne) break;
        y = cascade((double)sos_ext_val(a, e, buf[u], first, last));
        a.tmp[(long long)e * tstride + cg] = y;
      }
    };
    load(xa, 0);
    for (int e0 = 0; e0 < ne; e0 += 2 * K) {
      load(xb, e0 + K);
      run(xa, e0);
      load(xa, e0 + 2 * K);
      run(xb, e0 + K);
    }
  }
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    z[s][0] = a.zi[s][0] * y;
    z[s][1] = a.zi[s][1] * y;
  }
  // reversed pass over tmp (written by this thread only), every step-th sample out
  {
    double ta[K], tb[K];
    auto load = [&](double* buf, int e1) {
#pragma unroll
      for (int u = 0; u < K; ++u) buf[u] = a.tmp[(long long)max(e1 - u, 0) * tstride + cg];
    };
    auto run = [&](const double* buf, int e1) {
#pragma unroll
      for (int u = 0; u < K; ++u) {
        const int e = e1 - u;
        if (e < 0) break;
        const double xn = cascade(buf[u]);
        const int t = e - a.padlen;
        if (t >= 0 && t < a.nt && t % a.step == 0) {
          float v = (float)xn;
          if (a.sign) v = v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f);
          out[(long long)(t / a.step) * a.nch + c] = v;
        }
      }
    };
    load(ta, ne - 1);
    for (int e1 = ne - 1; e1 >= 0; e1 -= 2 * K) {
      load(tb, e1 - K);
      run(ta, e1);
      load(ta, e1 - 2 * K);
      run(tb, e1 - K);
    }
  }
}

// np.sign(x).astype(float32)
__global__ void k_sign(const float* x, long long n, float* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    out[i] = v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : (v == v ? 0.0f : v));
  }
}

// X(f) <- X(f) / ((|X(f)| + eps) nt) for bins in [k0, k1], 0 elsewhere (the
// nt is the unnormalised inverse transform's).  X: (nf, nch), channel fastest.
__global__ void k_whiten(hipfftComplex* X, int nf, int nch, int k0, int k1, float eps,
                         float scale) {
  const long long n = (long long)nf * nch;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(i / nch);
    hipfftComplex v = X[i];
    if (k < k0 || k > k1) {
      v.x = 0.0f;
      v.y = 0.0f;
    } else {
      const float d = (sqrtf(v.x * v.x + v.y * v.y) + eps) * scale;
      v.x = v.x / d;
      v.y = v.y / d;
    }
    X[i] = v;
  }
}

// F-K fan: keep (f, k) with k != 0 and vmin <= |f| / |k| <= vmax; X (nt, nx)
// row-major 2-D spectrum, frequencies fftfreq(nt, dt), wavenumbers fftfreq(nx, dx).
__global__ void k_fk_mask(hipfftComplex* X, int nt, int nx, double dt, double dx, double vmin,
                          double vmax, float scale) {
  const long long n = (long long)nt * nx;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int it = (int)(i / nx), ix = (int)(i - (long long)it * nx);
    const int ft = it <= (nt - 1) / 2 ? it : it - nt;
    const int kx = ix <= (nx - 1) / 2 ? ix : ix - nx;
    // numpy.fft.fftfreq: index * (1 / (n d))
    const double f = fabs((double)ft * (1.0 / (nt * dt))), k = fabs((double)kx * (1.0 / (nx * dx)));
    bool keep = false;
    if (kx != 0) {
      const double v = f / k;
      keep = v >= vmin && v <= vmax;
    }
    hipfftComplex c = X[i];
    if (keep) {
      c.x *= scale;
      c.y *= scale;
    } else {
      c.x = 0.0f;
      c.y = 0.0f;
    }
    X[i] = c;
  }
}

__global__ void k_real_part(const hipfftComplex* X, long long n, float* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = X[i].x;
}

__global__ void k_to_complex(const float* x, long long n, hipfftComplex* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    hipfftComplex c;
    c.x = x[i];
    c.y = 0.0f;
    out[i] = c;
  }
}

static unsigned gridn(long long n) {
  return (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 65536));
}

#define TSW_FFT(call)                                                              \
  do {                                                                             \
    hipfftResult r_ = (call);                                                      \
    if (r_ != HIPFFT_SUCCESS) {                                                    \
      ::tsw::set_error("%s failed: hipfft status %d", #call, (int)r_);             \
      return ::tsw::kErrHip;                                                       \
    }                                                                              \
  } while (0)

// FFT plans and device buffers of the host-pointer DSP entry points, kept for
// the life of the process: repeated calls of one shape (a block after block of
// the same recording) make their plans once -- making a plan for a length such
// as 30000 costs far more than running it -- and the buffers grow to the
// largest call.  Every entry point holds the library mutex, which guards them.
struct DspCache {
  std::map<std::array<int, 3>, hipfftHandle> plans;  // (kind, n0, n1) -> plan
  DevBuf buf[4];
  // two pinned staging buffers for host blocks, and the event of the copy
  // that last read each
  void* pin[2] = {nullptr, nullptr};
  size_t pin_bytes = 0;
  hipEvent_t pin_done[2] = {nullptr, nullptr};
};

static DspCache& dsp_cache() {
  static DspCache* c = new DspCache();  // not destroyed: released with the process
  return *c;
}

// kind 0: nch real-to-complex transforms of length n0 = nt along axis 0 of an
// (nt, nch) array (n1 = nch); kind 1: the inverse (complex-to-real); kind 2: a
// 2D complex transform of an (n0, n1) array.
static int cached_plan(int kind, int n0, int n1, hipfftHandle* out) {
  auto& plans = dsp_cache().plans;
  const std::array<int, 3> key{kind, n0, n1};
  auto it = plans.find(key);
  if (it != plans.end()) {
    *out = it->second;
    return kOk;
  }
  hipfftHandle h = 0;
  if (kind == 2) {
    TSW_FFT(hipfftPlan2d(&h, n0, n1, HIPFFT_C2C));
  } else {
    const int nf = n0 / 2 + 1;
    int n[1] = {n0}, re[1] = {n0}, cx[1] = {nf};
    if (kind == 0) {
      TSW_FFT(hipfftPlanMany(&h, 1, n, re, n1, 1, cx, n1, 1, HIPFFT_R2C, n1));
    } else {
      TSW_FFT(hipfftPlanMany(&h, 1, n, cx, n1, 1, re, n1, 1, HIPFFT_C2R, n1));
    }
  }
  plans.emplace(key, h);
  *out = h;
  return kOk;
}

static DevBuf& scratch(int i) { return dsp_cache().buf[i]; }

// Copies nbytes from host memory (pageable) to d_dst on st through the pinned
// staging buffer `slot`: the host copy is split over several threads, and it
// waits only for the earlier copy that read the same buffer, so the host copy
// of one block overlaps the transfer of the one before.
static int staged_h2d(void* d_dst, const void* src, size_t nbytes, int slot, hipStream_t st) {
  DspCache& c = dsp_cache();
  if (c.pin_bytes < nbytes) {
    for (int i = 0; i < 2; ++i) {
      if (c.pin_done[i]) TSW_HIP(hipEventSynchronize(c.pin_done[i]));
      if (c.pin[i]) TSW_HIP(hipHostFree(c.pin[i]));
      c.pin[i] = nullptr;
    }
    c.pin_bytes = 0;
    for (int i = 0; i < 2; ++i)
      if (hipHostMalloc(&c.pin[i], nbytes, hipHostMallocDefault) != hipSuccess) {
        set_error("pinned staging buffer of %zu bytes: allocation failed", nbytes);
        return kErrMem;
      }
    c.pin_bytes = nbytes;
  }
  if (!c.pin_done[slot]) TSW_HIP(hipEventCreateWithFlags(&c.pin_done[slot], hipEventDisableTiming));
  TSW_HIP(hipEventSynchronize(c.pin_done[slot]));
  char* dst = static_cast<char*>(c.pin[slot]);
  const char* from = static_cast<const char*>(src);
  const int nth = nbytes >= (64u << 20) ? 8 : 1;
  const size_t part = (nbytes + nth - 1) / nth;
  std::vector<std::thread> th;
  for (int t = 1; t < nth; ++t) {
    const size_t o = t * part;
    if (o < nbytes)
      th.emplace_back([=] { std::memcpy(dst + o, from + o, std::min(part, nbytes - o)); });
  }
  std::memcpy(dst, from, std::min(part, nbytes));
  for (auto& t : th) t.join();
  TSW_HIP(hipMemcpyAsync(d_dst, dst, nbytes, hipMemcpyHostToDevice, st));
  TSW_HIP(hipEventRecord(c.pin_done[slot], st));
  return kOk;
}

static int sos_check(int nt, int nch, const double* sos, int nsec, int padlen, int step) {
  TSW_ARG(sos != nullptr, "null argument");
  TSW_ARG(nt >= 1 && nch >= 1 && step >= 1, "bad sizes");
  TSW_ARG(nsec >= 1 && nsec <= kMaxSections, "1 to %d second-order sections", kMaxSections);
  TSW_ARG(padlen >= 0 && padlen < nt,
          "the input must be longer than padlen (%d samples, as sosfiltfilt requires)", padlen);
  for (int s = 0; s < nsec; ++s)
    TSW_ARG(sos[s * 6 + 3] == 1.0, "section %d is not normalised (a0 = %g)", s, sos[s * 6 + 3]);
  return kOk;
}

// Zero-phase filtering of the (nt, nch) device array d_x into d_out
// (ceil(nt / step) rows), optionally the sign of each output sample; with
// nblk > 1, nblk such arrays xs (input) and os (output) elements apart in one
// launch.  The padded FP64 work array is scratch buffer 1.
static int sos_run(const float* d_x, int nt, int nch, const double* sos, int nsec,
                   const double* zi, int padlen, int step, int sign, float* d_out,
                   hipStream_t st, int nblk = 1, long long xs = 0, long long os = 0) {
  TSW_ARG(zi != nullptr, "null argument");
  SosArgs a;
  std::memset(&a, 0, sizeof a);
  for (int s = 0; s < nsec; ++s) {
    for (int j = 0; j < 6; ++j) a.sos[s][j] = sos[s * 6 + j];
    a.zi[s][0] = zi[s * 2];
    a.zi[s][1] = zi[s * 2 + 1];
  }
  a.nt = nt;
  a.nch = nch;
  a.nsec = nsec;
  a.padlen = padlen;
  a.step = step;
  a.sign = sign ? 1 : 0;
  a.nblk = nblk;
  a.xs = xs;
  a.os = os;
  DevBuf& dtmp = scratch(1);
  TSW_TRY(dtmp.ensure((size_t)(nt + 2 * padlen) * nch * nblk * sizeof(double)));
  a.x = d_x;
  a.tmp = (double*)dtmp.p;
  a.out = d_out;
  const dim3 grid((nch * nblk + 63) / 64), block(64);
  switch (nsec) {  // the usual cascades (order 2 to 16) unrolled
    case 1: hipLaunchKernelGGL(k_sosfiltfilt<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_sosfiltfilt<2>, grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL(k_sosfiltfilt<3>, grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_sosfiltfilt<4>, grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL(k_sosfiltfilt<5>, grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL(k_sosfiltfilt<6>, grid, block, 0, st, a); break;
    case 8: hipLaunchKernelGGL(k_sosfiltfilt<8>, grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(k_sosfiltfilt<0>, grid, block, 0, st, a); break;
  }
  TSW_HIP(hipGetLastError());
  return kOk;
}

// Spectral whitening of the (nt, nch) device array d_x into d_out (may be
// d_x): X(f) / (|X(f)| + eps) for fmin <= f <= fmax, 0 outside; the spectrum
// is scratch buffer 3.
static int whiten_run(float* d_x, int nt, int nch, double fs, double fmin, double fmax,
                      double eps, float* d_out, hipStream_t st) {
  const int nf = nt / 2 + 1;
  // bins k with fmin <= k fs / nt <= fmax
  const int k0 = std::max(0, (int)std::ceil(fmin * nt / fs - 1e-9));
  const int k1 = std::min(nf - 1, (int)std::floor(fmax * nt / fs + 1e-9));
  DevBuf& dX = scratch(3);
  TSW_TRY(dX.ensure((size_t)nf * nch * sizeof(hipfftComplex)));
  hipfftHandle fwd, inv;
  TSW_TRY(cached_plan(0, nt, nch, &fwd));
  TSW_TRY(cached_plan(1, nt, nch, &inv));
  TSW_FFT(hipfftSetStream(fwd, st));
  TSW_FFT(hipfftSetStream(inv, st));
  TSW_FFT(hipfftExecR2C(fwd, d_x, (hipfftComplex*)dX.p));
  hipLaunchKernelGGL(k_whiten, dim3(gridn((long long)nf * nch)), dim3(256), 0, st,
                     (hipfftComplex*)dX.p, nf, nch, k0, k1, (float)eps, (float)nt);
  TSW_HIP(hipGetLastError());
  TSW_FFT(hipfftExecC2R(inv, (hipfftComplex*)dX.p, d_out));
  return kOk;
}

}  // namespace tsw

using namespace tsw;

extern "C" {

int tsw_sosfiltfilt(const float* x, int nt, int nch, const double* sos, int nsec,
                    const double* zi, int padlen, int step, int sign, float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(x && out && zi, "null argument");
  TSW_TRY(sos_check(nt, nch, sos, nsec, padlen, step));
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const int nout = (nt + step - 1) / step;
  DevBuf &dx = scratch(0), &dout = scratch(2);
  TSW_TRY(dx.ensure((size_t)nt * nch * sizeof(float)));
  TSW_TRY(dout.ensure((size_t)nout * nch * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(dx.p, x, (size_t)nt * nch * sizeof(float), hipMemcpyHostToDevice, st));
  TSW_TRY(sos_run(dx.f(), nt, nch, sos, nsec, zi, padlen, step, sign, dout.f(), st));
  TSW_HIP(hipMemcpyAsync(out, dout.p, (size_t)nout * nch * sizeof(float), hipMemcpyDeviceToHost,
                         st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_sign(const float* x, long long n, float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(x && out && n >= 0, "bad arguments");
  if (n == 0) return kOk;
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  DevBuf &a = scratch(0), &b = scratch(1);
  TSW_TRY(a.ensure((size_t)n * sizeof(float)));
  TSW_TRY(b.ensure((size_t)n * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(a.p, x, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_sign, dim3(gridn(n)), dim3(256), 0, st, a.f(), n, b.f());
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(out, b.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_spectral_whitening(const float* x, int nt, int nch, double fs, double fmin, double fmax,
                           double eps, float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(x && out && nt >= 2 && nch >= 1 && fs > 0 && fmin <= fmax && eps >= 0,
          "bad arguments");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  DevBuf& dx = scratch(0);
  TSW_TRY(dx.ensure((size_t)nt * nch * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(dx.p, x, (size_t)nt * nch * sizeof(float), hipMemcpyHostToDevice, st));
  TSW_TRY(whiten_run(dx.f(), nt, nch, fs, fmin, fmax, eps, dx.f(), st));
  TSW_HIP(hipMemcpyAsync(out, dx.p, (size_t)nt * nch * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_preprocess_device_blocks(int nblk, const float* const* x, int nt, int nch,
                                 const double* lp_sos, int lp_nsec, const double* lp_zi,
                                 int lp_padlen, int q, const double* bp_sos, int bp_nsec,
                                 const double* bp_zi, int bp_padlen, double fs_proc, double fmin,
                                 double fmax, double eps, float* const* d_out, void* stream) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(nblk >= 1 && x && d_out && bp_sos && bp_zi, "null argument or no block");
  for (int b = 0; b < nblk; ++b) TSW_ARG(x[b] && d_out[b], "block %d: null pointer", b);
  TSW_ARG(nt >= 1 && nch >= 1 && q >= 1, "bad sizes");
  TSW_ARG((long long)nch * nblk < (1LL << 31), "too many channels in one call");
  TSW_ARG(fs_proc > 0 && fmin <= fmax && eps >= 0, "bad band");
  const int nt1 = q > 1 ? (nt + q - 1) / q : nt;  // samples after decimation
  if (q > 1) {
    TSW_ARG(lp_sos && lp_zi, "null argument");
    TSW_TRY(sos_check(nt, nch, lp_sos, lp_nsec, lp_padlen, q));
  }
  TSW_TRY(sos_check(nt1, nch, bp_sos, bp_nsec, bp_padlen, 1));
  TSW_ARG(nt1 >= 2, "the decimated block needs at least 2 samples");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long xs = (long long)nt * nch, os = (long long)nt1 * nch;
  DevBuf &a = scratch(0), &b = scratch(2);
  TSW_TRY(a.ensure((size_t)(xs * nblk) * sizeof(float)));
  TSW_TRY(b.ensure((size_t)(os * nblk) * sizeof(float)));
  for (int k = 0; k < nblk; ++k)
    TSW_TRY(staged_h2d(a.f() + k * xs, x[k], (size_t)xs * sizeof(float), k & 1, st));
  float* cur = a.f();
  float* other = b.f();
  long long cs = xs;
  if (q > 1) {  // low-pass at the new Nyquist frequency, every q-th sample
    TSW_TRY(sos_run(a.f(), nt, nch, lp_sos, lp_nsec, lp_zi, lp_padlen, q, 0, b.f(), st, nblk,
                    xs, os));
    cur = b.f();
    other = a.f();
    cs = os;
  }
  // band-pass with the 1-bit normalisation of its output, then whitening
  TSW_TRY(sos_run(cur, nt1, nch, bp_sos, bp_nsec, bp_zi, bp_padlen, 1, 1, other, st, nblk, cs,
                  os));
  for (int k = 0; k < nblk; ++k)
    TSW_TRY(whiten_run(other + k * os, nt1, nch, fs_proc, fmin, fmax, eps, d_out[k], st));
  // the scratch buffers are free again once the stream has drained
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_preprocess_device(const float* x, int nt, int nch, const double* lp_sos, int lp_nsec,
                          const double* lp_zi, int lp_padlen, int q, const double* bp_sos,
                          int bp_nsec, const double* bp_zi, int bp_padlen, double fs_proc,
                          double fmin, double fmax, double eps, float* d_out, void* stream) {
  return tsw_preprocess_device_blocks(1, &x, nt, nch, lp_sos, lp_nsec, lp_zi, lp_padlen, q,
                                      bp_sos, bp_nsec, bp_zi, bp_padlen, fs_proc, fmin, fmax,
                                      eps, &d_out, stream);
}

int tsw_fk_filter(const float* g, int nt, int nx, double dt, double dx, double vmin, double vmax,
                  float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(g && out && nt >= 1 && nx >= 1 && dt > 0 && dx > 0 && vmin <= vmax, "bad arguments");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const long long n = (long long)nt * nx;
  DevBuf &dg = scratch(0), &dX = scratch(1);
  TSW_TRY(dg.ensure((size_t)n * sizeof(float)));
  TSW_TRY(dX.ensure((size_t)n * sizeof(hipfftComplex)));
  TSW_HIP(hipMemcpyAsync(dg.p, g, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_to_complex, dim3(gridn(n)), dim3(256), 0, st, dg.f(), n,
                     (hipfftComplex*)dX.p);
  hipfftHandle p;
  TSW_TRY(cached_plan(2, nt, nx, &p));
  TSW_FFT(hipfftSetStream(p, st));
  TSW_FFT(hipfftExecC2C(p, (hipfftComplex*)dX.p, (hipfftComplex*)dX.p, HIPFFT_FORWARD));
  hipLaunchKernelGGL(k_fk_mask, dim3(gridn(n)), dim3(256), 0, st, (hipfftComplex*)dX.p, nt, nx,
                     dt, dx, vmin, vmax, 1.0f / (float)n);
  TSW_HIP(hipGetLastError());
  TSW_FFT(hipfftExecC2C(p, (hipfftComplex*)dX.p, (hipfftComplex*)dX.p, HIPFFT_BACKWARD));
  hipLaunchKernelGGL(k_real_part, dim3(gridn(n)), dim3(256), 0, st, (const hipfftComplex*)dX.p,
                     n, dg.f());
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(out, dg.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

}  // extern "C"
