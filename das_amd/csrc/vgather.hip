// Virtual-shot gathers from continuous DAS data (the step before
// back-injection): cross-correlation of source channels with receiver
// channels (tsw_rtm/interferometry.py:26-79), time derivative to the
// empirical Green's function (:82-104), branch selection (:112-163) and linear
// resampling to the simulation time step (run_das_pipeline.py:48-55).
//
// The reference correlates through zero-padded FFTs; here the lag window is
// summed directly, c[lag] = sum_t s[t] r[t + lag] for |lag| <= nlag: FP32
// samples multiply exactly in FP64 and accumulate in FP64 in t order, so the
// result is the exact correlation rounded once to FP32 (the FFT route carries
// the FFT's rounding instead).  The cost, nt * (2 nlag + 1) per pair, is
// FP64-FMA bound; each workgroup is one wavefront computing up to 64 L lags of
// one source/receiver pair with an L-lag register window over LDS-staged
// samples.
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "tswrtm.h"

namespace tsw {

// columns ch[c] of data (nt, nch), row-major, into out (ncol, nt)
__global__ void k_take_cols(const float* data, int nt, int nch, const int* ch, int ncol,
                            float* out) {
  __shared__ float tile[32][33];
  const int t0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  for (int j = threadIdx.y; j < 32; j += blockDim.y) {
    const int t = t0 + j, c = c0 + threadIdx.x;
    tile[j][threadIdx.x] = (t < nt && c < ncol) ? data[(long long)t * nch + ch[c]] : 0.0f;
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += blockDim.y) {
    const int c = c0 + j, t = t0 + threadIdx.x;
    if (c < ncol && t < nt) out[(long long)c * nt + t] = tile[threadIdx.x][j];
  }
}

// samples per LDS chunk: with L = 6, 96 samples take 3.9 KB of LDS per wave
// (the chunk plus the 384-lag window, in FP64), so LDS does not cap the waves
// per CU; measured per 64 x 800-pair launch: 256 samples 22.1 ms, 128 21.4,
// 96 20.9, 64 21.8 (the smaller the chunk, the more staging and barriers)
constexpr int kXC = 96;

// rows of S (nrow, nt) float32 into S64 (nrow, ntp) float64, zero past nt
__global__ void k_rows_f64(const float* S, int nt, int ntp, int nrow, double* S64) {
  const long long n = (long long)nrow * ntp;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int row = (int)(e / ntp), t = (int)(e - (long long)row * ntp);
    S64[e] = t < nt ? (double)S[(long long)row * nt + t] : 0.0;
  }
}

// pairs[(src * nrec + rec) * nl + j] = sum_t S[src][t] * R[rec][t - nlag + j]
// (nl = 2 nlag + 1; samples outside [0, nt) are zero).  Grid (nrec, nsrc,
// lag tiles of 64 L), one wavefront per workgroup, L lags per lane.  The
// source samples are the same for every lane: they come from S64 (FP64, rows
// padded with zeros to ntp >= nt + chunk) through scalar loads, so
// only the receiver window is read from LDS.  The product of two FP32 samples
// is exact in FP64, so fma(s, r, acc) is the product added with one rounding:
// the sum is the same as a separate multiply and add, in one instruction.
template <int L>
__global__ __launch_bounds__(64) void k_xcorr(const double* __restrict__ S64, int ntp,
                                              const float* __restrict__ R, int nt, int nlag,
                                              int nrec, float* pairs) {
  constexpr int kW = 64 * L;
  constexpr int kC = (kXC / L) * L;  // chunk: a whole number of L-sample steps
  __shared__ double r_c[kC + kW + 2 * L];  // receiver samples converted once
  const int rec = blockIdx.x, src = blockIdx.y;
  const int nl = 2 * nlag + 1;
  const int j0 = blockIdx.z * kW;
  const int lag_lo = -nlag + j0;
  const double* s = S64 + (long long)src * ntp;
  const float* r = R + (long long)rec * nt;
  const int base = threadIdx.x * L;
  double acc[L];
#pragma unroll
  for (int q = 0; q < L; ++q) acc[q] = 0.0;
  for (int t0 = 0; t0 < nt; t0 += kC) {
    __syncthreads();
    for (int i = threadIdx.x; i < kC + kW + 2 * L; i += 64) {
      const long long t = (long long)t0 + lag_lo + i;
      r_c[i] = (t >= 0 && t < nt) ? (double)r[t] : 0.0;
    }
    __syncthreads();
    const double* sc = s + t0;
    double w[2 * L];
#pragma unroll
    for (int q = 0; q < L; ++q) w[q] = r_c[base + q];
    for (int i = 0; i < kC; i += L) {
#pragma unroll
      for (int q = 0; q < L; ++q) w[L + q] = r_c[i + base + L + q];
#pragma unroll
      for (int u = 0; u < L; ++u) {
        const double sv = sc[i + u];
#pragma unroll
        for (int q = 0; q < L; ++q) acc[q] = fma(sv, w[u + q], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < L; ++q) w[q] = w[L + q];
    }
  }
  float* out = pairs + ((long long)src * gridDim.x + rec) * nl;
#pragma unroll
  for (int q = 0; q < L; ++q) {
    const int j = j0 + base + q;
    if (j < nl) out[j] = (float)acc[q];
  }
}

// pairs (nsrc, nrec, nl) -> out (nsrc, nl, nrec); accumulate: out += value (FP32)
__global__ void k_lag_major(const float* pairs, int nrec, int nl, float* out, int accumulate) {
  __shared__ float tile[32][33];
  const int src = blockIdx.z;
  const int j0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const float* p = pairs + (long long)src * nrec * nl;
  float* o = out + (long long)src * nl * nrec;
  for (int y = threadIdx.y; y < 32; y += blockDim.y) {
    const int rec = r0 + y, j = j0 + threadIdx.x;
    tile[y][threadIdx.x] = (rec < nrec && j < nl) ? p[(long long)rec * nl + j] : 0.0f;
  }
  __syncthreads();
  for (int y = threadIdx.y; y < 32; y += blockDim.y) {
    const int j = j0 + y, rec = r0 + threadIdx.x;
    if (j < nl && rec < nrec) {
      const float v = tile[threadIdx.x][y];
      float* d = o + (long long)j * nrec + rec;
      *d = accumulate ? *d + v : v;
    }
  }
}

// np.gradient(f, dt, axis=0) of float32 f (nl, ncol) with edge_order 1:
// interior (f[i+1] - f[i-1]) / fl32(2 dt), edges one-sided over fl32(dt)
__global__ void k_gradient(const float* f, int nl, int ncol, float two_dt, float dt1, float* out) {
  const long long n = (long long)nl * ncol;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / ncol);
    const int c = (int)(e - (long long)i * ncol);
    float v;
    if (nl < 2) v = 0.0f;
    else if (i == 0) v = (f[(long long)ncol + c] - f[c]) / dt1;
    else if (i == nl - 1) v = (f[e] - f[e - ncol]) / dt1;
    else v = (f[e + ncol] - f[e - ncol]) / two_dt;
    out[e] = v;
  }
}

// Impulse response of irfft(i w rfft(., n), n): h[p] = -(2/n) sum_{k=1}^{n/2-1}
// w_k sin(2 pi k p / n), w_k = 2 pi k / (n dt) (the Nyquist term is imaginary
// and irfft drops it).  The sine argument is reduced exactly in integers.
__global__ void k_dspec(double* h, int n, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double acc = 0.0;
  for (int k = 1; k < n / 2; ++k) {
    const long long m = (2LL * k * p) % (2LL * n);
    const double w = 2.0 * M_PI * (double)k / ((double)n * dt);
    acc += w * sinpi((double)m / (double)n);
  }
  h[p] = -2.0 / (double)n * acc;
}

// d[m] = sum_j c[j] h[(m - j) mod n] for m < nl, per column (FP64 accumulation)
__global__ void k_dspec_apply(const float* c, int nl, int ncol, const double* h, int n,
                              float* out) {
  const int col = blockIdx.y * blockDim.x + threadIdx.x;
  const int m = blockIdx.x;
  if (col >= ncol) return;
  double acc = 0.0;
  for (int j = 0; j < nl; ++j) {
    int q = (m - j) % n;
    if (q < 0) q += n;
    acc += (double)c[(long long)j * ncol + col] * h[q];
  }
  out[(long long)m * ncol + col] = (float)acc;
}

// branch 0 causal d[z + i], 1 anticausal d[z - i], 2 symmetric 0.5 (d[z+i] + d[z-i])
__global__ void k_branch(const float* d, int ncol, int z, int branch, int nout, float* out) {
  const long long n = (long long)nout * ncol;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / ncol);
    const int c = (int)(e - (long long)i * ncol);
    float v;
    if (branch == 0) v = d[(long long)(z + i) * ncol + c];
    else if (branch == 1) v = d[(long long)(z - i) * ncol + c];
    else v = 0.5f * (d[(long long)(z + i) * ncol + c] + d[(long long)(z - i) * ncol + c]);
    out[e] = v;
  }
}

// scipy interp1d(kind='linear', bounds_error=False, fill_value=0) of g
// (nt_orig, ncol) sampled at t_i = i dt_orig, evaluated at t_j = j dt_sim:
// idx = searchsorted(t, x) clipped to [1, nt_orig-1], slope =
// fl32(y_hi - y_lo) / (t_hi - t_lo) in FP64, y = slope (x - t_lo) + y_lo,
// rounded to FP32; 0 outside [t_0, t_last].
__global__ void k_resample(const float* g, int nt_orig, int ncol, double dt_orig, double dt_sim,
                           int target_nt, float* out) {
  const long long n = (long long)target_nt * ncol;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(e / ncol);
    const int c = (int)(e - (long long)j * ncol);
    const double x = (double)j * dt_sim;
    const double xlast = (double)(nt_orig - 1) * dt_orig;
    if (x < 0.0 || x > xlast) {
      out[e] = 0.0f;
      continue;
    }
    long long i = (long long)floor(x / dt_orig) - 1;
    if (i < 0) i = 0;
    while (i < nt_orig && (double)i * dt_orig < x) ++i;  // first t_i >= x
    if (i < 1) i = 1;
    if (i > nt_orig - 1) i = nt_orig - 1;
    const double tlo = (double)(i - 1) * dt_orig, thi = (double)i * dt_orig;
    const float ylo = g[(i - 1) * ncol + c], yhi = g[i * ncol + c];
    const double slope = (double)(yhi - ylo) / (thi - tlo);
    out[e] = (float)(slope * (x - tlo) + (double)ylo);
  }
}

static unsigned grid1(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)std::max<long long>(1, std::min<long long>(b, 65536));
}

// Device areas of one correlation call: the channel list, the source and
// receiver columns (ncol, nt), the pair-major lag sums and the FP64 source rows.
struct XcorrAreas {
  int* chs;
  float* S;
  float* R;
  float* pairs;
  double* S64;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of each area, in the order they are carved from one block.
static void xcorr_area_bytes(int nt, int nsrc, int nrec, int nlag, size_t b[5]) {
  const size_t nl = 2 * (size_t)nlag + 1;
  b[0] = (size_t)(nsrc + nrec) * sizeof(int);
  b[1] = (size_t)nsrc * nt * sizeof(float);
  b[2] = (size_t)nrec * nt * sizeof(float);
  b[3] = (size_t)nsrc * nrec * nl * sizeof(float);
  b[4] = (size_t)nsrc * ((size_t)nt + kXC) * sizeof(double);
}

static size_t xcorr_scratch_bytes(int nt, int nsrc, int nrec, int nlag) {
  size_t b[5], total = 0;
  xcorr_area_bytes(nt, nsrc, nrec, nlag, b);
  for (size_t v : b) total += align256(v);
  return total;
}

static XcorrAreas xcorr_carve(char* base, int nt, int nsrc, int nrec, int nlag) {
  size_t b[5];
  xcorr_area_bytes(nt, nsrc, nrec, nlag, b);
  char* q[5];
  for (int i = 0; i < 5; ++i) {
    q[i] = base;
    base += align256(b[i]);
  }
  return XcorrAreas{(int*)q[0], (float*)q[1], (float*)q[2], (float*)q[3], (double*)q[4]};
}

// One correlation call on the areas `a`; returns after the stream has
// drained, so the areas are free for the next call.
static int xcorr_run(const float* d_data, int nt, int nch, const int* src_ch, int nsrc,
                     const int* rec_ch, int nrec, int nlag, int accumulate, float* d_out,
                     const XcorrAreas& a, hipStream_t st) {
  const int nl = 2 * nlag + 1;
  std::vector<int> h(src_ch, src_ch + nsrc);
  h.insert(h.end(), rec_ch, rec_ch + nrec);
  TSW_HIP(hipMemcpyAsync(a.chs, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_take_cols, dim3((nt + 31) / 32, (nsrc + 31) / 32), dim3(32, 8), 0, st,
                     d_data, nt, nch, a.chs, nsrc, a.S);
  hipLaunchKernelGGL(k_take_cols, dim3((nt + 31) / 32, (nrec + 31) / 32), dim3(32, 8), 0, st,
                     d_data, nt, nch, a.chs + nsrc, nrec, a.R);
  // lags per lane: the smallest of 4 / 6 / 8 whose 64-lane tiles cover nl
  // with the least idle lanes
  int L = 8;
  if (nl <= 256) L = 4;
  else if (nl <= 384) L = 6;
  const unsigned tiles = (unsigned)((nl + 64 * L - 1) / (64 * L));
  // the source rows in FP64, padded with zeros so that the last chunk (it
  // starts below nt and spans at most kXC samples) stays inside its row
  const int ntp = nt + kXC;
  hipLaunchKernelGGL(k_rows_f64, dim3(grid1((long long)nsrc * ntp)), dim3(256), 0, st, a.S, nt,
                     ntp, nsrc, a.S64);
  const double* s64 = a.S64;
  if (L == 4)
    hipLaunchKernelGGL(k_xcorr<4>, dim3(nrec, nsrc, tiles), dim3(64), 0, st, s64, ntp, a.R, nt,
                       nlag, nrec, a.pairs);
  else if (L == 6)
    hipLaunchKernelGGL(k_xcorr<6>, dim3(nrec, nsrc, tiles), dim3(64), 0, st, s64, ntp, a.R, nt,
                       nlag, nrec, a.pairs);
  else
    hipLaunchKernelGGL(k_xcorr<8>, dim3(nrec, nsrc, tiles), dim3(64), 0, st, s64, ntp, a.R, nt,
                       nlag, nrec, a.pairs);
  hipLaunchKernelGGL(k_lag_major, dim3((nl + 31) / 32, (nrec + 31) / 32, nsrc), dim3(32, 8), 0,
                     st, a.pairs, nrec, nl, d_out, accumulate);
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

static int xcorr_device(const float* d_data, int nt, int nch, const int* src_ch, int nsrc,
                        const int* rec_ch, int nrec, int nlag, int accumulate, float* d_out,
                        hipStream_t st) {
  // scratch kept across calls (grown as needed, held until the process ends);
  // the entry points hold the API lock and drain the stream before returning
  static DevBuf* scr = new DevBuf();
  TSW_TRY(scr->ensure(xcorr_scratch_bytes(nt, nsrc, nrec, nlag)));
  const XcorrAreas a = xcorr_carve((char*)scr->p, nt, nsrc, nrec, nlag);
  return xcorr_run(d_data, nt, nch, src_ch, nsrc, rec_ch, nrec, nlag, accumulate, d_out, a, st);
}

// Fixed scratch of the allocation-free entry point: one array per device,
// held from the library's load until the process ends.  Host code never uses
// the name `scratch` as an address; it uses the device address that
// hipGetSymbolAddress returns for the current device.
constexpr size_t kXcorrScratchMax = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kXcorrScratchMax];
constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices];

// Looks up (once per device, with that device current) the device address of
// this device's `scratch`.
int xcorr_scratch_setup(void** out) {
  int dev = 0;
  TSW_HIP(hipGetDevice(&dev));
  TSW_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev,
          kMaxDevices);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    TSW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TSW_ARG(p != nullptr, "no device address for the correlation scratch");
    g_scratch_dev[dev] = p;
  }
  if (out) *out = g_scratch_dev[dev];
  return kOk;
}

static int xcorr_args(int nt, int nch, const int* src_ch, int nsrc, const int* rec_ch, int nrec,
                      int nlag) {
  TSW_ARG(nt >= 1 && nch >= 1 && nsrc >= 1 && nrec >= 1 && nlag >= 0, "bad sizes");
  TSW_ARG(nsrc <= 65535, "at most 65535 source channels per call");
  TSW_ARG(src_ch && rec_ch, "null channel list");
  for (int i = 0; i < nsrc; ++i)
    TSW_ARG(src_ch[i] >= 0 && src_ch[i] < nch, "source channel %d outside [0, %d)", src_ch[i],
            nch);
  for (int i = 0; i < nrec; ++i)
    TSW_ARG(rec_ch[i] >= 0 && rec_ch[i] < nch, "receiver channel %d outside [0, %d)", rec_ch[i],
            nch);
  // the reference slices nlag either side of the centre of a 2^ceil(log2(2nt-1))
  // transform; beyond half of it the slice no longer holds the lag window
  long long nfft = 1;
  while (nfft < 2LL * nt - 1) nfft <<= 1;
  const long long c = nfft / 2, room = std::min(c, nfft - c - 1);
  TSW_ARG(nlag <= room, "nlag %d exceeds the correlation window (%lld)", nlag, room);
  return kOk;
}

}  // namespace tsw

using namespace tsw;

extern "C" {

int tsw_xcorr(const float* data, int nt, int nch, const int* src_ch, int nsrc, const int* rec_ch,
              int nrec, int nlag, int accumulate, float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_TRY(fail_injected());
  TSW_ARG(data && out, "null argument");
  TSW_TRY(xcorr_args(nt, nch, src_ch, nsrc, rec_ch, nrec, nlag));
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const size_t nout = (size_t)nsrc * (2 * nlag + 1) * nrec;
  DevBuf d, o;
  TSW_TRY(d.ensure((size_t)nt * nch * sizeof(float)));
  TSW_TRY(o.ensure(nout * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(d.p, data, (size_t)nt * nch * sizeof(float), hipMemcpyHostToDevice, st));
  if (accumulate)
    TSW_HIP(hipMemcpyAsync(o.p, out, nout * sizeof(float), hipMemcpyHostToDevice, st));
  TSW_TRY(xcorr_device(d.f(), nt, nch, src_ch, nsrc, rec_ch, nrec, nlag, accumulate, o.f(), st));
  TSW_HIP(hipMemcpyAsync(out, o.p, nout * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_xcorr_device(const float* d_data, int nt, int nch, const int* src_ch, int nsrc,
                     const int* rec_ch, int nrec, int nlag, int accumulate, float* d_out,
                     void* stream) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(d_data && d_out, "null argument");
  TSW_TRY(xcorr_args(nt, nch, src_ch, nsrc, rec_ch, nrec, nlag));
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  return xcorr_device(d_data, nt, nch, src_ch, nsrc, rec_ch, nrec, nlag, accumulate, d_out,
                      static_cast<hipStream_t>(stream));
}

int tsw_xcorr_device_noalloc(const float* d_data, int nt, int nch, const int* src_ch, int nsrc,
                             const int* rec_ch, int nrec, int nlag, int accumulate, float* d_out,
                             void* stream) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(d_data && d_out, "null argument");
  TSW_TRY(xcorr_args(nt, nch, src_ch, nsrc, rec_ch, nrec, nlag));
  const size_t need = xcorr_scratch_bytes(nt, nsrc, nrec, nlag);
  TSW_ARG(need <= kXcorrScratchMax,
          "correlation needs %zu bytes of scratch, more than the %zu of the allocation-free "
          "entry point (use tsw_xcorr_device, or fewer channels or samples per call)",
          need, kXcorrScratchMax);
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  void* base = nullptr;
  TSW_TRY(xcorr_scratch_setup(&base));
  const XcorrAreas a = xcorr_carve((char*)base, nt, nsrc, nrec, nlag);
  return xcorr_run(d_data, nt, nch, src_ch, nsrc, rec_ch, nrec, nlag, accumulate, d_out, a,
                   static_cast<hipStream_t>(stream));
}

long long tsw_xcorr_noalloc_scratch_bytes(int nt, int nsrc, int nrec, int nlag) {
  if (nt < 1 || nsrc < 1 || nrec < 1 || nlag < 0) return -1;
  return (long long)xcorr_scratch_bytes(nt, nsrc, nrec, nlag);
}

long long tsw_xcorr_noalloc_scratch_max(void) { return (long long)kXcorrScratchMax; }

int tsw_time_derivative(const float* in, int nl, int ncol, double dt, int scheme, float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(in && out && nl >= 1 && ncol >= 1 && dt > 0, "bad arguments");
  TSW_ARG(scheme == 0 || scheme == 1, "scheme must be 0 (central) or 1 (spectral)");
  TSW_ARG(scheme == 1 || nl >= 2, "the central scheme needs at least 2 samples");
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  const size_t n = (size_t)nl * ncol;
  DevBuf a, b;
  TSW_TRY(a.ensure(n * sizeof(float)));
  TSW_TRY(b.ensure(n * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(a.p, in, n * sizeof(float), hipMemcpyHostToDevice, st));
  if (scheme == 0) {
    // numpy evaluates the divisors in double and rounds them to float32
    hipLaunchKernelGGL(k_gradient, dim3(grid1((long long)n)), dim3(256), 0, st, a.f(), nl, ncol,
                       (float)(2.0 * dt), (float)dt, b.f());
  } else {
    int nfft = 1;
    while (nfft < nl) nfft <<= 1;
    DevBuf h;
    TSW_TRY(h.ensure((size_t)nfft * sizeof(double)));
    hipLaunchKernelGGL(k_dspec, dim3((nfft + 255) / 256), dim3(256), 0, st, (double*)h.p, nfft,
                       dt);
    hipLaunchKernelGGL(k_dspec_apply, dim3(nl, (ncol + 63) / 64), dim3(64), 0, st, a.f(), nl,
                       ncol, (const double*)h.p, nfft, b.f());
    TSW_HIP(hipGetLastError());
    TSW_HIP(hipStreamSynchronize(st));
  }
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(out, b.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_green_branch(const float* in, int nl, int ncol, int zero_idx, int branch, float* out,
                     int* nout) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(in && out && nout && nl >= 1 && ncol >= 1, "bad arguments");
  TSW_ARG(zero_idx >= 0 && zero_idx < nl, "zero index outside the lag axis");
  TSW_ARG(branch >= 0 && branch <= 2, "branch must be 0 causal, 1 anticausal, 2 symmetric");
  const int m = branch == 0 ? nl - zero_idx
                            : branch == 1 ? zero_idx + 1 : std::min(nl - zero_idx, zero_idx + 1);
  TSW_ARG(*nout == m, "output holds %d rows, the branch has %d", *nout, m);
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  DevBuf a, b;
  TSW_TRY(a.ensure((size_t)nl * ncol * sizeof(float)));
  TSW_TRY(b.ensure((size_t)m * ncol * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(a.p, in, (size_t)nl * ncol * sizeof(float), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_branch, dim3(grid1((long long)m * ncol)), dim3(256), 0, st, a.f(), ncol,
                     zero_idx, branch, m, b.f());
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(out, b.p, (size_t)m * ncol * sizeof(float), hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

int tsw_resample_linear(const float* in, int nt_orig, int ncol, double dt_orig, double dt_sim,
                        int target_nt, float* out) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  TSW_ARG(in && out && ncol >= 1 && target_nt >= 0, "bad arguments");
  TSW_ARG(nt_orig >= 2, "linear interpolation needs at least 2 samples");
  TSW_ARG(dt_orig > 0 && dt_sim > 0, "time steps must be positive");
  if (target_nt == 0) return kOk;
  Ctx* ctx;
  TSW_TRY(ctx_get(&ctx));
  hipStream_t st = ctx->stream;
  DevBuf a, b;
  TSW_TRY(a.ensure((size_t)nt_orig * ncol * sizeof(float)));
  TSW_TRY(b.ensure((size_t)target_nt * ncol * sizeof(float)));
  TSW_HIP(hipMemcpyAsync(a.p, in, (size_t)nt_orig * ncol * sizeof(float), hipMemcpyHostToDevice,
                         st));
  hipLaunchKernelGGL(k_resample, dim3(grid1((long long)target_nt * ncol)), dim3(256), 0, st,
                     a.f(), nt_orig, ncol, dt_orig, dt_sim, target_nt, b.f());
  TSW_HIP(hipGetLastError());
  TSW_HIP(hipMemcpyAsync(out, b.p, (size_t)target_nt * ncol * sizeof(float),
                         hipMemcpyDeviceToHost, st));
  TSW_HIP(hipStreamSynchronize(st));
  return kOk;
}

}  // extern "C"
