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
// Adjoint (exact transpose, backward in time), in gather form: with
// A^m = dJ/du^m, Lam^m = dJ/dP^m = A^m + R^T r_{m-1} and
// w^m = S.Q^T Lam^m on I (0 elsewhere),
//   A^n = 2 w^{n+1} + L^T (C.w^{n+1}) - w^{n+2} + [Q^T Lam^{n+3} on E]
//   g_C += w^{n+1} . (L u^n) on I,      dJ/dv = g_C . 2 v dt^2,
// so each step reads Lam^{n+1} (stencil), Lam^{n+2} and, on the border cells
// only, Lam^{n+3}, and writes Lam^n: four buffers rotate with period 4.  L is
// symmetric (central stencil), so L^T is the same stencil applied to the
// zero-extended field C.w, which each block stages in LDS with its halo
// (with w and S, which the block's own cells then read from LDS).
//
// The forward levels u^n are not stored: every ceil(sqrt(nt)) steps the three
// levels (u^{n-2}, u^{n-1}, u^n) are checkpointed, and the backward sweep
// recomputes each segment from its checkpoint (bit-identical to the first
// pass) before running that segment's adjoint steps.  The recompute keeps,
// for every step of the segment, L u^n on the interior (the Laplacian the
// step computes anyway) rather than u^n: the gradient term then reads one
// value per cell instead of a 9-point stencil.
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <vector>

#include "engine.h"
#include "tb.h"
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

// Per-shot constants of a batched launch (blockIdx.z = shot within a group).
struct FwiMeta {
  const void* src;   // (nt) source time function
  const void* res;   // (nt, nrec) residual
  void* pred;        // (nt, nrec) predicted data
  void* probe;       // (nt) adjoint at the source cell
  const int* roff;   // receiver CSR over columns (nx + 1)
  const int* ridx;
  int src_cell, riz, nrec, pad;
};

template <typename T>
struct FwdArgs {
  int nx, nz, fs;
  T rdx2, rdz2, inj;
  long long n;       // cells per shot: shot j's level is base + j * n
  const T* C;
  const T* S;
  const T* cur;
  const T* prv;
  T* nxt;
  T* save;           // non-null: save + j * n = L cur (L u^n) on the interior, 0 elsewhere
  const FwiMeta* meta;
  int src_it, rec_it;  // -1: none
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

// Blocks cover 64 columns x kTZ rows (256 threads, 4 rows each); the field
// read through the stencil is staged in LDS with a 2-cell halo.
constexpr int kTZ = 16;
constexpr int kLW = 64 + 4;
constexpr int kLH = kTZ + 4;

template <typename T>
__device__ __forceinline__ T lap4_lds(const T* t, int l, T rdx2, T rdz2) {
  const T c0 = FwiK<T>::c0, c1 = fwi_c1<T>(), c2 = fwi_c2<T>();
  const T pc = t[l];
  T lx = c2 * (t[l - 2] + t[l + 2]);
  lx = lx + c1 * (t[l - 1] + t[l + 1]);
  lx = (lx + c0 * pc) * rdx2;
  T lz = c2 * (t[l - 2 * kLW] + t[l + 2 * kLW]);
  lz = lz + c1 * (t[l - kLW] + t[l + kLW]);
  lz = (lz + c0 * pc) * rdz2;
  return lx + lz;
}

template <typename T>
__device__ __forceinline__ void fwd_io(const FwdArgs<T>& s, const FwiMeta& m, int iz, int ix,
                                       T& v, bool& touched) {
  if (s.rec_it >= 0 && iz == m.riz) {
    const int a = m.roff[ix], b = m.roff[ix + 1];
    T* rec = (T*)m.pred + (long long)s.rec_it * m.nrec;
    for (int r = a; r < b; ++r) rec[m.ridx[r]] = v;
  }
  if (s.src_it >= 0 && iz * s.nx + ix == m.src_cell) {
    v = v + ((const T*)m.src)[s.src_it] * s.inj;
    touched = true;
  }
}

// One forward step of every shot of a group (same operation order as the
// FP32 production kernel, so FP32 data are bit-identical to it).
template <typename T>
__global__ __launch_bounds__(256) void k_fwi_fwd(FwdArgs<T> s) {
  __shared__ T tile[kLH * kLW];
  const int nx = s.nx, nz = s.nz;
  const int x0 = blockIdx.x * 64, z0 = blockIdx.y * kTZ;
  const long long o = (long long)blockIdx.z * s.n;
  const T* cur = s.cur + o;
  for (int i = threadIdx.y * 64 + threadIdx.x; i < kLH * kLW; i += 256) {
    const int gz = z0 - 2 + i / kLW, gx = x0 - 2 + i % kLW;
    tile[i] = (gz >= 0 && gz < nz && gx >= 0 && gx < nx) ? cur[gz * nx + gx] : T(0);
  }
  __syncthreads();
  const int ix = x0 + threadIdx.x;
  if (ix >= nx) return;
  const FwiMeta& m = s.meta[blockIdx.z];
  T* nxt = s.nxt + o;
  for (int k = 0; k < kTZ / 4; ++k) {
    const int lz = threadIdx.y + 4 * k, iz = z0 + lz;
    if (iz >= nz) break;
    const int c = iz * nx + ix;
    const int l = (lz + 2) * kLW + threadIdx.x + 2;
    const T pc = tile[l];
    const bool interior = iz >= 2 && iz < nz - 2 && ix >= 2 && ix < nx - 2;
    // the gradient needs L u^n on the interior: saved instead of u^n (0 elsewhere)
    const T lap = interior ? lap4_lds(tile, l, s.rdx2, s.rdz2) : T(0);
    if (s.save) s.save[o + c] = lap;
    if (s.fs && iz < 2) continue;
    T v;
    if (interior) {
      v = (T(2) * pc - s.prv[o + c]) + s.C[c] * lap;
      v = v * s.S[c];
    } else {
      v = nxt[c];
    }
    bool t0 = false, t1 = false, t2 = false;
    if (s.fs && iz == 2) {
      T v0 = T(0);
      fwd_io(s, m, 0, ix, v0, t0);
      nxt[ix] = v0;
      T v1 = -v;
      fwd_io(s, m, 1, ix, v1, t1);
      nxt[nx + ix] = v1;
    }
    fwd_io(s, m, iz, ix, v, t2);
    if (interior || t2) nxt[c] = v;
  }
}

// iteration-0 injection: level[j][src_cell] += src_j[0] * inj, one thread per shot
template <typename T>
__global__ void k_fwi_inject0(T* level, long long n, const FwiMeta* meta, int count, T inj) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  T* p = level + j * n + meta[j].src_cell;
  *p = *p + ((const T*)meta[j].src)[0] * inj;
}

template <typename T>
struct AdjArgs {
  int nx, nz, fs;
  T rdx2, rdz2;
  long long n;
  const T* C;
  const T* S;            // sponge per cell (used when `table` is null)
  const float* table;    // FP32: sponge factor by depth, computed per cell instead of read
  int n_sponge;
  const T* lam1;  // Lam^{n+1}
  const T* lam2;  // Lam^{n+2}
  const T* lam3;  // Lam^{n+3}
  T* lam0;        // Lam^n (written)
  const T* lun;   // forward L u^n (the Laplacian the forward step used; saved by the recompute)
  T* grad;        // g_C, one slice per shot of the group
  const FwiMeta* meta;
  int res_it;     // residual row added to Lam^n (n-1), -1: none
  int probe_it;
  // two-step kernel: Lam^{n-1} too
  T* lamm1;       // Lam^{n-1} (written)
  const T* lum1;  // forward L u^{n-1}
  int res_it2;    // residual row added to Lam^{n-1} (n-2), -1: none
};

template <typename T>
__device__ __forceinline__ bool fwi_interior(int iz, int ix, int nz, int nx) {
  return iz >= 2 && iz < nz - 2 && ix >= 2 && ix < nx - 2;
}

// (Q^T lam) at cell c = (iz, ix): the free surface zeroes rows 0/1 and
// mirrors row 1 into row 2
template <typename T>
__device__ __forceinline__ T fwi_zeta(int fs, const T* lam, int c, int iz, int nx) {
  if (!fs) return lam[c];
  if (iz < 2) return T(0);
  if (iz == 2) return lam[c] - lam[c - nx];
  return lam[c];
}

template <typename T>
__device__ __forceinline__ T fwi_sponge(const AdjArgs<T>& a, int iz, int ix, int c) {
  if (a.table == nullptr) return a.S[c];
  const int d = sponge_depth(ix, iz, a.nx, a.nz, a.n_sponge, a.fs);
  return d > 0 ? (T)a.table[d] : T(1);
}

// An empty asm that takes and gives back `v` in a register: arithmetic on `v`
// after it cannot be moved up next to the read that produced it.
template <typename T>
__device__ __forceinline__ void opaque(T& v) {
  asm volatile("" : "+v"(v));
}

template <typename T, int TZ>
__global__ __launch_bounds__(256) void k_fwi_adj(AdjArgs<T> a) {
  constexpr int LH = TZ + 4;
  constexpr int NL = LH * kLW;
  constexpr int NST = (NL + 255) / 256;  // staging passes of the 256 threads
  __shared__ T cw[NL];  // C.w^{n+1}, zero outside I
  __shared__ T wl[NL];  // w^{n+1} = S.zeta(Lam^{n+1})
  __shared__ T sl[NL];  // S
  const int nx = a.nx, nz = a.nz;
  const int x0 = blockIdx.x * 64, z0 = blockIdx.y * TZ;
  const long long o = (long long)blockIdx.z * a.n;
  const T* lam1 = a.lam1 + o;
  const int tid = threadIdx.y * 64 + threadIdx.x;
  // Staging: every pass's reads are issued before any is used, so the passes
  // share one memory latency instead of one each.
  T cv[NST], lv[NST], sv[NST];
#pragma unroll
  for (int p = 0; p < NST; ++p) {
    const int i = tid + 256 * p;
    const int gz = z0 - 2 + i / kLW, gx = x0 - 2 + i % kLW;
    cv[p] = lv[p] = sv[p] = T(0);
    if (i < NL && fwi_interior<T>(gz, gx, nz, nx)) {
      const int c = gz * nx + gx;
      cv[p] = a.C[c];
      lv[p] = lam1[c];
      // table[0] = 1, so the table is read for every depth
      sv[p] = a.table ? (T)a.table[sponge_depth(gx, gz, nx, nz, a.n_sponge, a.fs)] : a.S[c];
    }
  }
#pragma unroll
  for (int p = 0; p < NST; ++p) {
    opaque(cv[p]);
    opaque(lv[p]);
    opaque(sv[p]);
  }
#pragma unroll
  for (int p = 0; p < NST; ++p) {
    const int i = tid + 256 * p;
    if (i < NL) {
      const int gz = z0 - 2 + i / kLW, gx = x0 - 2 + i % kLW;
      // zeta (Q^T) on the interior: the free surface changes row 2 only
      T z = lv[p];
      if (a.fs && gz == 2 && fwi_interior<T>(gz, gx, nz, nx)) z = z - lam1[gz * nx + gx - nx];
      const T w = sv[p] * z;  // 0 outside I (all factors read as 0 there)
      cw[i] = cv[p] * w;
      wl[i] = w;
      sl[i] = sv[p];
    }
  }
  __syncthreads();
  const int ix = x0 + threadIdx.x;
  if (ix >= nx) return;
  const FwiMeta& m = a.meta[blockIdx.z];
  const T c0 = FwiK<T>::c0, c1 = fwi_c1<T>(), c2 = fwi_c2<T>();
  constexpr int KR = TZ / 4;  // rows per thread, 4 apart
  // the interior cells' remaining reads of all rows, issued together
  const T* lam2 = a.lam2 + o;
  const T* lun = a.lun + o;
  T* grad = a.grad + o;
  // (read for every row, the row clamped into the grid, so no row branches;
  // only interior cells use them)
  T l2[KR], g[KR], lu[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int iz = min(z0 + (int)threadIdx.y + 4 * k, nz - 1);
    const unsigned c = (unsigned)(iz * nx + ix);
    l2[k] = lam2[c];
    g[k] = grad[c];
    lu[k] = lun[c];
  }
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    opaque(l2[k]);
    opaque(g[k]);
    opaque(lu[k]);
  }
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lz = threadIdx.y + 4 * k, iz = z0 + lz;
    if (iz >= nz) break;
    const int c = iz * nx + ix;
    const int l = (lz + 2) * kLW + threadIdx.x + 2;
    // L^T (C.w) at c
    T lx = c2 * (cw[l - 2] + cw[l + 2]);
    lx = lx + c1 * (cw[l - 1] + cw[l + 1]);
    lx = (lx + c0 * cw[l]) * a.rdx2;
    T lz2 = c2 * (cw[l - 2 * kLW] + cw[l + 2 * kLW]);
    lz2 = lz2 + c1 * (cw[l - kLW] + cw[l + kLW]);
    lz2 = (lz2 + c0 * cw[l]) * a.rdz2;
    T acc;
    if (fwi_interior<T>(iz, ix, nz, nx)) {
      const T w1 = wl[l];
      T z2 = l2[k];
      if (a.fs && iz == 2) z2 = z2 - lam2[c - nx];
      const T w2 = sl[l] * z2;
      acc = (T(2) * w1 + (lx + lz2)) - w2;
      grad[c] = g[k] + w1 * lu[k];
    } else {
      acc = lx + lz2;
      if (!(a.fs && iz < 2)) acc = acc + fwi_zeta(a.fs, a.lam3 + o, c, iz, nx);
    }
    // probe: A^n = dJ/du^n at the source cell (before the receiver term,
    // which belongs to dJ/dP^n)
    if (c == m.src_cell) ((T*)m.probe)[a.probe_it] = acc;
    if (a.res_it >= 0 && iz == m.riz) {
      const int r0 = m.roff[ix], r1 = m.roff[ix + 1];
      const T* row = (const T*)m.res + (long long)a.res_it * m.nrec;
      for (int r = r0; r < r1; ++r) acc = acc + row[m.ridx[r]];
    }
    a.lam0[o + c] = acc;
  }
}

// Two adjoint steps in one launch: Lam^n on the tile and a 2-cell halo, then
// Lam^{n-1} on the tile.  The step formulas, their FP order and the order of
// the two gradient additions are those of k_fwi_adj applied at n then n-1, so
// the result is bit-identical to two single steps.  Reads Lam^{n+1} (tile +
// 4), Lam^{n+2}, Lam^{n+3} (border cells), L u^n, L u^{n-1}; writes Lam^n,
// Lam^{n-1} into buffers no workgroup of the launch reads (five rotate).
template <typename T, int TZ>
__global__ __launch_bounds__(256) void k_fwi_adj2(AdjArgs<T> a) {
  constexpr int W1 = 64 + 4, H1 = TZ + 4;  // tile + 2
  constexpr int W2 = 64 + 8, H2 = TZ + 8;  // tile + 4
  __shared__ T cw1[H2 * W2];               // C.w^{n+1}
  __shared__ T ln[H1 * W1];                // Lam^n
  __shared__ T cw0[H1 * W1];               // C.w^n
  const int nx = a.nx, nz = a.nz;
  const int x0 = blockIdx.x * 64, z0 = blockIdx.y * TZ;
  const long long o = (long long)blockIdx.z * a.n;
  const T* lam1 = a.lam1 + o;
  const T* lam2 = a.lam2 + o;
  const FwiMeta& m = a.meta[blockIdx.z];
  const T c0 = FwiK<T>::c0, c1 = fwi_c1<T>(), c2 = fwi_c2<T>();
  const int tid = threadIdx.y * 64 + threadIdx.x;
  for (int i = tid; i < H2 * W2; i += 256) {
    const int gz = z0 - 4 + i / W2, gx = x0 - 4 + i % W2;
    T v = T(0);
    if (fwi_interior<T>(gz, gx, nz, nx)) {
      const int c = gz * nx + gx;
      v = a.C[c] * (fwi_sponge(a, gz, gx, c) * fwi_zeta(a.fs, lam1, c, gz, nx));
    }
    cw1[i] = v;
  }
  __syncthreads();
  // step n on tile + 2
  for (int i = tid; i < H1 * W1; i += 256) {
    const int lzz = i / W1, lxx = i % W1;
    const int gz = z0 - 2 + lzz, gx = x0 - 2 + lxx;
    if (gz < 0 || gz >= nz || gx < 0 || gx >= nx) {
      ln[i] = T(0);
      continue;
    }
    const int c = gz * nx + gx;
    const int l = (lzz + 2) * W2 + lxx + 2;
    T lx = c2 * (cw1[l - 2] + cw1[l + 2]);
    lx = lx + c1 * (cw1[l - 1] + cw1[l + 1]);
    lx = (lx + c0 * cw1[l]) * a.rdx2;
    T lz2 = c2 * (cw1[l - 2 * W2] + cw1[l + 2 * W2]);
    lz2 = lz2 + c1 * (cw1[l - W2] + cw1[l + W2]);
    lz2 = (lz2 + c0 * cw1[l]) * a.rdz2;
    T acc;
    if (fwi_interior<T>(gz, gx, nz, nx)) {
      const T sc = fwi_sponge(a, gz, gx, c);
      const T w1 = sc * fwi_zeta(a.fs, lam1, c, gz, nx);
      const T w2 = sc * fwi_zeta(a.fs, lam2, c, gz, nx);
      acc = (T(2) * w1 + (lx + lz2)) - w2;
    } else {
      acc = lx + lz2;
      if (!(a.fs && gz < 2)) acc = acc + fwi_zeta(a.fs, a.lam3 + o, c, gz, nx);
    }
    const bool core = lzz >= 2 && lzz < 2 + TZ && lxx >= 2 && lxx < 2 + 64;
    if (core && c == m.src_cell) ((T*)m.probe)[a.probe_it] = acc;
    if (a.res_it >= 0 && gz == m.riz) {
      const int r0 = m.roff[gx], r1 = m.roff[gx + 1];
      const T* row = (const T*)m.res + (long long)a.res_it * m.nrec;
      for (int r = r0; r < r1; ++r) acc = acc + row[m.ridx[r]];
    }
    ln[i] = acc;
    if (core) a.lam0[o + c] = acc;
  }
  __syncthreads();
  // C.w^n on tile + 2 from Lam^n (Q^T at row 2 reads row 1 of the same region)
  for (int i = tid; i < H1 * W1; i += 256) {
    const int gz = z0 - 2 + i / W1, gx = x0 - 2 + i % W1;
    T v = T(0);
    if (fwi_interior<T>(gz, gx, nz, nx)) {
      const int c = gz * nx + gx;
      T zt = ln[i];
      if (a.fs && gz == 2) zt = ln[i] - ln[i - W1];
      v = a.C[c] * (fwi_sponge(a, gz, gx, c) * zt);
    }
    cw0[i] = v;
  }
  __syncthreads();
  const int ix = x0 + threadIdx.x;
  if (ix >= nx) return;
  for (int k = 0; k < TZ / 4; ++k) {
    const int lzz = threadIdx.y + 4 * k, iz = z0 + lzz;
    if (iz >= nz) break;
    const int c = iz * nx + ix;
    const int l = (lzz + 2) * W1 + threadIdx.x + 2;
    T lx = c2 * (cw0[l - 2] + cw0[l + 2]);
    lx = lx + c1 * (cw0[l - 1] + cw0[l + 1]);
    lx = (lx + c0 * cw0[l]) * a.rdx2;
    T lz2 = c2 * (cw0[l - 2 * W1] + cw0[l + 2 * W1]);
    lz2 = lz2 + c1 * (cw0[l - W1] + cw0[l + W1]);
    lz2 = (lz2 + c0 * cw0[l]) * a.rdz2;
    T acc;
    if (fwi_interior<T>(iz, ix, nz, nx)) {
      const T sc = fwi_sponge(a, iz, ix, c);
      T zt = ln[l];
      if (a.fs && iz == 2) zt = ln[l] - ln[l - W1];
      const T w0 = sc * zt;
      const T wp = sc * fwi_zeta(a.fs, lam1, c, iz, nx);
      acc = (T(2) * w0 + (lx + lz2)) - wp;
      T g = a.grad[o + c];
      g = g + wp * a.lun[o + c];
      g = g + w0 * a.lum1[o + c];
      a.grad[o + c] = g;
    } else {
      acc = lx + lz2;
      if (!(a.fs && iz < 2)) acc = acc + fwi_zeta(a.fs, lam2, c, iz, nx);
    }
    if (c == m.src_cell) ((T*)m.probe)[a.probe_it - 1] = acc;
    if (a.res_it2 >= 0 && iz == m.riz) {
      const int r0 = m.roff[ix], r1 = m.roff[ix + 1];
      const T* row = (const T*)m.res + (long long)a.res_it2 * m.nrec;
      for (int r = r0; r < r1; ++r) acc = acc + row[m.ridx[r]];
    }
    a.lamm1[o + c] = acc;
  }
}

template <typename T>
__global__ void k_fwi_fill(T* p, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// mu^{nt} = R^T r_{nt-1}, shot j = blockIdx.y
template <typename T>
__global__ void k_fwi_seed(T* mu, long long n, int nx, const FwiMeta* meta, int row_it) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= nx) return;
  const FwiMeta& m = meta[blockIdx.y];
  const int r0 = m.roff[ix], r1 = m.roff[ix + 1];
  if (r0 == r1) return;
  const T* row = (const T*)m.res + (long long)row_it * m.nrec;
  T* p = mu + blockIdx.y * n + (long long)m.riz * nx + ix;
  T v = *p;
  for (int r = r0; r < r1; ++r) v = v + row[m.ridx[r]];
  *p = v;
}

template <typename T>
__global__ void k_fwi_coef(const T* v, T* C, long long n, T dt2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    C[i] = (v[i] * v[i]) * dt2;
}

// dJ/dv = (sum over the group slots of g_C) . 2 v dt^2, slots added in order
template <typename T>
__global__ void k_fwi_chain(const T* g, int slots, const T* v, T* out, long long n, T dt2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    T acc = g[i];
    for (int j = 1; j < slots; ++j) acc = acc + g[j * n + i];
    out[i] = acc * ((T(2) * v[i]) * dt2);
  }
}

constexpr int kSumBlocks = 512;

// res = pred - obs, and per-block partial sums of res^2 in double (fixed grid
// and tree order: deterministic)
template <typename T>
__global__ __launch_bounds__(256) void k_fwi_resid(const T* pred, const T* obs, T* res,
                                                   long long n, double* partial) {
  __shared__ double red[256];
  double acc = 0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const T r = pred[i] - obs[i];
    res[i] = r;
    acc += (double)r * (double)r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// loss[shot] = 1/2 sum of the partials, in order (one block)
__global__ void k_fwi_loss(const double* partial, double* loss) {
  __shared__ double red[256];
  double acc = 0;
  for (int i = threadIdx.x; i < kSumBlocks; i += 256) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = 0.5 * red[0];
}

// ---- Temporally blocked forward steps (FP32).  One launch advances every
// shot of a group by up to T steps.  A workgroup owns a core tile and loads it
// with a 2T-cell halo (clipped at the grid); each thread keeps its R rows'
// three time levels, coefficient and sponge factor in registers and exchanges
// neighbours through double-buffered LDS, one barrier per step.  After s steps
// the cells at least 2s from a cut tile edge are exact, so the core is exact
// after T steps.  The per-cell arithmetic is that of k_fwi_fwd<float> (the
// production FP32 association), so results are bit-identical to it.  The state
// at a block boundary (u^{n-2}, u^{n-1}, u^n) ping-pongs between two sets.
struct FwiTbArgs {
  int nx, nz, fs, n_sponge;
  float rdx2, rdz2, inj;
  const float* C;
  const float* table;        // sponge factor by depth
  long long n;               // cells per shot
  const float* rd[3];        // set read: u^{n-2}, u^{n-1}, u^n (shot j at + j * n)
  float* wr[3];              // set written
  float* save;               // non-null: save + (it - save_it0) * sstride + j * n = L u^it
  long long sstride;
  int save_it0;
  const FwiMeta* meta;
  int it0, nsub, nt, rec;    // rec: record the receivers (P^{it+1}, before injection)
  int tiles_x, tiles;        // tiles per shot
  int count;                 // shots: the grid is tiles * count workgroups
};

template <int T, int R, int WX, int WZ>
__global__ __launch_bounds__(64 * WX * WZ)
__attribute__((amdgpu_waves_per_eu(kTbWavesPerEu<R, WX, WZ>))) void k_fwi_tb(FwiTbArgs a) {
  using S = TbShape<T, R, WX, WZ>;
  __shared__ float lds[2][S::LR * S::LS];
  const int b = xcd_tile_tb(blockIdx.x, a.tiles * a.count);
  const int shot = b / a.tiles, t = b - shot * a.tiles;
  const FwiMeta& m = a.meta[shot];
  const long long o = (long long)shot * a.n;
  const int nx = a.nx, nz = a.nz;
  const int cx0 = (t % a.tiles_x) * S::W;
  const int cz0 = (t / a.tiles_x) * S::H;
  const int X0 = cx0 - 2 * T > 0 ? cx0 - 2 * T : 0;
  const int Z0 = cz0 - 2 * T > 0 ? cz0 - 2 * T : 0;
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int col = X0 + lx;
  const int row0 = Z0 + ly * R;
  const bool col_in = col < nx;
  const bool core_col = col_in && col >= cx0 && col < cx0 + S::W;
  const bool col_int = col >= 2 && col < nx - 2;
  unsigned m_in = 0, m_int = 0, m_core = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int row = row0 + q;
    if (col_in && row < nz) {
      m_in |= 1u << q;
      if (col_int && row >= 2 && row < nz - 2) m_int |= 1u << q;
      if (core_col && row >= cz0 && row < cz0 + S::H) m_core |= 1u << q;
    }
  }
  const unsigned m_fs = (a.fs && row0 == 0) ? (m_in & 3u) : 0u;
  const unsigned m_keep = m_in & ~m_int & ~m_fs;
  float u2[R], u1[R], c[R], cf[R], sp[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    u2[q] = u1[q] = c[q] = cf[q] = 0.0f;
    sp[q] = 1.0f;
    if (m_in >> q & 1u) {
      const int idx = (row0 + q) * nx + col;
      cf[q] = a.C[idx];
      const int d = sponge_depth(col, row0 + q, nx, nz, a.n_sponge, a.fs);
      if (d > 0) sp[q] = a.table[d];
      c[q] = a.rd[2][o + idx];
      u1[q] = a.rd[1][o + idx];
      if (m_keep >> q & 1u) u2[q] = a.rd[0][o + idx];
    }
  }
  float* X = lds[0];
#pragma unroll
  for (int q = 0; q < R; ++q) X[(ly * R + q + 2) * S::LS + lx + 2] = c[q];
  __syncthreads();
  const int src_row = m.src_cell / nx, src_col = m.src_cell - src_row * nx;
  const bool src_thread = src_col == col && src_row >= row0 && src_row < row0 + R;
  // receivers of this column recorded from this thread's rows (read once)
  int rec_r0 = 0, rec_r1 = 0;
  if (a.rec && m.riz >= row0 && m.riz < row0 + R && core_col) {
    rec_r0 = as_global(m.roff)[col];
    rec_r1 = as_global(m.roff)[col + 1];
  }
  for (int j = 0; j < a.nsub; ++j) {
    const int it = a.it0 + j;
    float* Y = lds[(j + 1) & 1];
    // the gradient needs L u^it on the interior: saved instead of u^it (0 elsewhere)
    float* sv = a.save ? a.save + (long long)(it - a.save_it0) * a.sstride + o : nullptr;
    float nv[R];
    const float* xr = X + 2;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int lr = (ly * R + q + 2) * S::LS + lx;
      // evaluated on every lane (all reads stay inside this LDS buffer), so the
      // reads of all rows can issue together
      const float pc = c[q];
      const float xm2 = xr[lr - 2], xm1 = xr[lr - 1], xp1 = xr[lr + 1], xp2 = xr[lr + 2];
      const float zm2 = q >= 2 ? c[q - 2] : xr[lr - 2 * S::LS];
      const float zm1 = q >= 1 ? c[q - 1] : xr[lr - S::LS];
      const float zp1 = q + 1 < R ? c[q + 1] : xr[lr + S::LS];
      const float zp2 = q + 2 < R ? c[q + 2] : xr[lr + 2 * S::LS];
      float lxx = TSW_C2 * (xm2 + xp2);
      lxx = lxx + TSW_C1 * (xm1 + xp1);
      lxx = (lxx + TSW_C0 * pc) * a.rdx2;
      float lzz = TSW_C2 * (zm2 + zp2);
      lzz = lzz + TSW_C1 * (zm1 + zp1);
      lzz = (lzz + TSW_C0 * pc) * a.rdz2;
      const float lap = lxx + lzz;
      if (sv && (m_core >> q & 1u))  // read once, by the adjoint sweep
        __builtin_nontemporal_store((m_int >> q & 1u) ? lap : 0.0f, sv + (row0 + q) * nx + col);
      float v = (2.0f * pc - u1[q]) + cf[q] * lap;
      v = v * sp[q];
      // cells the step never writes keep the level three buffers back
      nv[q] = (m_int >> q & 1u) ? v : u2[q];
    }
    if (m_fs) {  // rows 0, 1, 2 of this column belong to this thread
      nv[0] = 0.0f;
      nv[1] = -nv[2];
    }
    if (rec_r0 < rec_r1) {
      gfloat* rec = as_global((float*)m.pred) + (long long)it * m.nrec;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (row0 + q == m.riz && (m_core >> q & 1u))
          for (int r = rec_r0; r < rec_r1; ++r) rec[as_global(m.ridx)[r]] = nv[q];
    }
    if (src_thread && it + 1 < a.nt) {
      const float s = as_global((const float*)m.src)[it + 1];
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (row0 + q == src_row) nv[q] = nv[q] + s * a.inj;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      Y[(ly * R + q + 2) * S::LS + lx + 2] = nv[q];
      u2[q] = u1[q];
      u1[q] = c[q];
      c[q] = nv[q];
    }
    X = Y;
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (m_core >> q & 1u) {
      const int idx = (row0 + q) * nx + col;
      a.wr[2][o + idx] = c[q];
      a.wr[1][o + idx] = u1[q];
      a.wr[0][o + idx] = u2[q];
    }
  }
}

// Blocked configurations: steps per launch T, rows per thread R, wavefronts
// across WX and down WZ.
struct FwiTbCfg {
  int T, R, WX, WZ;
  void (*launch)(const FwiTbArgs&, int count, hipStream_t);
};
template <int T, int R, int WX, int WZ>
static void fwi_tb_launch(const FwiTbArgs& a, int count, hipStream_t st) {
  (void)count;
  hipLaunchKernelGGL((k_fwi_tb<T, R, WX, WZ>), dim3(a.tiles * a.count), dim3(64 * WX, WZ), 0,
                     st, a);
}
static const FwiTbCfg kFwiTb[] = {
    {4, 4, 1, 16, fwi_tb_launch<4, 4, 1, 16>}, {5, 4, 1, 16, fwi_tb_launch<5, 4, 1, 16>},
    {4, 8, 1, 8, fwi_tb_launch<4, 8, 1, 8>},   {4, 16, 1, 4, fwi_tb_launch<4, 16, 1, 4>},
    {3, 4, 1, 16, fwi_tb_launch<3, 4, 1, 16>}, {5, 8, 1, 8, fwi_tb_launch<5, 8, 1, 8>},
    {3, 8, 1, 8, fwi_tb_launch<3, 8, 1, 8>},   {4, 12, 1, 8, fwi_tb_launch<4, 12, 1, 8>},
    {6, 8, 1, 8, fwi_tb_launch<6, 8, 1, 8>},
};
constexpr int kFwiTbN = (int)(sizeof(kFwiTb) / sizeof(kFwiTb[0]));
constexpr int kFwiDefaultTb = 2;
constexpr int kFwiDefaultAdjTb = 2;

// Temporally blocked FP32 adjoint: nsub adjoint steps n = it0, it0-1, ... of
// every shot of a group in one launch, with the tiling of k_fwi_tb (2T halo).
// Each thread keeps Lam^{n+1}, Lam^{n+2}, Lam^{n+3}, C, S and its g_C in
// registers for its R rows; C.w^{n+1} moves through double-buffered LDS with
// one barrier per step.  Per step and cell it evaluates the formulas of
// k_fwi_adj in the same FP32 order (A^n, the probe, R^T r_{n-1}, the gradient
// term w^{n+1}.L u^n from the recompute's segment), so the gradient and the
// adjoint field are bit-identical to one launch per step.  The state at
// block boundaries ping-pongs between two sets, as in the forward.
struct FwiAdjTbArgs {
  int nx, nz, fs, n_sponge;
  float rdx2, rdz2;
  const float* C;
  const float* table;        // sponge factor by depth
  long long n;               // cells per shot
  const float* rd[3];        // set read: Lam^{n+1}, Lam^{n+2}, Lam^{n+3} (shot j at + j * n)
  float* wr[3];              // set written: the same for the next block
  const float* lu;           // segment of L u^it: lu + (it - lu_it0) * sstride + j * n
  long long sstride;
  int lu_it0;
  float* grad;               // g_C (shot j at + j * n)
  const FwiMeta* meta;
  int it0, nsub;
  int tiles_x, tiles;        // tiles per shot
  int count;                 // shots: the grid is tiles * count workgroups
};

template <int T, int R, int WX, int WZ>
__global__ __launch_bounds__(64 * WX * WZ)
__attribute__((amdgpu_waves_per_eu(kTbWavesPerEu<R, WX, WZ>))) void k_fwi_adj_tb(FwiAdjTbArgs a) {
  using S = TbShape<T, R, WX, WZ>;
  __shared__ float lds[2][S::LR * S::LS];
  const int b = xcd_tile_tb(blockIdx.x, a.tiles * a.count);
  const int shot = b / a.tiles, t = b - shot * a.tiles;
  const FwiMeta& m = a.meta[shot];
  const long long o = (long long)shot * a.n;
  const int nx = a.nx, nz = a.nz;
  const int cx0 = (t % a.tiles_x) * S::W;
  const int cz0 = (t / a.tiles_x) * S::H;
  const int X0 = cx0 - 2 * T > 0 ? cx0 - 2 * T : 0;
  const int Z0 = cz0 - 2 * T > 0 ? cz0 - 2 * T : 0;
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int col = X0 + lx;
  const int row0 = Z0 + ly * R;
  const bool col_in = col < nx;
  const bool core_col = col_in && col >= cx0 && col < cx0 + S::W;
  const bool col_int = col >= 2 && col < nx - 2;
  unsigned m_in = 0, m_int = 0, m_core = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int row = row0 + q;
    if (col_in && row < nz) {
      m_in |= 1u << q;
      if (col_int && row >= 2 && row < nz - 2) m_int |= 1u << q;
      if (core_col && row >= cz0 && row < cz0 + S::H) m_core |= 1u << q;
    }
  }
  // rows 0 and 1 under a free surface (their Q^T Lam is 0); row 2 is the
  // thread's row q = 2 exactly when row0 == 0, and its Q^T subtracts row 1
  const bool top = a.fs && row0 == 0;
  const unsigned m_out = m_in & ~m_int;  // in the grid, outside I
  float l1[R], l2[R], l3[R], cf[R], sp[R], g[R], cw[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    l1[q] = l2[q] = l3[q] = cf[q] = g[q] = 0.0f;
    sp[q] = 1.0f;
    if (m_in >> q & 1u) {
      const int idx = (row0 + q) * nx + col;
      cf[q] = a.C[idx];
      // table[0] = 1
      sp[q] = a.table[sponge_depth(col, row0 + q, nx, nz, a.n_sponge, a.fs)];
      l1[q] = a.rd[0][o + idx];
      l2[q] = a.rd[1][o + idx];
      if (m_out >> q & 1u) l3[q] = a.rd[2][o + idx];
      if ((m_core & m_int) >> q & 1u) g[q] = a.grad[o + idx];
    }
  }
  // C.w of a level on I (0 elsewhere): C * (S * zeta(Lam))
  auto cw_of = [&](const float* lv, int q) -> float {
    const float z = (top && q == 2) ? lv[2] - lv[1] : lv[q];
    return (m_int >> q & 1u) ? cf[q] * (sp[q] * z) : 0.0f;
  };
  // C.w is 0 outside the grid, and cells on the grid's edge read it there
  // (their A^n is L^T(C.w) + ...): the pad rows and columns of both LDS
  // buffers, which no thread writes, are zeroed once
  for (int i = ly * (64 * WX) + lx; i < 2 * S::LR * S::LS; i += S::THREADS) (&lds[0][0])[i] = 0.0f;
  __syncthreads();
  float* X = lds[0];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    cw[q] = cw_of(l1, q);
    X[(ly * R + q + 2) * S::LS + lx + 2] = cw[q];
  }
  __syncthreads();
  const float c0 = FwiK<float>::c0, c1 = fwi_c1<float>(), c2 = fwi_c2<float>();
  const int riz = m.riz, nrec = m.nrec;
  const bool rec_strip = riz >= row0 && riz < row0 + R && col_in;
  int rec_r0 = 0, rec_r1 = 0;
  if (rec_strip) {
    rec_r0 = as_global(m.roff)[col];
    rec_r1 = as_global(m.roff)[col + 1];
  }
  const int src_cell = m.src_cell;
  for (int j = 0; j < a.nsub; ++j) {
    const int it = a.it0 - j;
    float* Y = lds[(j + 1) & 1];
    const gfloat* lu = as_global(a.lu) + (long long)(it - a.lu_it0) * a.sstride + o;
    float lun[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const unsigned idx = (unsigned)((row0 + q) * nx + col);
      lun[q] = ((m_core & m_int) >> q & 1u) ? __builtin_nontemporal_load(lu + idx) : 0.0f;
    }
    float nv[R];
    const float* xr = X + 2;  // column offset of lx = 0
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int lr = (ly * R + q + 2) * S::LS + lx;
      // L^T (C.w^{n+1}), evaluated on every lane
      const float xm2 = xr[lr - 2], xm1 = xr[lr - 1], xp1 = xr[lr + 1], xp2 = xr[lr + 2];
      const float zm2 = q >= 2 ? cw[q - 2] : xr[lr - 2 * S::LS];
      const float zm1 = q >= 1 ? cw[q - 1] : xr[lr - S::LS];
      const float zp1 = q + 1 < R ? cw[q + 1] : xr[lr + S::LS];
      const float zp2 = q + 2 < R ? cw[q + 2] : xr[lr + 2 * S::LS];
      float lxx = c2 * (xm2 + xp2);
      lxx = lxx + c1 * (xm1 + xp1);
      lxx = (lxx + c0 * cw[q]) * a.rdx2;
      float lzz = c2 * (zm2 + zp2);
      lzz = lzz + c1 * (zm1 + zp1);
      lzz = (lzz + c0 * cw[q]) * a.rdz2;
      const float lt = lxx + lzz;
      float acc;
      if (m_int >> q & 1u) {
        const float z1 = (top && q == 2) ? l1[2] - l1[1] : l1[q];
        const float z2 = (top && q == 2) ? l2[2] - l2[1] : l2[q];
        const float w1 = sp[q] * z1;
        const float w2 = sp[q] * z2;
        acc = (2.0f * w1 + lt) - w2;
        if (m_core >> q & 1u) g[q] = g[q] + w1 * lun[q];
      } else if (m_in >> q & 1u) {
        acc = lt;
        if (!(top && q < 2)) acc = acc + ((top && q == 2) ? l3[2] - l3[1] : l3[q]);
      } else {
        acc = 0.0f;  // outside the grid
      }
      nv[q] = acc;
    }
    // probe: A^n at the source cell, before the receiver term
    if (m_core) {
#pragma unroll
      for (int q = 0; q < R; ++q)
        if ((m_core >> q & 1u) && (row0 + q) * nx + col == src_cell)
          as_global((float*)m.probe)[it] = nv[q];
    }
    // Lam^n = A^n + R^T r_{n-1}, receivers in order
    if (it >= 1 && rec_r0 < rec_r1) {
      const gfloat* row = as_global((const float*)m.res) + (long long)(it - 1) * nrec;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (row0 + q == riz)
          for (int r = rec_r0; r < rec_r1; ++r) nv[q] = nv[q] + row[as_global(m.ridx)[r]];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      l3[q] = l2[q];
      l2[q] = l1[q];
      l1[q] = nv[q];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cw[q] = cw_of(l1, q);
      Y[(ly * R + q + 2) * S::LS + lx + 2] = cw[q];
    }
    X = Y;
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (m_core >> q & 1u) {
      const int idx = (row0 + q) * nx + col;
      a.wr[0][o + idx] = l1[q];
      a.wr[1][o + idx] = l2[q];
      a.wr[2][o + idx] = l3[q];
      if (m_int >> q & 1u) a.grad[o + idx] = g[q];
    }
  }
}

struct FwiAdjTbCfg {
  int T, R, WX, WZ;
  void (*launch)(const FwiAdjTbArgs&, hipStream_t);
};
template <int T, int R, int WX, int WZ>
static void fwi_adj_tb_launch(const FwiAdjTbArgs& a, hipStream_t st) {
  hipLaunchKernelGGL((k_fwi_adj_tb<T, R, WX, WZ>), dim3(a.tiles * a.count), dim3(64 * WX, WZ),
                     0, st, a);
}
// the shapes of kFwiTb, in the same order
static const FwiAdjTbCfg kFwiAdjTb[] = {
    {4, 4, 1, 16, fwi_adj_tb_launch<4, 4, 1, 16>}, {5, 4, 1, 16, fwi_adj_tb_launch<5, 4, 1, 16>},
    {4, 8, 1, 8, fwi_adj_tb_launch<4, 8, 1, 8>},   {4, 16, 1, 4, fwi_adj_tb_launch<4, 16, 1, 4>},
    {3, 4, 1, 16, fwi_adj_tb_launch<3, 4, 1, 16>}, {5, 8, 1, 8, fwi_adj_tb_launch<5, 8, 1, 8>},
    {3, 8, 1, 8, fwi_adj_tb_launch<3, 8, 1, 8>},   {4, 12, 1, 8, fwi_adj_tb_launch<4, 12, 1, 8>},
    {6, 8, 1, 8, fwi_adj_tb_launch<6, 8, 1, 8>},
};
static_assert(sizeof(kFwiAdjTb) / sizeof(kFwiAdjTb[0]) == kFwiTbN, "one adjoint shape per forward shape");

struct FwiShotHost {
  int nt, six, siz, riz, nrec;
  DevBuf src, obs, pred, res, roff, ridx, probe;
};

// Shots of equal nt advance together, one launch per step.
struct FwiGroup {
  int nt = 0;
  std::vector<int> ids;
  DevBuf meta;  // FwiMeta[count]
  DevBuf ckpt;  // [nck][3][count][n]
  int count() const { return (int)ids.size(); }
};

struct Fwi {
  int nx = 0, nz = 0, fs = 1, prec = 4, kint = 1, n_sponge = 0, cmax = 1;
  double dx = 0, dz = 0, dt = 0, alpha = 0;
  std::vector<FwiShotHost> shots;
  std::vector<FwiGroup> groups;
  DevBuf S, C, buf, mu, seg, gC, partial, loss;
  int tb = 0;          // FP32 blocked forward: configuration kFwiTb[tb - 1]; 0 = per step
  int adj_steps = 1;   // adjoint steps per launch (1 or 2; 2 measured slower on MI355X)
  int atb = 0;         // FP32 blocked adjoint: configuration kFwiAdjTb[atb - 1]; 0 = per step
  DevBuf sets, table;  // blocked forward: two state sets [2][3][cmax][n], sponge table
  DevBuf asets;        // blocked adjoint: two state sets [2][3][cmax][n]
  // launch groups of the blocked passes: a group's shots split into `views`
  // contiguous sets, each set's launches on its own stream (set 0 on the
  // call's stream), forked from and joined back into the call's stream, so
  // that one set's launch start and tail overlap the other's work.  The
  // 16-shot 300x800 gradient took 0.1605 s in one set, 0.1455 s in 2, 0.186 s
  // in 4 and 0.196 s in 8 (the passes launch without graphs, and 4 or 8 sets
  // of short launches outrun the host's launch rate); results bit-identical.
  // The environment variable TSW_FWI_GROUPS (1..kFwiViews) sets the count.
  static constexpr int kFwiViews = 8;
  int views = 2;
  hipStream_t vs[kFwiViews] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kFwiViews] = {};
  double last_loss = 0;
  bool have_forward = false;
  ~Fwi() {
    for (int v = 1; v < kFwiViews; ++v) {
      if (vs[v]) (void)hipStreamDestroy(vs[v]);
      if (ev_join[v]) (void)hipEventDestroy(ev_join[v]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
  }
  size_t esz() const { return prec == 8 ? 8 : 4; }
  long long cells() const { return (long long)nx * nz; }
};

static unsigned g1(long long n) {
  long long b = (n + 255) / 256;
  return (unsigned)std::max<long long>(1, std::min<long long>(b, 65536));
}

template <typename T>
static FwdArgs<T> fwd_base(const Fwi* F, const FwiGroup& G) {
  FwdArgs<T> s;
  std::memset(&s, 0, sizeof s);
  const T dxT = (T)F->dx, dzT = (T)F->dz;
  s.nx = F->nx;
  s.nz = F->nz;
  s.fs = F->fs;
  s.rdx2 = T(1) / (dxT * dxT);
  s.rdz2 = T(1) / (dzT * dzT);
  s.inj = (T)(F->dt * F->dt);
  s.n = F->cells();
  s.C = (const T*)F->C.p;
  s.S = (const T*)F->S.p;
  s.meta = (const FwiMeta*)G.meta.p;
  return s;
}

// level b (0..2) of the group's working buffers
template <typename T>
static T* lvl(const Fwi* F, int b) {
  return (T*)F->buf.p + (long long)b * F->cmax * F->cells();
}
constexpr int kAdjRing = 5;  // adjoint levels: Lam^{n-1} .. Lam^{n+3}

template <typename T>
static T* adj(const Fwi* F, int b) {
  return (T*)F->mu.p + (long long)(b % kAdjRing) * F->cmax * F->cells();
}

template <typename T>
static void fwd_reset(const Fwi* F, const FwiGroup& G, hipStream_t st) {
  const long long n = F->cells(), cn = (long long)G.count() * n;
  for (int b = 0; b < 3; ++b)
    hipLaunchKernelGGL(k_fwi_fill<T>, dim3(g1(cn)), dim3(256), 0, st, lvl<T>(F, b), cn, T(0));
  hipLaunchKernelGGL(k_fwi_inject0<T>, dim3((G.count() + 63) / 64), dim3(64), 0, st,
                     lvl<T>(F, 1), n, (const FwiMeta*)G.meta.p, G.count(), (T)(F->dt * F->dt));
}

template <typename T>
static void fwd_step(const Fwi* F, const FwiGroup& G, FwdArgs<T>& s, int it, T* save, bool rec,
                     hipStream_t st) {
  s.prv = lvl<T>(F, it % 3);
  s.cur = lvl<T>(F, (it + 1) % 3);
  s.nxt = lvl<T>(F, (it + 2) % 3);
  s.save = save;
  s.src_it = it + 1 < G.nt ? it + 1 : -1;
  s.rec_it = rec ? it : -1;
  hipLaunchKernelGGL(k_fwi_fwd<T>, dim3((F->nx + 63) / 64, (F->nz + kTZ - 1) / kTZ, G.count()),
                     dim3(64, 4), 0, st, s);
}

// level l (0: u^{n-2}, 1: u^{n-1}, 2: u^n) of blocked-forward state set k
static float* tb_set(const Fwi* F, int k, int l) {
  return (float*)F->sets.p + (long long)(k * 3 + l) * F->cmax * F->cells();
}

static FwiTbArgs tb_base(const Fwi* F, const FwiGroup& G) {
  FwiTbArgs a;
  std::memset(&a, 0, sizeof a);
  const float dxf = (float)F->dx, dzf = (float)F->dz;
  a.nx = F->nx;
  a.nz = F->nz;
  a.fs = F->fs;
  a.n_sponge = F->n_sponge;
  a.rdx2 = 1.0f / (dxf * dxf);
  a.rdz2 = 1.0f / (dzf * dzf);
  a.inj = (float)(F->dt * F->dt);
  a.C = (const float*)F->C.p;
  a.table = F->table.f();
  a.n = F->cells();
  a.meta = (const FwiMeta*)G.meta.p;
  a.nt = G.nt;
  a.count = G.count();
  const FwiTbCfg& c = kFwiTb[F->tb - 1];
  const int W = 64 * c.WX - 4 * c.T, H = c.WZ * c.R - 4 * c.T;
  a.tiles_x = (F->nx + W - 1) / W;
  a.tiles = a.tiles_x * ((F->nz + H - 1) / H);
  return a;
}

// Blocked steps [i0, i1) of a group from state set *k (updated), recording
// the receivers or saving L u^it of every step into `save` (segment starting at
// save_it0); blocks never cross a multiple of `cut` (checkpoint boundaries,
// 0: none) so that the state is in memory there.
// A contiguous set of a group's shots: positions [p0, p0 + cnt).  Every
// per-shot array of the blocked passes is indexed by position (shot j at
// + j * n), so a view is the same launch with its pointers moved by p0.
struct FwiView {
  int p0, cnt;
};
static std::vector<FwiView> views_of(const Fwi* F, const FwiGroup& G) {
  const int nv = std::max(1, std::min(F->views, G.count()));
  std::vector<FwiView> v(nv);
  for (int i = 0; i < nv; ++i) {
    v[i].p0 = (int)((long long)i * G.count() / nv);
    v[i].cnt = (int)((long long)(i + 1) * G.count() / nv) - v[i].p0;
  }
  return v;
}
static hipStream_t view_stream(const Fwi* F, int v, hipStream_t st) { return v ? F->vs[v] : st; }
static int views_fork(Fwi* F, int nv, hipStream_t st) {
  if (nv <= 1) return kOk;
  TSW_HIP(hipEventRecord(F->ev_fork, st));
  for (int v = 1; v < nv; ++v) TSW_HIP(hipStreamWaitEvent(F->vs[v], F->ev_fork, 0));
  return kOk;
}
static int views_join(Fwi* F, int nv, hipStream_t st) {
  for (int v = 1; v < nv; ++v) {
    TSW_HIP(hipEventRecord(F->ev_join[v], F->vs[v]));
    TSW_HIP(hipStreamWaitEvent(st, F->ev_join[v], 0));
  }
  return kOk;
}

// Blocked steps [i0, i1) of the views V of a group, each on its stream, the
// views' launches interleaved; state set *k (updated) as for one launch list.
static void tb_run(const Fwi* F, const FwiGroup& G, const std::vector<FwiView>& V, int i0, int i1,
                   int* k, bool rec, float* save, int save_it0, hipStream_t st) {
  FwiTbArgs a = tb_base(F, G);
  const FwiTbCfg& c = kFwiTb[F->tb - 1];
  const long long n = F->cells();
  a.rec = rec ? 1 : 0;
  a.save_it0 = save_it0;
  a.sstride = (long long)G.count() * n;
  for (int it = i0; it < i1;) {
    const int len = std::min(c.T, i1 - it);
    a.it0 = it;
    a.nsub = len;
    for (size_t v = 0; v < V.size(); ++v) {
      const long long off = (long long)V[v].p0 * n;
      for (int l = 0; l < 3; ++l) {
        a.rd[l] = tb_set(F, *k, l) + off;
        a.wr[l] = tb_set(F, *k ^ 1, l) + off;
      }
      a.save = save ? save + off : nullptr;
      a.meta = (const FwiMeta*)G.meta.p + V[v].p0;
      a.count = V[v].cnt;
      c.launch(a, V[v].cnt, view_stream(F, (int)v, st));
    }
    *k ^= 1;
    it += len;
  }
}

// state set k := (0, 0, src[0] * inj at the source cell) for every shot
static void tb_reset(const Fwi* F, const FwiGroup& G, int k, hipStream_t st) {
  const long long n = F->cells(), cn = (long long)G.count() * n;
  for (int l = 0; l < 3; ++l)
    hipLaunchKernelGGL(k_fwi_fill<float>, dim3(g1(cn)), dim3(256), 0, st, tb_set(F, k, l), cn,
                       0.0f);
  hipLaunchKernelGGL(k_fwi_inject0<float>, dim3((G.count() + 63) / 64), dim3(64), 0, st,
                     tb_set(F, k, 2), n, (const FwiMeta*)G.meta.p, G.count(),
                     (float)(F->dt * F->dt));
}

template <typename T>
static int fwi_forward_t(Fwi* F, const T* d_vel, hipStream_t st, bool store_all) {
  const long long n = F->cells();
  const T dtT = (T)F->dt;
  hipLaunchKernelGGL(k_fwi_coef<T>, dim3(g1(n)), dim3(256), 0, st, d_vel, (T*)F->C.p, n,
                     T(dtT * dtT));
  const int K = F->kint;
  for (auto& G : F->groups) {
    const long long cn = (long long)G.count() * n;
    if (std::is_same<T, float>::value && F->tb > 0) {
      int k = 0;
      tb_reset(F, G, k, st);
      const std::vector<FwiView> V = views_of(F, G);
      TSW_TRY(views_fork(F, (int)V.size(), st));
      for (int it = 0; it < G.nt;) {
        int end = G.nt;
        if (!store_all) {
          T* ck = (T*)G.ckpt.p + (long long)(it / K) * 3 * cn;  // it % K == 0 here
          for (size_t v = 0; v < V.size(); ++v) {
            const long long off = (long long)V[v].p0 * n;
            for (int l = 0; l < 3; ++l)
              TSW_HIP(hipMemcpyAsync(ck + l * cn + off, tb_set(F, k, l) + off,
                                     V[v].cnt * n * sizeof(T), hipMemcpyDeviceToDevice,
                                     view_stream(F, (int)v, st)));
          }
          end = std::min(G.nt, it + K);
        }
        tb_run(F, G, V, it, end, &k, true, nullptr, 0, st);
        it = end;
      }
      TSW_TRY(views_join(F, (int)V.size(), st));
    } else {
    fwd_reset<T>(F, G, st);
    FwdArgs<T> s = fwd_base<T>(F, G);
    for (int it = 0; it < G.nt; ++it) {
      if (!store_all && it % K == 0) {  // checkpoint (u^{n-2}, u^{n-1}, u^n)
        T* ck = (T*)G.ckpt.p + (long long)(it / K) * 3 * cn;
        TSW_HIP(hipMemcpyAsync(ck, lvl<T>(F, (it + 2) % 3), cn * sizeof(T),
                               hipMemcpyDeviceToDevice, st));
        TSW_HIP(hipMemcpyAsync(ck + cn, lvl<T>(F, it % 3), cn * sizeof(T),
                               hipMemcpyDeviceToDevice, st));
        TSW_HIP(hipMemcpyAsync(ck + 2 * cn, lvl<T>(F, (it + 1) % 3), cn * sizeof(T),
                               hipMemcpyDeviceToDevice, st));
      }
      fwd_step<T>(F, G, s, it, nullptr, true, st);
    }
    }
    for (int j = 0; j < G.count(); ++j) {
      const int id = G.ids[j];
      const FwiShotHost& sh = F->shots[id];
      const long long nr = (long long)sh.nt * sh.nrec;
      hipLaunchKernelGGL(k_fwi_resid<T>, dim3(kSumBlocks), dim3(256), 0, st, (const T*)sh.pred.p,
                         (const T*)sh.obs.p, (T*)sh.res.p, nr, (double*)F->partial.p);
      hipLaunchKernelGGL(k_fwi_loss, dim3(1), dim3(256), 0, st, (const double*)F->partial.p,
                         (double*)F->loss.p + id);
    }
    TSW_HIP(hipGetLastError());
  }
  std::vector<double> l(F->shots.size());
  TSW_HIP(hipMemcpyAsync(l.data(), F->loss.p, l.size() * sizeof(double), hipMemcpyDeviceToHost,
                         st));
  TSW_HIP(hipStreamSynchronize(st));
  double total = 0;
  for (double x : l) total += x;
  F->last_loss = total;
  F->have_forward = true;
  return kOk;
}

static float* atb_set(const Fwi* F, int k, int l) {
  return (float*)F->asets.p + (long long)(k * 3 + l) * F->cmax * F->cells();
}

// Blocked adjoint steps n1-1 down to n0 of a group from state set *k
// (updated); `lu` holds L u^it of the segment starting at n0.
static void adj_tb_run(const Fwi* F, const FwiGroup& G, const std::vector<FwiView>& V, int n0,
                       int n1, int* k, const float* lu, hipStream_t st) {
  const FwiAdjTbCfg& c = kFwiAdjTb[F->atb - 1];
  FwiAdjTbArgs a;
  std::memset(&a, 0, sizeof a);
  const float dxf = (float)F->dx, dzf = (float)F->dz;
  a.nx = F->nx;
  a.nz = F->nz;
  a.fs = F->fs;
  a.n_sponge = F->n_sponge;
  a.rdx2 = 1.0f / (dxf * dxf);
  a.rdz2 = 1.0f / (dzf * dzf);
  a.C = (const float*)F->C.p;
  a.table = F->table.f();
  a.n = F->cells();
  a.lu = lu;
  a.sstride = (long long)G.count() * F->cells();
  a.lu_it0 = n0;
  a.grad = (float*)F->gC.p;
  a.meta = (const FwiMeta*)G.meta.p;
  a.count = G.count();
  const int W = 64 * c.WX - 4 * c.T, H = c.WZ * c.R - 4 * c.T;
  a.tiles_x = (F->nx + W - 1) / W;
  a.tiles = a.tiles_x * ((F->nz + H - 1) / H);
  const long long n = F->cells();
  for (int it = n1 - 1; it >= n0;) {
    const int len = std::min(c.T, it - n0 + 1);
    a.it0 = it;
    a.nsub = len;
    for (size_t v = 0; v < V.size(); ++v) {
      const long long off = (long long)V[v].p0 * n;
      for (int l = 0; l < 3; ++l) {
        a.rd[l] = atb_set(F, *k, l) + off;
        a.wr[l] = atb_set(F, *k ^ 1, l) + off;
      }
      a.lu = lu + off;
      a.grad = (float*)F->gC.p + off;
      a.meta = (const FwiMeta*)G.meta.p + V[v].p0;
      a.count = V[v].cnt;
      c.launch(a, view_stream(F, (int)v, st));
    }
    *k ^= 1;
    it -= len;
  }
}

template <typename T>
static int fwi_backward_t(Fwi* F, const T* d_vel, T* d_grad, hipStream_t st, bool store_all) {
  const long long n = F->cells();
  const T dtT = (T)F->dt;
  const T dt2 = dtT * dtT;
  const long long gn = (long long)F->cmax * n;
  hipLaunchKernelGGL(k_fwi_fill<T>, dim3(g1(gn)), dim3(256), 0, st, (T*)F->gC.p, gn, T(0));
  for (auto& G : F->groups) {
    const int nt = G.nt, cnt = G.count();
    const long long cn = (long long)cnt * n;
    // Lam^{nt} = R^T r_{nt-1}; Lam^{nt+1..nt+3} = 0: in the ring of the per-step
    // kernels, or in state set 0 (Lam^{n+1}, Lam^{n+2}, Lam^{n+3}) of the blocked one
    const bool blocked_adj = std::is_same<T, float>::value && F->atb > 0;
    int ak = 0;
    if (blocked_adj) {
      for (int l = 0; l < 3; ++l)
        hipLaunchKernelGGL(k_fwi_fill<float>, dim3(g1(cn)), dim3(256), 0, st, atb_set(F, 0, l),
                           cn, 0.0f);
      hipLaunchKernelGGL(k_fwi_seed<float>, dim3((F->nx + 255) / 256, cnt), dim3(256), 0, st,
                         atb_set(F, 0, 0), n, F->nx, (const FwiMeta*)G.meta.p, nt - 1);
    } else {
      for (int b = 0; b < kAdjRing; ++b)
        hipLaunchKernelGGL(k_fwi_fill<T>, dim3(g1(cn)), dim3(256), 0, st, adj<T>(F, b), cn,
                           T(0));
      hipLaunchKernelGGL(k_fwi_seed<T>, dim3((F->nx + 255) / 256, cnt), dim3(256), 0, st,
                         adj<T>(F, nt), n, F->nx, (const FwiMeta*)G.meta.p, nt - 1);
    }
    AdjArgs<T> a;
    std::memset(&a, 0, sizeof a);
    const T dxT = (T)F->dx, dzT = (T)F->dz;
    a.nx = F->nx;
    a.nz = F->nz;
    a.fs = F->fs;
    a.rdx2 = T(1) / (dxT * dxT);
    a.rdz2 = T(1) / (dzT * dzT);
    a.n = n;
    a.C = (const T*)F->C.p;
    a.S = (const T*)F->S.p;
    if (std::is_same<T, float>::value && F->table.p) {
      a.table = F->table.f();
      a.n_sponge = F->n_sponge;
    }
    a.grad = (T*)F->gC.p;
    a.meta = (const FwiMeta*)G.meta.p;
    FwdArgs<T> s = fwd_base<T>(F, G);
    const int K = store_all ? nt : F->kint;
    const int nseg = (nt + K - 1) / K;
    // the blocked recompute and the blocked adjoint run the group's shots as
    // views on their own streams from here to the join below
    const bool views = std::is_same<T, float>::value && F->tb > 0 && blocked_adj && !store_all;
    const std::vector<FwiView> V = views ? views_of(F, G) : std::vector<FwiView>{{0, cnt}};
    if (views) TSW_TRY(views_fork(F, (int)V.size(), st));
    for (int sg = nseg - 1; sg >= 0; --sg) {
      const int n0 = sg * K, n1 = std::min(nt, n0 + K);
      if (std::is_same<T, float>::value && F->tb > 0) {
        int k = 0;
        if (store_all) {
          tb_reset(F, G, k, st);
        } else {
          const T* ck = (const T*)G.ckpt.p + (long long)sg * 3 * cn;
          for (size_t v = 0; v < V.size(); ++v) {
            const long long off = (long long)V[v].p0 * n;
            for (int l = 0; l < 3; ++l)
              TSW_HIP(hipMemcpyAsync(tb_set(F, k, l) + off, ck + l * cn + off,
                                     V[v].cnt * n * sizeof(T), hipMemcpyDeviceToDevice,
                                     view_stream(F, (int)v, st)));
          }
        }
        tb_run(F, G, V, n0, n1, &k, false, (float*)F->seg.p, n0, st);
      } else {
      if (store_all) {
        fwd_reset<T>(F, G, st);  // one segment: the whole forward again
      } else {
        // restore (u^{n0-2}, u^{n0-1}, u^{n0}) into the buffers the step at n0 uses
        const T* ck = (const T*)G.ckpt.p + (long long)sg * 3 * cn;
        TSW_HIP(hipMemcpyAsync(lvl<T>(F, (n0 + 2) % 3), ck, cn * sizeof(T),
                               hipMemcpyDeviceToDevice, st));
        TSW_HIP(hipMemcpyAsync(lvl<T>(F, n0 % 3), ck + cn, cn * sizeof(T),
                               hipMemcpyDeviceToDevice, st));
        TSW_HIP(hipMemcpyAsync(lvl<T>(F, (n0 + 1) % 3), ck + 2 * cn, cn * sizeof(T),
                               hipMemcpyDeviceToDevice, st));
      }
      for (int it = n0; it < n1; ++it)
        fwd_step<T>(F, G, s, it, (T*)F->seg.p + (long long)(it - n0) * cn, false, st);
      }
      if (blocked_adj) {
        adj_tb_run(F, G, V, n0, n1, &ak, (const float*)F->seg.p, st);
        continue;
      }
      for (int it = n1 - 1; it >= n0;) {
        a.lam0 = adj<T>(F, it);
        a.lam1 = adj<T>(F, it + 1);
        a.lam2 = adj<T>(F, it + 2);
        a.lam3 = adj<T>(F, it + 3);
        a.lun = (const T*)F->seg.p + (long long)(it - n0) * cn;
        a.res_it = it >= 1 ? it - 1 : -1;
        a.probe_it = it;
        if (F->adj_steps == 2 && it - 1 >= n0) {
          a.lamm1 = adj<T>(F, it + 4);  // Lam^{it-1}: the ring slot of Lam^{it+4}
          a.lum1 = (const T*)F->seg.p + (long long)(it - 1 - n0) * cn;
          a.res_it2 = it >= 2 ? it - 2 : -1;
          hipLaunchKernelGGL((k_fwi_adj2<T, 16>), dim3((F->nx + 63) / 64, (F->nz + 15) / 16, cnt),
                             dim3(64, 4), 0, st, a);
          it -= 2;
        } else {
          hipLaunchKernelGGL((k_fwi_adj<T, 16>), dim3((F->nx + 63) / 64, (F->nz + 15) / 16, cnt),
                             dim3(64, 4), 0, st, a);
          it -= 1;
        }
      }
    }
    if (views) TSW_TRY(views_join(F, (int)V.size(), st));
    TSW_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fwi_chain<T>, dim3(g1(n)), dim3(256), 0, st, (const T*)F->gC.p, F->cmax,
                     d_vel, d_grad, n, dt2);
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
  for (int i = 0; i < nshot; ++i) {
    const tsw_fwi_shot& d = shots[i];
    TSW_ARG(d.nt >= 1 && d.src && d.d_obs && d.rix && d.nrec >= 1,
            "shot %d: need nt >= 1, nrec >= 1 and non-null arrays", i);
    TSW_ARG(d.six >= 0 && d.six < nx && d.siz >= 0 && d.siz < nz,
            "shot %d: source (%d, %d) outside the %dx%d grid", i, d.siz, d.six, nz, nx);
    TSW_ARG(d.riz >= 0 && d.riz < nz, "shot %d: receiver row %d outside the grid", i, d.riz);
    for (int r = 0; r < d.nrec; ++r)
      TSW_ARG(d.rix[r] >= 0 && d.rix[r] < nx, "shot %d: receiver %d column %d outside [0, %d)",
              i, r, d.rix[r], nx);
  }
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
  if (const char* ev = std::getenv("TSW_FWI_GROUPS")) {
    const int v = std::atoi(ev);
    if (v >= 1 && v <= Fwi::kFwiViews) F->views = v;
  }
  if (hipEventCreateWithFlags(&F->ev_fork, hipEventDisableTiming) != hipSuccess)
    return fail((set_error("event creation failed"), kErrHip));
  for (int v = 1; v < Fwi::kFwiViews; ++v)
    if (hipStreamCreateWithFlags(&F->vs[v], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&F->ev_join[v], hipEventDisableTiming) != hipSuccess)
      return fail((set_error("stream creation failed"), kErrHip));
  F->alpha = sponge_alpha;
  const long long n = F->cells();
  const size_t es = F->esz();
  int ntmax = 1;
  for (int i = 0; i < nshot; ++i) ntmax = std::max(ntmax, shots[i].nt);
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
    if (hipMemcpy(h.src.p, d.src, d.nt * es, hipMemcpyHostToDevice) ||
        hipMemcpy(h.obs.p, d.d_obs, nr, hipMemcpyHostToDevice) ||
        hipMemcpy(h.roff.p, m.off.data(), m.off.size() * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(h.ridx.p, m.idx.data(), m.idx.size() * sizeof(int), hipMemcpyHostToDevice))
      return fail((set_error("shot %d upload failed", i), kErrHip));
    // group by nt, in order of first appearance
    FwiGroup* G = nullptr;
    for (auto& g : F->groups)
      if (g.nt == d.nt) G = &g;
    if (!G) {
      F->groups.emplace_back();
      G = &F->groups.back();
      G->nt = d.nt;
    }
    G->ids.push_back(i);
  }
  F->cmax = 1;
  for (auto& G : F->groups) {
    F->cmax = std::max(F->cmax, G.count());
    std::vector<FwiMeta> meta(G.count());
    for (int j = 0; j < G.count(); ++j) {
      const FwiShotHost& h = F->shots[G.ids[j]];
      FwiMeta& m = meta[j];
      std::memset(&m, 0, sizeof m);
      m.src = h.src.p;
      m.res = h.res.p;
      m.pred = h.pred.p;
      m.probe = h.probe.p;
      m.roff = h.roff.i();
      m.ridx = h.ridx.i();
      m.src_cell = h.siz * nx + h.six;
      m.riz = h.riz;
      m.nrec = h.nrec;
    }
    if ((rc = G.meta.ensure(meta.size() * sizeof(FwiMeta)))) return fail(rc);
    if (hipMemcpy(G.meta.p, meta.data(), meta.size() * sizeof(FwiMeta), hipMemcpyHostToDevice))
      return fail((set_error("metadata upload failed"), kErrHip));
    if (!store_all) {
      const int nck = (G.nt + F->kint - 1) / F->kint;
      if ((rc = G.ckpt.ensure((size_t)nck * 3 * G.count() * n * es))) return fail(rc);
    }
  }
  const size_t cn = (size_t)F->cmax * n * es;
  if ((rc = F->C.ensure(n * es)) || (rc = F->gC.ensure(cn)) || (rc = F->buf.ensure(3 * cn)) ||
      (rc = F->mu.ensure(kAdjRing * cn)) ||
      (rc = F->seg.ensure((size_t)(store_all ? ntmax : std::min(F->kint, ntmax)) * cn)) ||
      (rc = F->partial.ensure(kSumBlocks * sizeof(double))) ||
      (rc = F->loss.ensure(nshot * sizeof(double))))
    return fail(rc);
  F->kint = store_all ? -1 : F->kint;
  if (precision == 4) {  // blocked forward by default in FP32
    std::vector<float> tab;
    sponge_table(tab, n_sponge, (float)sponge_alpha);
    if ((rc = F->sets.ensure(6 * cn)) || (rc = F->asets.ensure(6 * cn)) ||
        (rc = F->table.ensure(tab.size() * sizeof(float))))
      return fail(rc);
    if (hipMemcpy(F->table.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice))
      return fail((set_error("sponge table upload failed"), kErrHip));
    F->tb = kFwiDefaultTb;
    F->atb = kFwiDefaultAdjTb;
  }
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
  if (rc == kOk && loss_out) *loss_out = F->last_loss;
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

int tsw_fwi_set_blocking(tsw_fwi_t h, int config) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F, "null argument");
  TSW_ARG(config >= 0 && config <= kFwiTbN, "blocking configuration %d outside [0, %d]", config,
          kFwiTbN);
  TSW_ARG(config == 0 || F->prec == 4, "the blocked forward is FP32 only");
  F->tb = config;
  F->have_forward = false;  // checkpoints of another path: run the forward again
  return kOk;
}

int tsw_fwi_set_adjoint_blocking(tsw_fwi_t h, int config) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F, "null argument");
  TSW_ARG(config >= 0 && config <= kFwiTbN, "adjoint blocking configuration %d outside [0, %d]",
          config, kFwiTbN);
  TSW_ARG(config == 0 || F->prec == 4, "the blocked adjoint is FP32 only");
  F->atb = config;
  return kOk;
}

int tsw_fwi_set_adjoint_steps(tsw_fwi_t h, int steps) {
  Fwi* F = reinterpret_cast<Fwi*>(h);
  TSW_ARG(F && (steps == 1 || steps == 2), "adjoint steps per launch must be 1 or 2");
  F->adj_steps = steps;
  return kOk;
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
  if (!h) return kOk;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  (void)hipDeviceSynchronize();
  delete reinterpret_cast<Fwi*>(h);
  return kOk;
}

}  // extern "C"
