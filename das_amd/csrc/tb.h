// Temporally blocked acoustic RTM step kernel: one launch advances T time
// steps of every active shot.
//
// A workgroup owns a core tile of W x H cells and loads it with a 2T-cell
// halo (clipped at the grid).  Thread (lx, ly) owns column X0+lx and the R
// consecutive rows Z0+ly*R .. Z0+ly*R+R-1 of the extended tile; it keeps the
// three most recent time levels of its cells, the model coefficient, the
// sponge factor and the image / illumination accumulator in registers.
// Neighbours along x, and along z across thread boundaries, are exchanged
// through double-buffered LDS with one barrier per sub-step.  After s sub-steps
// the cells whose distance to a cut (non-grid) tile edge is >= 2s are exact,
// so after T sub-steps the core is exact and is written back.
//
// Per sub-step the kernel reproduces the reference loop exactly
// (tsw_rtm/wave_solver.py:157-183 forward, :210-242 backward,
// tsw_rtm/c_engine/wave2d.c:19-82): snapshot + illumination or zero-lag
// imaging of the post-injection level, the 4th-order update with the
// reference's FP32 association, free-surface rows (row 0 = 0, row 1 = -row 2,
// all three rows owned by one thread), cells outside [2,nz-2) x [2,nx-2) that
// keep the level from three buffers back, and injection of the next
// iteration's samples.  Results are bit-identical to the per-step kernel.
#pragma once
#include "batch.h"

namespace tsw {

// One block of sub-steps of one shot in one launch.
struct TbEntry {
  int shot;
  int phase;    // 0 forward (iterations it0, it0+1, ...), 1 back-propagation (it0, it0-1, ...)
  int it0;
  int nsub;     // sub-steps in this block (<= T)
  int flags;    // kTbFirst: state generated (zero + iteration-0 injection), not loaded
                // kTbLast: last block of the phase (no state write-back; illum scaled)
  int rd;       // state set read (0/1)
  int tiles_x;
  int pad;
};
constexpr int kTbFirst = 1;
constexpr int kTbLast = 2;

// State of a shot for the blocked kernel: two sets of three levels
// (level t-2, t-1, t) used alternately as input and output.
struct TbShot {
  int nx, nz, nt, k;
  int src_cell, riz, nrec;
  int pad;
  float* set[2][3];
  const float* coef_f;
  const float* coef_b;
  const float* src;
  const float* data;
  const int* roff;
  const int* ridx;
  float* snaps;
  float* illum;
  float* image;
};

__device__ __forceinline__ int xcd_tile_tb(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, l = b >> 3;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

// Configuration `id` (1-based) of the instantiated blocked kernels: steps per
// launch and core tile; and a launch of it (defined in batch.hip).
int tb_config(int id, int* T, int* W, int* H);
// Widest sponge strip (cells) configuration `id` supports.
int tb_max_sponge(int id);
void tb_launch(int id, const BatchConst& bc, const TbShot* shots, const TbEntry* ents,
               const int* pre, int nent, int ntiles, hipStream_t st);

template <int T, int R, int WX, int WZ>
struct TbShape {
  static constexpr int CX = 64 * WX;     // extended columns
  static constexpr int CZ = WZ * R;      // extended rows
  static constexpr int W = CX - 4 * T;   // core columns
  static constexpr int H = CZ - 4 * T;   // core rows
  static constexpr int LS = CX + 4;      // LDS row stride (2 pad columns each side)
  static constexpr int LR = CZ + 4;      // LDS rows (2 pad rows each side)
  static constexpr int THREADS = 64 * WX * WZ;
  static_assert(W > 0 && H > 0, "halo larger than the tile");
  static_assert(R >= 3, "rows 0..2 of a free surface must belong to one thread");
};

// Waves per SIMD the register allocation must leave room for: two 1024-thread
// workgroups per CU (8 waves per SIMD) for the 4-row shapes, which then fit
// in 64 VGPRs; no constraint for the others.
template <int R, int WX, int WZ>
constexpr int kTbWavesPerEu = R <= 4 && WX * WZ >= 16 ? 8 : 1;

template <int T, int R, int WX, int WZ>
__global__ __launch_bounds__(64 * WX * WZ)
__attribute__((amdgpu_waves_per_eu(kTbWavesPerEu<R, WX, WZ>))) void k_tb(
    BatchConst bc, const TbShot* shots, const TbEntry* ents, const int* pre, int nent,
    int ntiles) {
  using S = TbShape<T, R, WX, WZ>;
  __shared__ float lds[2][S::LR * S::LS];

  const int b = xcd_tile_tb(blockIdx.x, ntiles);
  int lo = 0, hi = nent - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const TbEntry E = ents[lo];
  const TbShot& P = shots[E.shot];
  const int t = b - pre[lo];
  const int nx = P.nx, nz = P.nz;
  const int cx0 = (t % E.tiles_x) * S::W;
  const int cz0 = (t / E.tiles_x) * S::H;
  const int X0 = cx0 - 2 * T > 0 ? cx0 - 2 * T : 0;
  const int Z0 = cz0 - 2 * T > 0 ? cz0 - 2 * T : 0;
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int col = X0 + lx;
  const int row0 = Z0 + ly * R;
  const bool fwd = E.phase == 0;
  const bool first = (E.flags & kTbFirst) != 0;
  const bool last = (E.flags & kTbLast) != 0;
  const long long ncell = (long long)nx * nz;

  const bool col_in = col < nx;
  const bool core_col = col_in && col >= cx0 && col < cx0 + S::W;
  const bool col_int = col >= 2 && col < nx - 2;
  // per-row classification bit masks (the row part is uniform over a wave)
  unsigned m_in = 0, m_int = 0, m_core = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int row = row0 + q;
    const bool rin = row < nz;
    m_in |= (unsigned)(col_in && rin) << q;
    m_int |= (unsigned)(col_int && row >= 2 && row < nz - 2) << q;
    m_core |= (unsigned)(core_col && rin && row >= cz0 && row < cz0 + S::H) << q;
  }
  // rows 0/1 under a free surface are rewritten every step, never "kept"
  const unsigned m_fs = (bc.fs && row0 == 0) ? (m_in & 3u) : 0u;
  const unsigned m_keep = m_in & ~m_int & ~m_fs;

  const gfloat* coef = as_global(fwd ? P.coef_f : P.coef_b);
  const gfloat* tab = as_global(bc.table);
  float a[R], bl[R], c[R], cf[R], sp[R], acc[R];

  // ---- load (or generate) the three input levels.  Lanes past the grid read
  // the clamped last column / row and discard it, so that the loads of all
  // rows are issued together without branches; 32-bit element offsets (the
  // batch admits grids of fewer than 2^31 cells).
  const int colc = col_in ? col : nx - 1;
  const int rd = E.rd;
  const gfloat* lv2 = as_global(P.set[rd][2]);
  const gfloat* lv1 = as_global(P.set[rd][1]);
  const gfloat* lv0 = as_global(P.set[rd][0]);
  const gfloat* accp = as_global(fwd ? P.illum : P.image);
  unsigned idx[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int row = row0 + q;
    const int rowc = row < nz ? row : nz - 1;
    idx[q] = (unsigned)(rowc * nx + colc);
    const bool in = m_in >> q & 1u;
    const float cfv = coef[idx[q]];
    // depth of a clamped cell is at most n_sponge; table[0] = 1
    const float spv = tab[sponge_depth(colc, rowc, nx, nz, bc.n_sponge, bc.fs)];
    cf[q] = in ? cfv : 0.0f;
    sp[q] = in ? spv : 1.0f;
    a[q] = 0.0f;
    bl[q] = 0.0f;
    c[q] = 0.0f;
    acc[q] = 0.0f;
  }
  if (first) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (m_in >> q & 1u) {
        const int row = row0 + q;
        if (fwd) {
          if ((int)idx[q] == P.src_cell) c[q] = 0.0f + as_global(P.src)[0] * bc.inj;
        } else if (row == P.riz) {
          const int r0 = as_global(P.roff)[col], r1 = as_global(P.roff)[col + 1];
          const gfloat* drow = as_global(P.data) + (long long)(P.nt - 1) * P.nrec;
          float v = 0.0f;
          for (int r = r0; r < r1; ++r) v = v + drow[as_global(P.ridx)[r]] * bc.inj;
          c[q] = v;
        }
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool in = m_in >> q & 1u;
      const float cv = lv2[idx[q]], bv = lv1[idx[q]];
      c[q] = in ? cv : 0.0f;
      bl[q] = in ? bv : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (m_keep >> q & 1u) a[q] = lv0[idx[q]];
      if (m_core >> q & 1u) acc[q] = accp[idx[q]];
    }
  }
  float* X = lds[0];
#pragma unroll
  for (int q = 0; q < R; ++q) X[(ly * R + q + 2) * S::LS + lx + 2] = c[q];
  __syncthreads();

  // Shot fields used every sub-step, held in registers: read through `P`
  // they would be re-read after each store the compiler cannot disambiguate.
  const int k = P.k, nt = P.nt, nrec = P.nrec, riz = P.riz;
  const gfloat* src = as_global(P.src);
  const gfloat* data = as_global(P.data);
  gfloat* snaps = as_global(P.snaps);
  const int src_row = P.src_cell / nx, src_col = P.src_cell - src_row * nx;
  const bool src_thread = fwd && src_col == col && src_row >= row0 && src_row < row0 + R;
  // receiver of this column on the back-propagation row: its index when the
  // column has exactly one, -1 with none, -2 with several
  int rsel = -1;
  if (!fwd && riz >= row0 && riz < row0 + R && col_in) {
    const int r0 = as_global(P.roff)[col], r1 = as_global(P.roff)[col + 1];
    rsel = r1 - r0 == 1 ? as_global(P.ridx)[r0] : (r1 > r0 ? -2 : -1);
  }
  // iteration `it` as phase it % k and snapshot level it / k, stepped by one
  // per sub-step instead of divided
  int ph = E.it0 % k, lvl = E.it0 / k;

  const float w = bc.w;
  // The sub-steps are unrolled (at most T of them), so the rotation of the
  // three levels is a renaming of registers and each LDS buffer's address is
  // fixed per sub-step.
#pragma unroll
  for (int j = 0; j < T; ++j) {
    if (j == E.nsub) break;
    const int it = fwd ? E.it0 + j : E.it0 - j;
    float* Y = lds[(j + 1) & 1];
    // side work on the post-injection level `c`
    if (ph == 0 && m_core) {
      gfloat* snap = snaps + (long long)lvl * ncell;
      if (fwd) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (m_core >> q & 1u) {
            // snapshots stream through HBM once each way: keep them out of the caches
            __builtin_nontemporal_store(c[q], snap + idx[q]);
            acc[q] = acc[q] + c[q] * c[q];
          }
        }
      } else {
        // every row's load is issued before the first use, so the rows share
        // one memory latency instead of queueing one after another
        float sv[R];
#pragma unroll
        for (int q = 0; q < R; ++q)
          sv[q] = (m_core >> q & 1u) ? __builtin_nontemporal_load(snap + idx[q]) : 0.0f;
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (m_core >> q & 1u) acc[q] = acc[q] + (sv[q] * c[q]) * w;
      }
    }
    // leapfrog update
    float n[R];
    const float* xr = X + 2;  // column offset of lx = 0
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int lr = (ly * R + q + 2) * S::LS + lx;
      // evaluated on every lane (all reads stay inside this LDS buffer), so the
      // reads of all rows issue together; non-interior lanes then take `a`
      const float pc = c[q];
      const float xm2 = xr[lr - 2], xm1 = xr[lr - 1], xp1 = xr[lr + 1], xp2 = xr[lr + 2];
      const float zm2 = q >= 2 ? c[q - 2] : xr[lr - 2 * S::LS];
      const float zm1 = q >= 1 ? c[q - 1] : xr[lr - S::LS];
      const float zp1 = q + 1 < R ? c[q + 1] : xr[lr + S::LS];
      const float zp2 = q + 2 < R ? c[q + 2] : xr[lr + 2 * S::LS];
      float lx2 = TSW_C2 * (xm2 + xp2);
      lx2 = lx2 + TSW_C1 * (xm1 + xp1);
      lx2 = (lx2 + TSW_C0 * pc) * bc.rdx2;
      float lz2 = TSW_C2 * (zm2 + zp2);
      lz2 = lz2 + TSW_C1 * (zm1 + zp1);
      lz2 = (lz2 + TSW_C0 * pc) * bc.rdz2;
      float v = (2.0f * pc - bl[q]) + cf[q] * (lx2 + lz2);
      v = v * sp[q];
      // kept cells (and cells outside the grid, which stay 0) keep `a`
      n[q] = (m_int >> q & 1u) ? v : a[q];
    }
    if (m_fs) {  // this thread holds rows 0, 1 and 2 of its column
      n[0] = 0.0f;
      n[1] = -n[2];
    }
    // injection of the next iteration's samples
    if (fwd) {
      if (src_thread && it + 1 < nt) {
        const float s = src[it + 1];
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (row0 + q == src_row) n[q] = n[q] + s * bc.inj;
      }
    } else if (rsel != -1 && it >= 1) {
      const gfloat* drow = data + (long long)(it - 1) * nrec;
      if (rsel >= 0) {
        const float d = drow[rsel];
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (row0 + q == riz) n[q] = n[q] + d * bc.inj;
      } else {
        const int r0 = as_global(P.roff)[col], r1 = as_global(P.roff)[col + 1];
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (row0 + q == riz)
            for (int r = r0; r < r1; ++r) n[q] = n[q] + drow[as_global(P.ridx)[r]] * bc.inj;
      }
    }
    // next iteration's phase and snapshot level
    if (fwd) {
      if (++ph == k) {
        ph = 0;
        ++lvl;
      }
    } else if (ph-- == 0) {
      ph = k - 1;
      --lvl;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      Y[(ly * R + q + 2) * S::LS + lx + 2] = n[q];
      a[q] = bl[q];
      bl[q] = c[q];
      c[q] = n[q];
    }
    X = Y;
    __syncthreads();
  }

  // ---- write back the core
  const int wr = rd ^ 1;
  gfloat* acco = as_global(fwd ? P.illum : P.image);
  gfloat* wv2 = as_global(P.set[wr][2]);
  gfloat* wv1 = as_global(P.set[wr][1]);
  gfloat* wv0 = as_global(P.set[wr][0]);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (m_core >> q & 1u) {
      if (!last) {
        wv2[idx[q]] = c[q];
        wv1[idx[q]] = bl[q];
        if (m_keep >> q & 1u) wv0[idx[q]] = a[q];
      }
      acco[idx[q]] = (fwd && last) ? acc[q] * w : acc[q];
    }
  }
}

}  // namespace tsw
