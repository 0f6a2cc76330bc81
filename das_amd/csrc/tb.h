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
  int snap;     // bit j: iteration it0 +- j is a snapshot iteration (it % k == 0)
  int lvl;      // snapshot level of the first such iteration
  int inj;      // bit j: sub-step j is followed by an injection (a source sample
                // it+1 < nt forward, a data row it-1 >= 0 back-propagating)
  int pad;
};
constexpr int kTbFirst = 1;
constexpr int kTbLast = 2;

// Fills the schedule fields of a block of `e.nsub` sub-steps from `e.it0` in
// phase `e.phase` of a shot with `nt` iterations and snapshot stride `k`.
inline void tb_schedule(TbEntry& e, int nt, int k) {
  e.snap = 0;
  e.inj = 0;
  e.lvl = -1;
  for (int j = 0; j < e.nsub; ++j) {
    const int it = e.phase == 0 ? e.it0 + j : e.it0 - j;
    if (it % k == 0) {
      e.snap |= 1 << j;
      if (e.lvl < 0) e.lvl = it / k;
    }
    if (e.phase == 0 ? it + 1 < nt : it >= 1) e.inj |= 1 << j;
  }
  if (e.lvl < 0) e.lvl = 0;
}

// One workgroup's tile: the index of its entry in the launch and the tile's
// column and row in the shot's grid of tiles (tx | tz << 16).  The host lists
// them per workgroup, the XCD-aware order of the tiles already applied.
struct TbTile {
  int ent;
  int txz;
};

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

__host__ __device__ inline int xcd_tile_tb(int b, int n) {
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
               const TbTile* tiles, int ntiles, hipStream_t st);

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

// Tile (tx, tz) of entry `E`.
template <int T, int R, int WX, int WZ>
__device__ __forceinline__ void tb_tile(const BatchConst& bc, const TbEntry& E, const TbShot& P,
                                        int tx, int tz, float* lds) {
  using S = TbShape<T, R, WX, WZ>;
  const bool fwd = E.phase == 0;
  const int nx = P.nx, nz = P.nz;
  const int cx0 = tx * S::W, cz0 = tz * S::H;
  const int X0 = cx0 - 2 * T > 0 ? cx0 - 2 * T : 0;
  const int Z0 = cz0 - 2 * T > 0 ? cz0 - 2 * T : 0;
  const int lx = threadIdx.x;
  // blockDim.x is a multiple of 64, so the lanes of a wave share threadIdx.y:
  // the thread's rows are uniform over the wave and live in scalar registers
  const int ly = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int col = X0 + lx;
  const int row0 = Z0 + ly * R;
  const bool first = (E.flags & kTbFirst) != 0;
  const bool last = (E.flags & kTbLast) != 0;
  const int nsub = E.nsub;

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
  const bool fs0 = bc.fs && row0 == 0;
  const unsigned m_keep = m_in & ~m_int & ~(fs0 ? 3u : 0u);
  // whether the wave holds any cell outside [2,nz-2) x [2,nx-2) (free-surface
  // rows and cells past the grid included); uniform
  const bool edge = !(X0 >= 2 && X0 + S::CX <= nx - 2 && row0 >= 2 && row0 + R <= nz - 2);
#define TB_IN(q) (m_in >> (q) & 1u)
#define TB_INT(q) (m_int >> (q) & 1u)
#define TB_CORE(q) (m_core >> (q) & 1u)
#define TB_KEEP(q) (m_keep >> (q) & 1u)

  float a[R], bl[R], c[R], cf[R], sp[R], acc[R];

  // ---- load (or generate) the three input levels.  Lanes past the grid read
  // the clamped last column / row: their values reach only cells outside
  // [2,nz-2) x [2,nx-2), whose update is discarded, and are never written
  // back, so the loads of all rows issue together without branches.  Every
  // access is at the lane's byte offset (vector register) plus the row's
  // byte offset (scalar register, uniform over the wave).
  const int colc = col_in ? col : nx - 1;
  const unsigned lane4 = (unsigned)colc * 4u;
  const unsigned rowb = (unsigned)nx * 4u;
  unsigned roff[R];  // byte offset of row min(row0 + q, nz - 1)
#pragma unroll
  for (int q = 0; q < R; ++q) roff[q] = (unsigned)(row0 + q < nz ? row0 + q : nz - 1) * rowb;
  const int rd = E.rd;
  // sponge depth (the larger of the column's and the row's, which is how the
  // reference's maximum over the four edges splits): a clamped cell's depth
  // is at most n_sponge, and table[0] = 1
  const int ns = bc.n_sponge;
  int dcol = ns - colc;
  dcol = dcol > colc - (nx - 1 - ns) ? dcol : colc - (nx - 1 - ns);
  dcol = dcol > 0 ? dcol : 0;
  {
    const float* const pcoef = fwd ? P.coef_f : P.coef_b;
    const float* const ptab = bc.table;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int row = row0 + q;
      const int rowc = row < nz ? row : nz - 1;
      int drow = rowc - (nz - 1 - ns);
      if (!bc.fs) drow = drow > ns - rowc ? drow : ns - rowc;
      const int d = dcol > drow ? dcol : drow;
      cf[q] = ld_at(pcoef, lane4, roff[q]);
      sp[q] = ld_at(ptab, (unsigned)d * 4u, 0);
      a[q] = 0.0f;
      bl[q] = 0.0f;
      c[q] = 0.0f;
      acc[q] = 0.0f;
    }
  }
  if (first) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (TB_IN(q)) {
        const int row = row0 + q;
        if (fwd) {
          if (row * nx + col == P.src_cell) c[q] = 0.0f + as_global(P.src)[0] * bc.inj;
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
    const float* const p2 = P.set[rd][2];
    const float* const p1 = P.set[rd][1];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      c[q] = ld_at(p2, lane4, roff[q]);
      bl[q] = ld_at(p1, lane4, roff[q]);
    }
    const float* const p0 = P.set[rd][0];
    const float* const pacc = fwd ? P.illum : P.image;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (TB_KEEP(q)) a[q] = ld_at(p0, lane4, roff[q]);
      if (TB_CORE(q)) acc[q] = ld_at(pacc, lane4, roff[q]);
    }
  }
  // ---- the wave's schedule, uniform and fixed before the sub-steps.
  // Local rows [core_lo - e, core_hi + e) of level j+1, e = 2(T-1-j), are the
  // ones the core at level T depends on; rows past the grid feed only rows
  // outside [2, nz-2).  These ranges shrink with j: from the first sub-step
  // `jend` whose range misses all of the wave's rows (then it has no core
  // rows) the wave only joins the remaining barriers, and what it left in LDS
  // is read only by cells that are not needed either.
  int jend = nsub;
  {
    const int core_lo = cz0 - Z0, core_hi = core_lo + S::H, grid_hi = nz - Z0;
#pragma unroll
    for (int j = T - 1; j >= 0; --j) {
      const int e = 2 * (T - 1 - j);
      const int hi = core_hi + e < grid_hi ? core_hi + e : grid_hi;
      if (ly * R >= hi || ly * R + R <= core_lo - e) jend = j;
    }
  }
  // snapshot iterations (bits of E.snap) and their levels, counted from E.lvl
  const int snap = E.snap;
  int lvl = E.lvl;
  // injections: only the wave holding the source / receiver row takes them
  const int src_row = P.src_cell / nx, src_col = P.src_cell - src_row * nx;
  const int riz = P.riz;
  const int inj_row = fwd ? src_row : riz;
  const int inj = inj_row >= row0 && inj_row < row0 + R ? E.inj : 0;
  // receiver of this column on the back-propagation row: its index when the
  // column has exactly one, -1 with none, -2 with several
  int rsel = -1;
  if (!fwd && inj != 0 && col_in) {
    const int r0 = as_global(P.roff)[col], r1 = as_global(P.roff)[col + 1];
    rsel = r1 - r0 == 1 ? as_global(P.ridx)[r0] : (r1 > r0 ? -2 : -1);
  }
  // the thread's first cell in LDS buffer 0; buffer 1 lies BUF floats
  // further, so every LDS access is this one address plus a constant
  constexpr int BUF = S::LR * S::LS;
  float* const B = lds + (ly * R + 2) * S::LS + lx + 2;
#pragma unroll
  for (int q = 0; q < R; ++q) B[q * S::LS] = c[q];
  // The wave holding the source / receiver row loads the samples it injects
  // after each sub-step here, with the levels, into LDS slots of its own
  // lanes: no memory latency inside the sub-steps.
  float* const injv = lds + 2 * BUF + lx;
  if (inj != 0) {
#pragma unroll
    for (int j = 0; j < T; ++j) {
      if (inj >> j & 1) {
        float v = 0.0f;
        if (fwd) v = as_global(P.src)[E.it0 + j + 1];
        else if (rsel >= 0) v = as_global(P.data)[(long long)(E.it0 - j - 1) * P.nrec + rsel];
        injv[j * S::CX] = v;
      }
    }
  }
  __syncthreads();

  float* snaps = P.snaps;
  const long long ncell = (long long)nx * nz;
  const float w = bc.w;

  // The sub-steps are unrolled (at most T of them), so the rotation of the
  // three levels is a renaming of registers and each LDS buffer's address is
  // fixed per sub-step.
#pragma unroll
  for (int j = 0; j < T; ++j) {
    if (j == nsub) break;
    if (j == jend) {
      for (int r = j; r < nsub; ++r) __syncthreads();
      return;
    }
    // side work on the post-injection level `c`
    if (snap >> j & 1) {
      // snapshots stream through HBM once each way: keep them out of the caches
      float* const ps = snaps + (long long)lvl * ncell;
      lvl += fwd ? 1 : -1;
      if (fwd) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (TB_CORE(q)) {
            st_at_nt(c[q], ps, lane4, roff[q]);
            acc[q] = acc[q] + c[q] * c[q];
          }
        }
      } else {
        // every row's load is issued before the first use, so the rows share
        // one memory latency instead of queueing one after another
        float sv[R];
#pragma unroll
        for (int q = 0; q < R; ++q) sv[q] = TB_CORE(q) ? ld_at_nt(ps, lane4, roff[q]) : 0.0f;
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (TB_CORE(q)) acc[q] = acc[q] + (sv[q] * c[q]) * w;
      }
    }
    // leapfrog update, evaluated on every lane (all reads stay inside this
    // LDS buffer)
    float n[R];
    const float* const Xb = B + (j & 1) * BUF;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const float* const xr = Xb + q * S::LS;
      const float pc = c[q];
      const float xm2 = xr[-2], xm1 = xr[-1], xp1 = xr[1], xp2 = xr[2];
      const float zm2 = q >= 2 ? c[q - 2] : xr[-2 * S::LS];
      const float zm1 = q >= 1 ? c[q - 1] : xr[-S::LS];
      const float zp1 = q + 1 < R ? c[q + 1] : xr[S::LS];
      const float zp2 = q + 2 < R ? c[q + 2] : xr[2 * S::LS];
      float lx2 = TSW_C2 * (xm2 + xp2);
      lx2 = lx2 + TSW_C1 * (xm1 + xp1);
      lx2 = (lx2 + TSW_C0 * pc) * bc.rdx2;
      float lz2 = TSW_C2 * (zm2 + zp2);
      lz2 = lz2 + TSW_C1 * (zm1 + zp1);
      lz2 = (lz2 + TSW_C0 * pc) * bc.rdz2;
      float v = (2.0f * pc - bl[q]) + cf[q] * (lx2 + lz2);
      n[q] = v * sp[q];
    }
    if (edge) {
      // the empty volatile asm keeps this a branch that only waves holding
      // cells outside [2,nz-2) x [2,nx-2) take, instead of selects on every wave
      asm volatile("");
      // kept cells (and cells outside the grid, which stay 0) keep `a`
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (!TB_INT(q)) n[q] = a[q];
      if (fs0) {  // this thread holds rows 0, 1 and 2 of its column
        n[0] = 0.0f;
        n[1] = -n[2];
      }
    }
    // injection of the next iteration's samples
    if (inj >> j & 1) {
      asm volatile("");
      const float smp = injv[j * S::CX];
      if (fwd) {
        if (col == src_col) {
#pragma unroll
          for (int q = 0; q < R; ++q)
            if (row0 + q == src_row) n[q] = n[q] + smp * bc.inj;
        }
      } else if (rsel >= 0) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (row0 + q == riz) n[q] = n[q] + smp * bc.inj;
      } else if (rsel == -2) {
        const gfloat* drow = as_global(P.data) + (long long)(E.it0 - j - 1) * P.nrec;
        const int r0 = as_global(P.roff)[col], r1 = as_global(P.roff)[col + 1];
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (row0 + q == riz)
            for (int r = r0; r < r1; ++r) n[q] = n[q] + drow[as_global(P.ridx)[r]] * bc.inj;
      }
    }
    float* const Yb = B + ((j + 1) & 1) * BUF;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      Yb[q * S::LS] = n[q];
      a[q] = bl[q];
      bl[q] = c[q];
      c[q] = n[q];
    }
    __syncthreads();
  }

  // ---- write back the core
  const int wr = rd ^ 1;
  float* const pacc = fwd ? P.illum : P.image;
  float* const pw2 = P.set[wr][2];
  float* const pw1 = P.set[wr][1];
  float* const pw0 = P.set[wr][0];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (TB_CORE(q)) {
      if (!last) {
        st_at(c[q], pw2, lane4, roff[q]);
        st_at(bl[q], pw1, lane4, roff[q]);
        if (TB_KEEP(q)) st_at(a[q], pw0, lane4, roff[q]);
      }
      st_at((fwd && last) ? acc[q] * w : acc[q], pacc, lane4, roff[q]);
    }
  }
#undef TB_IN
#undef TB_INT
#undef TB_CORE
#undef TB_KEEP
}

template <int T, int R, int WX, int WZ>
__global__ __launch_bounds__(64 * WX * WZ)
__attribute__((amdgpu_waves_per_eu(kTbWavesPerEu<R, WX, WZ>))) void k_tb(
    BatchConst bc, const TbShot* shots, const TbEntry* ents, const TbTile* tiles) {
  using S = TbShape<T, R, WX, WZ>;
  __shared__ float lds[2 * S::LR * S::LS + T * S::CX];
  const TbTile tile = tiles[blockIdx.x];
  const TbEntry& E = ents[tile.ent];
  const TbShot& P = shots[E.shot];
  const int tx = tile.txz & 0xffff, tz = tile.txz >> 16;
  tb_tile<T, R, WX, WZ>(bc, E, P, tx, tz, lds);
}

}  // namespace tsw
