// Temporally blocked acoustic RTM step kernel with the time levels in LDS:
// one launch advances T time steps of every active shot, like k_tb (tb.h),
// on taller tiles.
//
// A workgroup of 16 waves owns a core tile of W x H cells (W = 64 - 4T,
// H = 16R - 4T) and loads it with a 2T-cell halo (clipped at the grid); the
// lanes of a wave are the 64 columns of the extended tile and wave `ly` owns
// its rows ly*R .. ly*R+R-1.  LDS holds three time levels of the extended
// tile, level L in buffer L mod 3, as the reference's three rotating
// wavefield buffers do: a sub-step reads level j (neighbours) and writes
// level j+1 over level j-2.  Cells outside [2,nz-2) x [2,nx-2) keep the value
// of three levels back (wave2d.c:19-82), which is what their slot of the
// target buffer already holds, so no thread keeps that level in registers;
// registers hold the current level of the thread's cells, the model
// coefficient, the sponge factor and the image / illumination accumulator,
// and the previous level is read back from the thread's own LDS slots.  One
// barrier per sub-step; after s sub-steps the cells whose distance to a cut
// (non-grid) tile edge is >= 2s are exact, so after T sub-steps the core is
// exact and is written back.
//
// Per sub-step the kernel reproduces the reference loop exactly
// (tsw_rtm/wave_solver.py:157-183 forward, :210-242 backward): snapshot +
// illumination or zero-lag imaging of the post-injection level, the 4th-order
// update with the reference's FP32 association, free-surface rows (row 0 = 0,
// row 1 = -row 2, all three rows owned by one thread), kept cells, and
// injection of the next iteration's samples.  Results are bit-identical to
// the per-step kernel.
#pragma once
#include "tb.h"

namespace tsw {

template <int T, int R>
struct TcShape {
  static constexpr int CX = 64;             // extended columns (one wave)
  static constexpr int CZ = 16 * R;         // extended rows (16 waves)
  static constexpr int W = CX - 4 * T;      // core columns
  static constexpr int H = CZ - 4 * T;      // core rows
  static constexpr int BUF = CZ * CX;       // one time level, rows of CX floats, no padding
  // Two rows of slack before the first and after the last buffer keep the
  // reads of the tile's outermost rows in the allocation; lanes 0, 1, 62, 63
  // read the neighbouring row's ends.  Both feed only halo cells that the
  // core never depends on, or kept cells whose update is discarded.
  static constexpr int SLACK = 2 * CX;
  static constexpr int INJ = SLACK + 3 * BUF + SLACK;  // T slots of CX injection samples
  static constexpr int LDS = INJ + T * CX;             // floats
  static_assert(W > 0 && H > 0, "halo larger than the tile");
  static_assert(R >= 3, "rows 0..2 of a free surface must belong to one thread");
};

using lfloat = __attribute__((address_space(3))) float;
// `p` as a value the compiler takes as given: it keeps one register for it
// instead of folding each access's constant into a new address
__device__ __forceinline__ const lfloat* lds_reg(const lfloat* p) {
  asm volatile("" : "+v"(p));
  return p;
}

template <int T, int R>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_tc(
    BatchConst bc, const TbShot* shots, const TbEntry* ents, const TbTile* tiles) {
  using S = TcShape<T, R>;
  __shared__ float lds[S::LDS];
  const TbTile tile = tiles[blockIdx.x];
  const TbEntry& E = ents[tile.ent];
  const TbShot& P = shots[E.shot];
  const int tx = tile.txz & 0xffff, tz = tile.txz >> 16;
  const bool fwd = E.phase == 0;
  const int nx = P.nx, nz = P.nz;
  const int cx0 = tx * S::W, cz0 = tz * S::H;
  const int X0 = cx0 - 2 * T > 0 ? cx0 - 2 * T : 0;
  const int Z0 = cz0 - 2 * T > 0 ? cz0 - 2 * T : 0;
  const int lx = threadIdx.x;
  // the lanes of a wave share threadIdx.y: the rows are uniform over the wave
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
  unsigned m_keep = m_in & ~m_int & ~(fs0 ? 3u : 0u);
  // whether the wave holds any cell outside [2,nz-2) x [2,nx-2) (free-surface
  // rows and cells past the grid included); uniform
  const bool edge = !(X0 >= 2 && X0 + S::CX <= nx - 2 && row0 >= 2 && row0 + R <= nz - 2);
#define TC_INT(q) (m_int >> (q) & 1u)
#define TC_CORE(q) (m_core >> (q) & 1u)
#define TC_KEEP(q) (m_keep >> (q) & 1u)

  // the thread's first cell in buffer 0; level L lies in buffer L mod 3, and
  // every LDS access is this address plus a constant
  float* const B0 = lds + S::SLACK + ly * R * S::CX + lx;
#define TC_LEVEL(L) (B0 + (((L) % 3 + 3) % 3) * S::BUF)

  float bl[R], c[R], cf[R], sp[R], acc[R];

  // ---- load (or generate) the input levels 0, -1 (registers; kept cells also
  // in LDS) and -2 (kept cells, LDS only).  Lanes past the grid read the
  // clamped last column / row: their values reach only cells outside
  // [2,nz-2) x [2,nx-2), whose update is discarded, and are never written
  // back.  Every access is at the lane's byte offset (vector register) plus
  // the row's byte offset (scalar register).
  const int colc = col_in ? col : nx - 1;
  const unsigned lane4 = (unsigned)colc * 4u;
  const unsigned rowb = (unsigned)nx * 4u;
  unsigned roff[R];  // byte offset of row min(row0 + q, nz - 1)
#pragma unroll
  for (int q = 0; q < R; ++q) roff[q] = (unsigned)(row0 + q < nz ? row0 + q : nz - 1) * rowb;
  const int rd = E.rd;
  // sponge depth: the larger of the column's and the row's (the reference's
  // maximum over the four edges, split); table[0] = 1
  const int ns = bc.n_sponge;
  int dcol = ns - colc;
  dcol = dcol > colc - (nx - 1 - ns) ? dcol : colc - (nx - 1 - ns);
  dcol = dcol > 0 ? dcol : 0;
  {
    const float* const pcoef = fwd ? P.coef_f : P.coef_b;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int row = row0 + q;
      const int rowc = row < nz ? row : nz - 1;
      int drow = rowc - (nz - 1 - ns);
      if (!bc.fs) drow = drow > ns - rowc ? drow : ns - rowc;
      const int d = dcol > drow ? dcol : drow;
      cf[q] = ld_at(pcoef, lane4, roff[q]);
      sp[q] = ld_at(bc.table, (unsigned)d * 4u, 0);
      bl[q] = 0.0f;
      c[q] = 0.0f;
      acc[q] = 0.0f;
    }
  }
  float a0[R];  // level -2 of kept cells
#pragma unroll
  for (int q = 0; q < R; ++q) a0[q] = 0.0f;
  if (first) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (m_in >> q & 1u) {
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
      if (TC_KEEP(q)) a0[q] = ld_at(p0, lane4, roff[q]);
      if (TC_CORE(q)) acc[q] = ld_at(pacc, lane4, roff[q]);
    }
  }

  // ---- the wave's schedule (uniform).  Local rows [core_lo - e, core_hi + e)
  // of level j+1, e = 2(T-1-j), are the ones the core at level T depends on;
  // rows past the grid feed only rows outside [2, nz-2).  From the first
  // sub-step `jend` whose range misses all of the wave's rows (then it has no
  // core rows) the wave only joins the remaining barriers; what it left in
  // LDS is read only by cells that are not needed either.
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

  // levels 0 and -1 (every cell; a sub-step reads level j-1 of its own
  // cells back) and -2 (the kept cells read it back)
#pragma unroll
  for (int q = 0; q < R; ++q) {
    TC_LEVEL(0)[q * S::CX] = c[q];
    TC_LEVEL(-1)[q * S::CX] = bl[q];
  }
  if (edge) {
#pragma unroll
    for (int q = 0; q < R; ++q) TC_LEVEL(-2)[q * S::CX] = a0[q];
  }
  // the samples the injecting wave adds after each sub-step, in LDS slots of
  // its own lanes: no memory latency inside the sub-steps
  float* const injv = lds + S::INJ + lx;
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
  // Back-propagation reads a snapshot level every k sub-steps: its rows are
  // loaded one sub-step ahead (the first in the prologue), so that the
  // imaging sub-step does not wait on HBM.
  float sv[R];
#pragma unroll
  for (int q = 0; q < R; ++q) sv[q] = 0.0f;
  if (!fwd && (snap & 1)) {
    const float* const ps = snaps + (long long)lvl * ncell;
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (TC_CORE(q)) sv[q] = ld_at_nt(ps, lane4, roff[q]);
  }
  // The sub-steps are unrolled (at most T of them), so the rotation of the
  // register levels is a renaming and each buffer's address is fixed.
#pragma unroll
  for (int j = 0; j < T; ++j) {
    if (j == nsub) break;
    // the row masks as values the compiler cannot carry from one sub-step to
    // the next: each sub-step tests their bits again instead of holding six
    // rows' lane masks in scalar registers through the whole loop
    asm volatile("" : "+v"(m_int), "+v"(m_core), "+v"(m_keep));
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
          if (TC_CORE(q)) {
            st_at_nt(c[q], ps, lane4, roff[q]);
            acc[q] = acc[q] + c[q] * c[q];
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (TC_CORE(q)) acc[q] = acc[q] + (sv[q] * c[q]) * w;
      }
    }
    if (!fwd && (snap >> (j + 1) & 1) && j + 1 < nsub) {
      // the next imaging sub-step's snapshot rows (level `lvl` by now)
      const float* const ps = snaps + (long long)lvl * ncell;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (TC_CORE(q)) sv[q] = ld_at_nt(ps, lane4, roff[q]);
    }
    // leapfrog update, evaluated on every lane
    const float* const X = TC_LEVEL(j);
    const float* const Xp = TC_LEVEL(j - 1);  // level j-1 of the thread's own cells
    float* const Y = TC_LEVEL(j + 1);
    // row bases for the LDS reads: rows -2..1 from Xa, 2..5 from Xb, 6.. from
    // Xc, so each row's reads are small positive offsets of one register
    const lfloat* const Xa = lds_reg((const lfloat*)(X - 2 * S::CX - 2));
    const lfloat* const Xb = lds_reg((const lfloat*)(X + 2 * S::CX - 2));
    const lfloat* const Xc = lds_reg((const lfloat*)(X + 6 * S::CX - 2));
    float n[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const lfloat* const xr = (q < 2 ? Xa + (q + 2) * S::CX : q < 6 ? Xb + (q - 2) * S::CX : Xc + (q - 6) * S::CX) + 2;
      const float pc = c[q];
      const float xm2 = xr[-2], xm1 = xr[-1], xp1 = xr[1], xp2 = xr[2];
      const float zm2 = q >= 2 ? c[q - 2] : xr[-2 * S::CX];
      const float zm1 = q >= 1 ? c[q - 1] : xr[-S::CX];
      const float zp1 = q + 1 < R ? c[q + 1] : xr[S::CX];
      const float zp2 = q + 2 < R ? c[q + 2] : xr[2 * S::CX];
      float lx2 = TSW_C2 * (xm2 + xp2);
      lx2 = lx2 + TSW_C1 * (xm1 + xp1);
      lx2 = (lx2 + TSW_C0 * pc) * bc.rdx2;
      float lz2 = TSW_C2 * (zm2 + zp2);
      lz2 = lz2 + TSW_C1 * (zm1 + zp1);
      lz2 = (lz2 + TSW_C0 * pc) * bc.rdz2;
      float v = (2.0f * pc - Xp[q * S::CX]) + cf[q] * (lx2 + lz2);
      n[q] = v * sp[q];
    }
    if (edge) {
      // the empty volatile asm keeps this a branch that only waves holding
      // cells outside [2,nz-2) x [2,nx-2) take
      asm volatile("");
      // kept cells (and cells outside the grid) keep level j-2, which their
      // slot of the target buffer holds
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (!TC_INT(q)) n[q] = Y[q * S::CX];
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
#pragma unroll
    for (int q = 0; q < R; ++q) {
      Y[q * S::CX] = n[q];
      c[q] = n[q];
    }
    __syncthreads();
  }

  // ---- write back the core: level nsub from registers, level nsub-1 and
  // (kept cells) nsub-2 from the thread's own LDS slots
  const int wr = rd ^ 1;
  float* const pacc = fwd ? P.illum : P.image;
  float* const pw2 = P.set[wr][2];
  float* const pw1 = P.set[wr][1];
  float* const pw0 = P.set[wr][0];
  const float* const Ak = TC_LEVEL(nsub - 2);
  const float* const Bk = TC_LEVEL(nsub - 1);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (TC_CORE(q)) {
      if (!last) {
        st_at(c[q], pw2, lane4, roff[q]);
        st_at(Bk[q * S::CX], pw1, lane4, roff[q]);
        if (TC_KEEP(q)) st_at(Ak[q * S::CX], pw0, lane4, roff[q]);
      }
      st_at((fwd && last) ? acc[q] * w : acc[q], pacc, lane4, roff[q]);
    }
  }
#undef TC_INT
#undef TC_CORE
#undef TC_KEEP
#undef TC_LEVEL
}

template <int T, int R>
void launch_tc(const BatchConst& bc, const TbShot* shots, const TbEntry* ents, const TbTile* tiles,
               int ntiles, hipStream_t st) {
  hipLaunchKernelGGL((k_tc<T, R>), dim3(ntiles), dim3(64, 16), 0, st, bc, shots, ents, tiles);
}

}  // namespace tsw
