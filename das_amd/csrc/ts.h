// Temporally blocked acoustic RTM step kernel that streams down z: one launch
// advances T time steps of every active shot, like k_tb, without a z-halo
// around every tile.
//
// A workgroup owns a strip of CX = 64 WX columns of one shot (a core of
// W = CX - 4T columns plus a 2T-column halo each side, as k_tb) over a depth
// segment of H core rows, and has one wavefront per sub-step and 64 columns:
// the waves of stage m (1..T) compute level m.  Thread (lx, m-1) owns column
// X0 + lx of level m.  Rows are cut into chunks of two, starting at an odd
// row so that rows 1 and 2 (the free surface's mirror pair) share a chunk.
// At step s, stage m computes chunk j = s - 2m from level m-1 of chunks
// j-1, j, j+1, all of which earlier steps produced, so the T stages of a step
// are independent and one barrier per step orders everything.  Stage m only
// visits the chunks the segment's core depends on (2 (T-m) rows either side
// of it, clipped to the grid).
//
// Levels pass between stages through LDS rings of chunks (slot = chunk mod 8,
// column-contiguous rows with two pad columns each side): stage m writes its
// chunk once; stage m+1 reads it as the z+1 neighbour of its chunk (into a
// three-chunk register window) and, a step later, its x-neighbours; stage m+2
// reads it as the leapfrog level two back and stage m+3 as the level three
// back that kept cells carry.  Stage 1's waves also bring the input levels
// (two state levels and the level three back) and the image / illumination
// accumulator of each chunk in from HBM, one chunk ahead; the accumulator
// sits in its own ring (slot = chunk mod 16) while the imaging stages pass
// over it, and stage T writes back the core.
//
// Per sub-step every cell follows the reference loop exactly, as k_tb does
// (tsw_rtm/wave_solver.py:157-183 forward, :210-242 backward,
// tsw_rtm/c_engine/wave2d.c:19-82): side work on the post-injection level,
// the FP32 update in the reference's association, cells outside
// [2,nz-2) x [2,nx-2) keeping the level three back, the free-surface rows,
// next-iteration injection.  A block of nsub < T sub-steps runs the stages
// past nsub as copies of the level three back, which rotates the three state
// levels by (T - nsub) mod 3; the write-back undoes the rotation.  Results are
// bit-identical to k_tb and to the per-step kernel.
#pragma once
#include <type_traits>

#include "tb.h"

namespace tsw {

// f(integral_constant<int, I>) for I = B .. E-1, unrolled at compile time.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Largest sponge strip the streaming kernel's LDS copy of the table holds.
constexpr int kTsMaxSponge = 255;

// Steps ahead of use at which chunk inputs are loaded from HBM.
constexpr int kTsDepth = 4;

template <int T, int WX>
struct TsShape {
  static constexpr int CX = 64 * WX;    // strip columns
  static constexpr int W = CX - 4 * T;  // core columns
  static constexpr int CP = CX + 4;     // LDS row stride (2 pad columns each side)
  static_assert(W > 0, "halo wider than the strip");
  static_assert(T >= 4 && T <= 7, "four input arrays over the stage waves; accumulator ring of 16 chunks");
};

// The 4th-order leapfrog update of one cell in the reference's FP32
// association (wave2d.c:49-65), before the sponge factor.
__device__ __forceinline__ float ts_update(float xm2, float xm1, float xp1, float xp2, float zm2,
                                           float zm1, float zp1, float zp2, float c, float pv,
                                           float cf, float rdx2, float rdz2) {
  float lx2 = TSW_C2 * (xm2 + xp2);
  lx2 = lx2 + TSW_C1 * (xm1 + xp1);
  lx2 = (lx2 + TSW_C0 * c) * rdx2;
  float lz2 = TSW_C2 * (zm2 + zp2);
  lz2 = lz2 + TSW_C1 * (zm1 + zp1);
  lz2 = (lz2 + TSW_C0 * c) * rdz2;
  return (2.0f * c - pv) + cf * (lx2 + lz2);
}

// LDS of a k_ts workgroup.
template <int T, int WX>
struct TsLds {
  static constexpr int CX = 64 * WX, CP = CX + 4;
  // chunk rings, [slot][row][column + 2]: level l = 0..T-1 in slots
  // 8 l + (chunk & 7), level T in 8 T + (chunk & 3), level -1 in
  // 8 T + 4 + (chunk & 7), level -2 in 8 T + 12 + (chunk & 3)
  static constexpr int NSLOT = 8 * T + 16;
  float lv[NSLOT][2][CP];
  // accumulator: [chunk & 15][row][column]; a column's threads alone touch it
  float acr[16][2][CX];
  // injected sample of sub-step m (index m-1) per column
  float sj[T][CX];
  float tab[kTsMaxSponge + 1];
};

// One workgroup's block of one shot: phase (FWD), first block of the phase
// (FIRST: the state is generated, not loaded) and columns with several
// receivers in the strip (DUP, back-propagation only: their samples are read
// inside the step loop) are compile-time.  Wave
// row y < T runs stage m = y + 1 (stage 1 also brings the input chunks in);
// wave row T writes everything back to HBM, so the stage waves issue loads
// only, a fixed number per step, and each wait covers exactly the loads it
// needs.
template <int T, int WX, int H, bool FWD, bool FIRST, bool DUP>
__device__ __forceinline__ void ts_block(const BatchConst& bc, const TbShot& P, const TbEntry& E,
                                         const int t, TsLds<T, WX>& sh) {
  using S = TsShape<T, WX>;
  constexpr int CX = S::CX, W = S::W, CP = S::CP, SL = 2 * CP;
  constexpr int D = kTsDepth;
  const int m = __builtin_amdgcn_readfirstlane(threadIdx.y) + 1;  // stage; T+1: the writer
  const bool writer = m > T;
  const int lx = threadIdx.x;
  const int nx = P.nx, nz = P.nz, nt = P.nt, k = P.k, nrec = P.nrec, riz = P.riz;
  const int cx0 = (t % E.tiles_x) * W;
  const int cz0 = (t / E.tiles_x) * H;
  const int cz1 = cz0 + H < nz ? cz0 + H : nz;
  const int X0 = cx0 - 2 * T > 0 ? cx0 - 2 * T : 0;
  const int col = X0 + lx;
  const bool last = (E.flags & kTbLast) != 0;
  const int nsub = E.nsub;
  const int it0 = E.it0;
  const long long ncell = (long long)nx * nz;

  // chunk j holds rows rbase + 2j and rbase + 2j + 1 (rbase odd, <= the
  // first row level 0 needs); level l is needed on rows
  // [zl(l), zh(l)) = [cz0 - 2(T-l), cz1 + 2(T-l)) clipped to the grid, so
  // stage l visits the chunks that start in [zl(l) - 1, zh(l))
  const int zlo = cz0 - 2 * T > 0 ? cz0 - 2 * T : 0;
  const int rbase = zlo - ((zlo & 1) ^ 1);
  const auto zl = [=](int l) __attribute__((always_inline)) { return cz0 - 2 * (T - l) > 0 ? cz0 - 2 * (T - l) : 0; };
  const auto zh = [=](int l) __attribute__((always_inline)) { return cz1 + 2 * (T - l) < nz ? cz1 + 2 * (T - l) : nz; };
  const auto jlo_of = [=](int l) __attribute__((always_inline)) { return (zl(l) - rbase) >> 1; };
  const auto jhi_of = [=](int l) __attribute__((always_inline)) { return (zh(l) - rbase + 1) >> 1; };
  const int ncl = jhi_of(0);  // chunks of level 0 (loaded)
  const int ms = writer ? T : m;
  const int jlo = jlo_of(ms), jhi = jhi_of(ms);  // this stage's chunks
  // every wave takes every step; the writer stores chunk j of level T at
  // step j + 2T + 1.  A multiple of D, so that every step of the unrolled
  // loop issues its loads.
  const int nsteps = (jhi_of(T) + 2 * T + 1 + D - 1) / D * D;

  const int rd = E.rd, wr = rd ^ 1;
  const gfloat* coef = uniform_ptr(as_global(FWD ? P.coef_f : P.coef_b));
  gfloat* wC = uniform_ptr(as_global(P.set[wr][2]));
  gfloat* wB = uniform_ptr(as_global(P.set[wr][1]));
  gfloat* wA = uniform_ptr(as_global(P.set[wr][0]));
  const gfloat* sC = uniform_ptr(as_global(P.set[rd][2]));
  const gfloat* sB = uniform_ptr(as_global(P.set[rd][1]));
  const gfloat* sA0 = uniform_ptr(as_global(P.set[rd][0]));
  gfloat* acc_g = uniform_ptr(as_global(FWD ? P.illum : P.image));
  gfloat* snaps = as_global(P.snaps);
  const gfloat* data = as_global(P.data);
  const gfloat* src = as_global(P.src);
  const gint* roff = as_global(P.roff);
  const gint* ridx = as_global(P.ridx);

  const bool col_in = col < nx;
  const int colc = col_in ? col : nx - 1;  // clamped column of the loads
  const unsigned cb = (unsigned)colc * 4u;  // its byte offset in a row
  // byte offset of this lane's cell of row `row` in a (nz, nx) array
  const auto cell = [=](int row) __attribute__((always_inline)) {
    return __builtin_amdgcn_readfirstlane((unsigned)(row * nx) * 4u) + cb;
  };
  const bool core_col = col_in && col >= cx0 && col < cx0 + W;
  const bool col_int = col >= 2 && col < nx - 2;
  // the strip holds edge columns (cells kept on every row)
  const bool edge_strip = X0 < 2 || X0 + CX > nx - 2;
  const int ns = bc.n_sponge;
  int dcol = ns - col;
  dcol = col - (nx - 1 - ns) > dcol ? col - (nx - 1 - ns) : dcol;
  dcol = dcol > 0 ? dcol : 0;
  dcol = dcol < ns ? dcol : ns;  // columns past the grid: any entry

  // injection: row and per-lane samples of every sub-step
  const int src_row = P.src_cell / nx, src_col = P.src_cell - src_row * nx;
  const int inj_row = FWD ? src_row : riz;
  int r0 = 0, r1 = 0;
  if (!FWD && col_in) {
    r0 = roff[col];
    r1 = roff[col + 1];
  }
  const bool inj_lane = FWD ? (col_in && src_col == col) : r1 > r0;
  const bool dup = !FWD && r1 - r0 > 1;
  const bool inj_ok = inj_lane && !dup;
  const float inj = bc.inj, w = bc.w;
  // iteration of sub-step l; whether sub-step l images / injects
  const auto iter = [=](int l) __attribute__((always_inline)) { return FWD ? it0 + (l - 1) : it0 - (l - 1); };
  const auto images = [=](int l) __attribute__((always_inline)) { return l <= nsub && iter(l) % k == 0; };
  const int itm = iter(m);
  const bool img = !writer && images(m);
  const bool injs = !writer && m <= nsub && (FWD ? itm + 1 < nt : itm >= 1);
  if (m == 1) {  // samples of every sub-step for this column
    const int rsel = r1 - r0 == 1 ? ridx[r0] : 0;
    for (int l = 1; l <= T; ++l) {
      const int it = iter(l);
      float v = 0.0f;
      if (l <= nsub) {
        if (FWD) {
          if (it + 1 < nt) v = src[it + 1] * inj;
        } else if (it >= 1 && !dup) {
          v = data[(long long)(it - 1) * nrec + rsel] * inj;
        }
      }
      sh.sj[l - 1][lx] = v;
    }
  }
  // the value that generates iteration 0 in the first block of a phase
  float g0 = 0.0f;
  if (FIRST && m == 1 && inj_lane) {
    if (FWD) {
      g0 = 0.0f + src[0] * inj;
    } else {
      const gfloat* drow = data + (long long)(nt - 1) * nrec;
      for (int r = r0; r < r1; ++r) g0 = g0 + drow[ridx[r]] * inj;
    }
  }
  const int tid = threadIdx.y * CX + lx;
  constexpr int NTH = CX * (T + 1);
  for (int i = tid; i <= ns && i <= kTsMaxSponge; i += NTH) sh.tab[i] = bc.table[i];
  // pad columns of the rings: read by the strip's outermost lanes, whose
  // results never reach the core
  for (int i = tid; i < TsLds<T, WX>::NSLOT * 2 * 4; i += NTH) {
    const int r = i >> 2, c = i & 3;
    sh.lv[r >> 1][r & 1][c < 2 ? c : CX + c] = 0.0f;
  }
  // several receivers on one column of this strip: the slow injection path
  const bool dup_block = __syncthreads_or(dup && inj_lane) != 0;
  if constexpr (!FWD && !DUP) {
    if (dup_block) {
      ts_block<T, WX, H, FWD, FIRST, true>(bc, P, E, t, sh);
      return;
    }
  }
  const float sp_col = sh.tab[dcol];  // entry 0 is expf(-0) = 1: no sponge

  // row classification (uniform over the workgroup)
  const auto row_int = [=](int row) __attribute__((always_inline)) { return row >= 2 && row < nz - 2; };
  const auto row_core = [=](int row) __attribute__((always_inline)) { return row >= cz0 && row < cz1; };
  // kept cell: in the grid, outside the interior, not a free-surface row
  const auto kept = [=](int row) __attribute__((always_inline)) {
    return col_in && row >= 0 && row < nz && !(col_int && row_int(row)) && !(bc.fs && row < 2);
  };
  // depth into the top / bottom sponge strip of a row (<= 0 outside them)
  const auto drow_of = [=](int row) __attribute__((always_inline)) {
    int d = row - (nz - 1 - ns);
    if (!bc.fs && ns - row > d) d = ns - row;
    return d;
  };

  // ring slot of level l (-2..T) chunk c; this lane's word of it in row q
  const auto slot = [=](int l, int c) __attribute__((always_inline)) {
    return l >= 0 ? (l < T ? 8 * l + (c & 7) : 8 * T + (c & 3))
                  : (l == -1 ? 8 * T + 4 + (c & 7) : 8 * T + 12 + (c & 3));
  };
  float* const own = &sh.lv[0][0][lx + 2];  // this lane's column in slot 0, row 0
  const auto at_lv = [=](int sl, int q) __attribute__((always_inline)) -> float& { return own[sl * SL + q * CP]; };

  // ---- the input chunks (levels 0, -1, -2 and the accumulator, one array
  // per stage wave 1..4; wave 5 and up re-read level 0) and every stage's
  // coefficient and (back-propagation imaging) snapshot rows are loaded D
  // steps ahead of use, into register set (step mod D).  Every wave issues
  // the same loads every step (rows clamped into the grid; a chunk that needs
  // no level -2, accumulator or snapshot re-reads the line of level 0 or of
  // the coefficient instead), so the waits of the step loop cover exactly
  // the loads they need.
  float pin[D][2], cf[D][2], sv[D][2];
#pragma unroll
  for (int d = 0; d < D; ++d)
#pragma unroll
    for (int q = 0; q < 2; ++q) pin[d][q] = cf[d][q] = sv[d][q] = 0.0f;
  // this wave's input array, and the chunks for which it needs its own one
  const gfloat* const in_arr = m == 1 ? sC : (m == 2 ? sB : (m == 3 ? sA0 : (m == 4 ? (const gfloat*)acc_g : sC)));
  const gfloat* snap_m = FWD || !img ? coef : uniform_ptr(snaps + (long long)(itm / k) * ncell);

  // sub-steps that image (bit l) and the snapshot level of the first of
  // them; the next ones are one level further each
  unsigned imgbits = 0;
  int lvl1 = 0;
#pragma unroll
  for (int l = T; l >= 1; --l)
    if (images(l)) {
      imgbits |= 1u << l;
      lvl1 = iter(l) / k;
    }
  const auto snap_of = [=](int l) __attribute__((always_inline)) {
    const int n = __builtin_popcount(imgbits & ((1u << l) - 1u));  // imaging sub-steps before l
    return uniform_ptr(snaps + (long long)(FWD ? lvl1 + n : lvl1 - n) * ncell);
  };
  const auto core_chunk = [=](int row0) __attribute__((always_inline)) {
    return row0 + 1 >= cz0 && row0 < cz1;
  };

  if (writer) {
    // ---- the writer: snapshots and the core of level T, one step after
    // they are produced (core chunks are visited by every stage)
    for (int s = 0; s < nsteps; ++s) {
      if (FWD) {  // snapshots of level l when sub-step l+1 images
#pragma unroll
        for (int l = 0; l < T; ++l) {
          const int c = s - 2 * l - 1;
          const int row0 = rbase + 2 * c;
          if ((imgbits >> (l + 1) & 1u) && core_chunk(row0) && core_col) {
            gfloat* snap = snap_of(l + 1);
            if (row_core(row0)) __builtin_nontemporal_store(at_lv(slot(l, c), 0), at(snap, cell(row0)));
            if (row_core(row0 + 1))
              __builtin_nontemporal_store(at_lv(slot(l, c), 1), at(snap, cell(row0 + 1)));
          }
        }
      }
      // the core of level T's chunk; the state is (level nsub, nsub-1,
      // nsub-2), which the copies past nsub rotated
      const int c = s - 2 * T - 1;
      const int row0 = rbase + 2 * c;
      if (core_chunk(row0) && core_col) {
        const int rot = (T - nsub) % 3;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int row = row0 + q;
          if (row_core(row)) {
            const float a = sh.acr[c & 15][q][lx];
            const float l0 = at_lv(slot(T, c), q), l1 = at_lv(slot(T - 1, c), q),
                        l2 = at_lv(slot(T - 2, c), q);
            const unsigned off = cell(row);
            *at(acc_g, off) = (FWD && last) ? a * w : a;
            if (!last) {
              *at(wC, off) = rot == 0 ? l0 : (rot == 1 ? l1 : l2);
              *at(wB, off) = rot == 0 ? l1 : (rot == 1 ? l2 : l0);
              if (kept(row)) *at(wA, off) = rot == 0 ? l2 : (rot == 1 ? l0 : l1);
            }
          }
        }
      }
      __syncthreads();
    }
    return;
  }

  // ---- a stage wave.  Its rings: input level m-1, levels m-2 (leapfrog)
  // and m-3 (kept cells), output level m; ring (l) starts at slot
  // ring_base(l) and has (ring_mask(l) + 1) slots.
  const auto ring_base = [=](int l) __attribute__((always_inline)) {
    return l >= 0 ? (l < T ? 8 * l : 8 * T) : (l == -1 ? 8 * T + 4 : 8 * T + 12);
  };
  const auto ring_mask = [=](int l) __attribute__((always_inline)) { return l >= T || l == -2 ? 3 : 7; };
  float* const r_in = own + ring_base(m - 1) * SL;
  float* const r_pv = own + ring_base(m - 2) * SL;
  float* const r_kv = own + ring_base(m - 3) * SL;
  float* const r_out = own + ring_base(m) * SL;
  const int mk_kv = ring_mask(m - 3), mk_out = ring_mask(m);
  float* const r_acc = &sh.acr[0][0][lx];
  // rows where the update needs no per-row care: interior, outside the top
  // and bottom sponge strips, below the free-surface rows
  const int pr_lo = bc.fs ? 2 : (ns > 2 ? ns : 2);
  const int pr_hi = ns > 2 ? nz - ns : nz - 2;
  // a clamped row's byte offset in a row-major (nz, nx) array
  const auto row_off = [=](int row) __attribute__((always_inline)) {
    return (unsigned)((row < 0 ? 0 : (row < nz ? row : nz - 1)) * nx) * 4u;
  };

  // Steps -D .. -1 only issue the first loads.
  float zu0 = 0.0f, zu1 = 0.0f, zc0 = 0.0f, zc1 = 0.0f, zd0 = 0.0f, zd1 = 0.0f;  // window
  for (int s0 = -D; s0 < nsteps; s0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int s = s0 + u;
      const int j = s - 2 * m;
      if (s >= 0) {
        // waves 1..4: their array of chunk s in (generated in the first
        // block: level 0 holds iteration 0's injection, the rest is zero)
        if (m <= 4 && s < ncl) {
          const int row0 = rbase + 2 * s;
          float* dst = m == 4 ? &sh.acr[s & 15][0][lx] : &at_lv(slot(1 - m, s), 0);
          const int qs = m == 4 ? CX : CP;  // row stride of the target
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            float c = pin[u][q];
            if (FIRST) c = (m == 1 && inj_lane && row0 + q == inj_row) ? g0 : 0.0f;
            dst[q * qs] = c;
          }
        }
        // level m-1 of chunk j+1 into the window (written at step s-1)
        const int o1 = ((j + 1) & 7) * SL;
        zu0 = zc0;
        zu1 = zc1;
        zc0 = zd0;
        zc1 = zd1;
        zd0 = r_in[o1];
        zd1 = r_in[o1 + CP];
        if (j >= jlo && j < jhi) {
          const int row0 = rbase + 2 * j;
          const int o = (j & 7) * SL, okv = (j & mk_kv) * SL;
          float n0, n1;
          const float kv0 = r_kv[okv], kv1 = r_kv[okv + CP];
          if (m > nsub) {  // past the block's sub-steps: copy of the level three back
            n0 = kv0;
            n1 = kv1;
          } else {
            const float pv0 = r_pv[o], pv1 = r_pv[o + CP];
            // side work on the post-injection level of this sub-step's iteration
            if (img) {
              float* a = r_acc + (j & 15) * 2 * CX;
              if (FWD) {
                a[0] = a[0] + zc0 * zc0;
                a[CX] = a[CX] + zc1 * zc1;
              } else {
                a[0] = a[0] + (sv[u][0] * zc0) * w;
                a[CX] = a[CX] + (sv[u][1] * zc1) * w;
              }
            }
            const float* xr = r_in + o;  // level m-1 of chunk j, this column
            const float v0 = ts_update(xr[-2], xr[-1], xr[1], xr[2], zu0, zu1, zc1, zd0, zc0,
                                       pv0, cf[u][0], bc.rdx2, bc.rdz2);
            const float v1 = ts_update(xr[CP - 2], xr[CP - 1], xr[CP + 1], xr[CP + 2], zu1,
                                       zc0, zd0, zd1, zc1, pv1, cf[u][1], bc.rdx2, bc.rdz2);
            if (row0 >= pr_lo && row0 + 1 < pr_hi) {
              // kept cells (edge columns, cells outside the grid) carry the
              // level three back
              n0 = col_int ? v0 * sp_col : kv0;
              n1 = col_int ? v1 * sp_col : kv1;
            } else {
              // sponge factor table[max(dcol, drow)] (wave2d.c:247-262)
              const int d0 = drow_of(row0), d1 = drow_of(row0 + 1);
              float sp0 = sp_col, sp1 = sp_col;
              if (d0 > 0 || d1 > 0) {
                int e0 = d0 > dcol ? d0 : dcol, e1 = d1 > dcol ? d1 : dcol;
                e0 = e0 < ns ? e0 : ns;
                e1 = e1 < ns ? e1 : ns;
                sp0 = sh.tab[e0];
                sp1 = sh.tab[e1];
              }
              n0 = (row_int(row0) && col_int) ? v0 * sp0 : kv0;
              n1 = (row_int(row0 + 1) && col_int) ? v1 * sp1 : kv1;
              if (bc.fs && row0 <= 1) {  // row 0 = 0, row 1 = -row 2 on the grid's columns
                if (row0 == -1 && col_in) n1 = 0.0f;
                if (row0 == 1 && col_in) n0 = -n1;
              }
            }
            // injection of the next iteration's samples
            if (injs && (unsigned)(inj_row - row0) < 2u) {
              const bool q0 = inj_row == row0;
              float x = q0 ? n0 : n1;
              if (DUP && dup) {
                // a column with several receivers adds them in receiver order
                const gfloat* drow = data + (long long)(itm - 1) * nrec;
                for (int r = r0; r < r1; ++r) x = x + drow[ridx[r]] * inj;
              }
              if (inj_ok) x = x + sh.sj[m - 1][lx];
              if (q0) n0 = x;
              else n1 = x;
            }
          }
          const int oo = (j & mk_out) * SL;
          r_out[oo] = n0;
          r_out[oo + CP] = n1;
        }
      }
      // loads for step s + D: this wave's input array (rows of chunk s + D)
      // and the coefficient / snapshot rows of its chunk j + D
      if (!FIRST) {
        const int row0 = rbase + 2 * (s + D);
        // level -2 only where the chunk can hold kept cells, the
        // accumulator only where it holds core rows
        const bool need = m == 3   ? edge_strip || !row_int(row0) || !row_int(row0 + 1)
                          : m == 4 ? core_chunk(row0)
                                   : true;
        const gfloat* arr = need ? in_arr : sC;
        pin[u][0] = *at(uniform_ptr(arr + row_off(row0) / 4), fresh(cb));
        pin[u][1] = *at(uniform_ptr(arr + row_off(row0 + 1) / 4), fresh(cb));
      }
      {
        const int row0 = rbase + 2 * (j + D);
        const unsigned f0 = row_off(row0), f1 = row_off(row0 + 1);
        cf[u][0] = *at(uniform_ptr(coef + f0 / 4), fresh(cb));
        cf[u][1] = *at(uniform_ptr(coef + f1 / 4), fresh(cb));
        if (!FWD) {
          sv[u][0] = __builtin_nontemporal_load(at(uniform_ptr(snap_m + f0 / 4), fresh(cb)));
          sv[u][1] = __builtin_nontemporal_load(at(uniform_ptr(snap_m + f1 / 4), fresh(cb)));
        }
      }
      if (s >= 0) __syncthreads();
    }
  }
}

template <int T, int WX, int H>
__global__ __launch_bounds__(64 * WX * (T + 1)) void k_ts(BatchConst bc, const TbShot* shots,
                                                    const TbEntry* ents, const TbTile* tiles) {
  __shared__ TsLds<T, WX> sh;
  const TbTile tile = tiles[blockIdx.x];
  const TbEntry E = ents[tile.ent];
  const TbShot& P = shots[E.shot];
  const int t = (tile.txz >> 16) * E.tiles_x + (tile.txz & 0xffff);
  if (E.phase == 0) {
    if (E.flags & kTbFirst) ts_block<T, WX, H, true, true, false>(bc, P, E, t, sh);
    else ts_block<T, WX, H, true, false, false>(bc, P, E, t, sh);
  } else {
    if (E.flags & kTbFirst) ts_block<T, WX, H, false, true, false>(bc, P, E, t, sh);
    else ts_block<T, WX, H, false, false, false>(bc, P, E, t, sh);
  }
}

template <int T, int WX, int H>
void launch_ts(const BatchConst& bc, const TbShot* shots, const TbEntry* ents, const TbTile* tiles,
               int ntiles, hipStream_t st) {
  hipLaunchKernelGGL((k_ts<T, WX, H>), dim3(ntiles), dim3(64 * WX, T + 1), 0, st, bc, shots, ents,
                     tiles);
}

// The instantiated streaming configurations (T sub-steps, WX waves across,
// H core rows per depth segment), in table order after k_tb's; the kernels
// are compiled in ts_a.hip / ts_b.hip.
#define TSW_TS_CONFIGS_A(X) X(5, 1, 100) X(5, 2, 100) X(4, 1, 100) X(5, 1, 150)
#define TSW_TS_CONFIGS_B(X) X(6, 1, 100) X(5, 1, 75) X(4, 2, 100) X(5, 1, 300)

#define TSW_TS_EXTERN(T, WX, H)                                                              \
  extern template void launch_ts<T, WX, H>(const BatchConst&, const TbShot*, const TbEntry*, \
                                           const TbTile*, int, hipStream_t);
#define TSW_TS_INSTANTIATE(T, WX, H)                                                  \
  template void launch_ts<T, WX, H>(const BatchConst&, const TbShot*, const TbEntry*, \
                                    const TbTile*, int, hipStream_t);

}  // namespace tsw
