// Temporally blocked acoustic RTM step kernel that streams down z: one launch
// advances T time steps of every active shot, like k_tb, with no z-halo.
//
// A workgroup owns a strip of CX = 64 WX columns of one shot (a core of
// CX - 4T columns plus a 2T-column halo each side, as k_tb) over the shot's
// whole depth.  Thread lx owns column X0 + lx.  The depth is processed in
// chunks of R rows, top to bottom, as a pipeline over the T sub-steps: at
// step u, stage m (sub-step m, producing level m) advances chunk u - m + 1,
// so stage m finds the level m-1 chunk below it (its z+1 and z+2 neighbours)
// computed by stage m-1 earlier in the same step.  Each thread keeps, per
// level, the last four chunks of its column in registers (z-neighbours, the
// level two back, and the level three back that kept cells carry), in slots
// fixed relative to the stage, so the loop body is one step.  The
// x-neighbours of each level's newest chunk go through double-buffered LDS,
// with one barrier per step for all T stages.  The coefficient and the
// image / illumination accumulator of a chunk sit in per-thread LDS rings
// while the stages pass over it.
//
// The pipeline's fill and drain run on chunks outside the grid: their loads
// are clamped into the grid and their values never reach a cell of the grid
// (rows outside it are never read by an interior cell), and every store to
// device memory is masked to the core.  So the steady state carries no
// per-stage tests beyond the block's sub-step count.
//
// Per sub-step every cell follows the reference loop exactly, as k_tb does
// (tsw_rtm/wave_solver.py:157-183 forward, :210-242 backward,
// tsw_rtm/c_engine/wave2d.c:19-82): side work on the post-injection level,
// the FP32 update in the reference's association, the free-surface rows,
// kept edge cells, next-iteration injection.  Results are bit-identical to
// k_tb and to the per-step kernel.
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

template <int T, int R, int WX>
struct TsShape {
  static constexpr int CX = 64 * WX;    // strip columns
  static constexpr int W = CX - 4 * T;  // core columns
  static constexpr int LS = CX + 4;     // LDS row stride (2 pad columns each side)
  // chunk slots of the per-thread rings: at least T + 2, a power of two
  static constexpr int NR = T + 2 <= 4 ? 4 : (T + 2 <= 8 ? 8 : 16);
  static_assert(W > 0, "halo wider than the strip");
  static_assert(R >= 3, "rows 0..2 of a free surface must be in one chunk");
};

template <int T, int R, int WX, int H>
__global__ __launch_bounds__(64 * WX) void k_ts(BatchConst bc, const TbShot* shots,
                                                const TbEntry* ents, const int* pre, int nent,
                                                int ntiles) {
  using S = TsShape<T, R, WX>;
  constexpr int CX = S::CX, LS = S::LS, NR = S::NR;
  // x-exchange: level l's newest chunk, two buffers by chunk parity
  __shared__ float xs[T][2][R * LS];
  // per-thread rings (thread lx alone touches column lx): coefficient and
  // accumulator of the chunks the stages are passing over
  __shared__ float cfr[NR * R * CX];
  __shared__ float acr[NR * R * CX];
  __shared__ float tab[kTsMaxSponge + 1];
  __shared__ int dup_any;

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
  const int nx = P.nx, nz = P.nz, nt = P.nt, k = P.k, nrec = P.nrec, riz = P.riz;
  const int src_cell = P.src_cell;
  // strip and depth segment of this workgroup: core columns [cx0, cx0 + W),
  // core rows [cz0, cz1); a segment computes 2 nsub rows either side of its
  // core (the halo the sub-steps need), clipped to the grid
  const int cx0 = (t % E.tiles_x) * S::W;
  const int cz0 = (t / E.tiles_x) * H;
  const int cz1 = cz0 + H < nz ? cz0 + H : nz;
  const int zlo = cz0 - 2 * E.nsub > 0 ? cz0 - 2 * E.nsub : 0;
  const int zhi = cz1 + 2 * E.nsub < nz ? cz1 + 2 * E.nsub : nz;
  const int rbase = zlo / R * R;                 // first row of local chunk 0
  const int ncl = (zhi - rbase + R - 1) / R;     // chunks of the segment
  const int X0 = cx0 - 2 * T > 0 ? cx0 - 2 * T : 0;
  const int lx = threadIdx.x;
  const int col = X0 + lx;
  const bool fwd = E.phase == 0;
  const bool first = (E.flags & kTbFirst) != 0;
  const bool last = (E.flags & kTbLast) != 0;
  const int nsub = E.nsub;
  const int it0 = E.it0;
  const long long ncell = (long long)nx * nz;

  const int rd = E.rd, wr = rd ^ 1;
  const gfloat* coef = as_global(fwd ? P.coef_f : P.coef_b);
  const gfloat* sC = as_global(P.set[rd][2]);
  const gfloat* sB = as_global(P.set[rd][1]);
  const gfloat* sA = as_global(P.set[rd][0]);
  gfloat* wC = as_global(P.set[wr][2]);
  gfloat* wB = as_global(P.set[wr][1]);
  gfloat* wA = as_global(P.set[wr][0]);
  gfloat* acc_g = as_global(fwd ? P.illum : P.image);
  gfloat* snaps = as_global(P.snaps);
  const gfloat* data = as_global(P.data);
  const gfloat* src = as_global(P.src);
  const gint* roff = as_global(P.roff);
  const gint* ridx = as_global(P.ridx);

  const bool col_in = col < nx;
  const int colc = col_in ? col : nx - 1;  // clamped column of the loads
  const bool core_col = col_in && col >= cx0 && col < cx0 + S::W;
  const bool col_int = col >= 2 && col < nx - 2;
  const int ns = bc.n_sponge;
  int dcol = ns - col;
  dcol = col - (nx - 1 - ns) > dcol ? col - (nx - 1 - ns) : dcol;
  dcol = dcol > 0 ? dcol : 0;
  dcol = dcol < ns ? dcol : ns;  // columns past the grid: any entry

  // injection: row, this lane's samples for every sub-step, and the value
  // that generates iteration 0 in the first block of a phase
  const int src_row = src_cell / nx, src_col = src_cell - src_row * nx;
  const int inj_row = fwd ? src_row : riz;
  int r0 = 0, r1 = 0;
  if (!fwd && col_in) {
    r0 = roff[col];
    r1 = roff[col + 1];
  }
  const bool inj_lane = fwd ? (col_in && src_col == col) : r1 > r0;
  const bool dup = !fwd && r1 - r0 > 1;
  const int rsel = r1 - r0 == 1 ? ridx[r0] : 0;
  const float inj = bc.inj, w = bc.w;
  auto iter = [&](int m) { return fwd ? it0 + (m - 1) : it0 - (m - 1); };
  float sinj[T + 1];
  static_for<1, T + 1>([&](auto m_c) {
    constexpr int m = decltype(m_c)::value;
    const int it = iter(m);
    float v = 0.0f;
    if (m <= nsub) {
      if (fwd) {
        if (it + 1 < nt) v = src[it + 1] * inj;
      } else if (it >= 1 && !dup) {
        v = data[(long long)(it - 1) * nrec + rsel] * inj;
      }
    }
    sinj[m] = v;
  });
  float g0 = 0.0f;
  if (first && inj_lane) {
    if (fwd) {
      g0 = 0.0f + src[0] * inj;
    } else {
      const gfloat* drow = data + (long long)(nt - 1) * nrec;
      for (int r = r0; r < r1; ++r) g0 = g0 + drow[ridx[r]] * inj;
    }
  }
  // sub-steps that inject (the next iteration exists), imaging sub-steps and
  // their snapshot levels
  const bool inj_ok = inj_lane && !dup;
  unsigned injm = 0;
  static_for<1, T + 1>([&](auto m_c) {
    constexpr int m = decltype(m_c)::value;
    const int it = iter(m);
    if (m <= nsub && (fwd ? it + 1 < nt : it >= 1)) injm |= 1u << m;
  });
  unsigned img = 0;
  int lvl[T + 1];
  static_for<1, T + 1>([&](auto m_c) {
    constexpr int m = decltype(m_c)::value;
    const int it = iter(m);
    lvl[m] = 0;
    if (m <= nsub && it % k == 0) {
      img |= 1u << m;
      lvl[m] = it / k;
    }
  });

  if (lx == 0) dup_any = 0;
  for (int i = lx; i <= ns && i <= kTsMaxSponge; i += CX) tab[i] = bc.table[i];
  __syncthreads();
  if (dup && inj_lane) dup_any = 1;  // several receivers on one column of this strip
  const float sp_col = tab[dcol];    // entry 0 is expf(-0) = 1: no sponge

  // ---- chunk loads (clamped into the grid; masked when committed), two
  // chunks in flight in alternating register sets
  float pC[2][R], pB[2][R], pA[2][R], pF[2][R], pG[2][R];
  auto load_chunk = [&](int jl, auto e_c) {
    constexpr int e = decltype(e_c)::value;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int row = rbase + jl * R + q;
      row = row < nz ? row : nz - 1;
      const unsigned idx = (unsigned)(row * nx + colc);
      pF[e][q] = coef[idx];
      pC[e][q] = sC[idx];
      pB[e][q] = sB[idx];
      pA[e][q] = sA[idx];
      pG[e][q] = acc_g[idx];
    }
  };

  // Register windows: slot 3 holds a level's newest chunk, slots 2, 1, 0 the
  // three before it; a chunk enters by shifting the others down.  When stage
  // m runs on chunk jl, level m-1's newest is jl+1, level m-2's is jl+2 and
  // level m-3's is jl+3, so every slot a stage reads is fixed.
  float lw[T][4][R];  // levels 0..T-1
  float lb[4][R];     // level -1
  float la[2][R];     // level -2 (kept cells)
  float sv[T + 1][R]; // back-propagation: snapshot rows of stage m's next chunk
  auto shift_in = [&](float (&win)[4][R], const float (&v)[R]) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      win[0][q] = win[1][q];
      win[1][q] = win[2][q];
      win[2][q] = win[3][q];
      win[3][q] = v[q];
    }
  };

  auto ring_of = [&](int jl) { return jl & (NR - 1); };
  // chunk jl (loaded into register set e) enters levels 0, -1, -2 and the rings
  auto commit = [&](int jl, auto e_c) {
    constexpr int e = decltype(e_c)::value;
    const int ring = ring_of(jl);
    float c[R], bb[R], a[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int row = rbase + jl * R + q;
      const bool in = col_in && row < nz;
      const bool interior = col_int && row >= 2 && row < nz - 2;
      const bool kept = in && !interior && !(bc.fs && row < 2);
      const bool core = core_col && row >= cz0 && row < cz1;
      const float pc = pC[e][q], pb = pB[e][q], pa = pA[e][q], pf = pF[e][q], pg = pG[e][q];
      c[q] = first ? ((inj_lane && row == inj_row) ? g0 : 0.0f) : (in ? pc : 0.0f);
      bb[q] = (!first && in) ? pb : 0.0f;
      a[q] = (!first && kept) ? pa : 0.0f;
      cfr[(ring * R + q) * CX + lx] = in ? pf : 0.0f;
      acr[(ring * R + q) * CX + lx] = (!first && core) ? pg : 0.0f;
    }
    shift_in(lw[0], c);
    shift_in(lb, bb);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      la[0][q] = la[1][q];
      la[1][q] = a[q];
    }
  };
  // level 0 of chunk jl (window slot s) out to the LDS exchange
  auto put_level0 = [&](int jl, int s) {
#pragma unroll
    for (int q = 0; q < R; ++q) xs[0][jl & 1][q * LS + lx + 2] = s == 2 ? lw[0][2][q] : lw[0][3][q];
  };
  // snapshot rows of stage m's chunk jl, when sub-step m images on the
  // back-propagation (clamped rows; only core rows of the grid are used)
  auto prefetch_snap = [&](int m, int jl) {
    if (!fwd && (img >> m & 1u)) {
      const gfloat* snap = snaps + (long long)lvl[m] * ncell;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        int row = rbase + jl * R + q;
        row = row < 0 ? 0 : (row < nz ? row : nz - 1);
        sv[m][q] = __builtin_nontemporal_load(snap + (unsigned)(row * nx + colc));
      }
    }
  };
  // x-neighbours (xm2, xm1, xp1, xp2) and coefficient of stage m's chunk
  auto read_x = [&](int m, int jl, float (&xn)[R][4], float (&cv)[R]) {
    const float* xr = &xs[m - 1][jl & 1][2 + lx];
    const int ring = ring_of(jl);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      xn[q][0] = xr[q * LS - 2];
      xn[q][1] = xr[q * LS - 1];
      xn[q][2] = xr[q * LS + 1];
      xn[q][3] = xr[q * LS + 2];
      cv[q] = cfr[(ring * R + q) * CX + lx];
    }
  };

  // ---- prologue: chunks -2 and -1 are zero; chunks 0 and 1 in; level 0 of
  // chunk 0 in LDS; chunks 2 and 3 and stage 1's first snapshot rows on their way
#pragma unroll
  for (int l = 0; l < T; ++l)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (lx < 2) xs[l][p][q * LS + lx] = 0.0f;
        if (lx >= CX - 2) xs[l][p][q * LS + lx + 4] = 0.0f;
      }
#pragma unroll
  for (int l = 0; l < T; ++l)
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int q = 0; q < R; ++q) lw[l][s][q] = 0.0f;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int q = 0; q < R; ++q) lb[s][q] = 0.0f;
#pragma unroll
  for (int q = 0; q < R; ++q) la[0][q] = la[1][q] = 0.0f;
#pragma unroll
  for (int m = 0; m <= T; ++m)
#pragma unroll
    for (int q = 0; q < R; ++q) sv[m][q] = 0.0f;
  using E0 = std::integral_constant<int, 0>;
  using E1 = std::integral_constant<int, 1>;
  load_chunk(0, E0{});
  commit(0, E0{});
  load_chunk(1, E1{});
  commit(1, E1{});
  put_level0(0, 2);  // chunk 0 sits in slot 2 below chunk 1
  load_chunk(2, E0{});
  load_chunk(3, E1{});
  prefetch_snap(1, 0);
  __syncthreads();
  const bool dup_block = dup_any != 0;

  // ---- one pipeline step: stage m advances chunk u - m + 1 (outside the
  // grid during fill and drain)
  const int nsteps = ncl + nsub - 1;
  // the register set of a chunk's loads is its parity, so steps come in pairs
  auto step = [&](int u, auto e_c) {
    float xa[R][4], ca[R], xb[R][4], cb[R];
    float wn[R], wl[R], wp[R];  // the last sub-step's chunk: new level, level before, two before
    read_x(1, u, xa, ca);
    static_for<1, T + 1>([&](auto m_c) {
      constexpr int m = decltype(m_c)::value;
      // this stage's x-neighbours and coefficients; the next stage's are read now
      float(&xn)[R][4] = (m & 1) ? xa : xb;
      float(&cv)[R] = (m & 1) ? ca : cb;
      if (m < T && m + 1 <= nsub) read_x(m + 1, u - m, (m & 1) ? xb : xa, (m & 1) ? cb : ca);
      if (m <= nsub) {
        const int jl = u - m + 1;
        const int row0 = rbase + jl * R;
        const int ring = ring_of(jl);
        const float(&Lc)[R] = lw[m - 1][2];
        const float(&Lu)[R] = lw[m - 1][1];
        const float(&Ld)[R] = lw[m - 1][3];
        float pv[R], kv[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          pv[q] = m == 1 ? lb[2][q] : lw[m >= 2 ? m - 2 : 0][1][q];
          kv[q] = m == 1 ? la[0][q] : (m == 2 ? lb[1][q] : lw[m >= 3 ? m - 3 : 0][0][q]);
        }
        // side work on the post-injection level of this sub-step's iteration
        if (img >> m & 1u) {
          if (fwd) {
            gfloat* snap = snaps + (long long)lvl[m] * ncell;
#pragma unroll
            for (int q = 0; q < R; ++q) {
              const int row = row0 + q;
              if (core_col && row >= cz0 && row < cz1)
                __builtin_nontemporal_store(Lc[q], snap + (unsigned)(row * nx + col));
            }
          }
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const int row = row0 + q;
            const bool core = core_col && row >= cz0 && row < cz1;
            const float a = acr[(ring * R + q) * CX + lx];
            const float term = fwd ? Lc[q] * Lc[q] : (sv[m][q] * Lc[q]) * w;
            acr[(ring * R + q) * CX + lx] = core ? a + term : a;
          }
        }
        // sponge factor table[max(dcol, drow)] (wave2d.c:247-262); rows
        // above the bottom strip (and below the top one without a free
        // surface) take the column's factor
        float spq[R];
        if (row0 + R - 1 < nz - 1 - ns && (bc.fs || row0 > ns)) {
#pragma unroll
          for (int q = 0; q < R; ++q) spq[q] = sp_col;
        } else {
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const int row = row0 + q;
            int d = row - (nz - 1 - ns);
            if (!bc.fs && ns - row > d) d = ns - row;
            d = d > dcol ? d : dcol;
            d = d < ns ? d : ns;
            spq[q] = tab[d];
          }
        }
        // leapfrog update of the chunk
        float n[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int row = row0 + q;
          const float pc = Lc[q];
          const float xm2 = xn[q][0], xm1 = xn[q][1], xp1 = xn[q][2], xp2 = xn[q][3];
          const float zm2 = q >= 2 ? Lc[q >= 2 ? q - 2 : 0] : Lu[q + R - 2];
          const float zm1 = q >= 1 ? Lc[q >= 1 ? q - 1 : 0] : Lu[R - 1];
          const float zp1 = q + 1 < R ? Lc[q + 1 < R ? q + 1 : 0] : Ld[0];
          const float zp2 = q + 2 < R ? Lc[q + 2 < R ? q + 2 : 0] : Ld[q + 2 - R];
          float lx2 = TSW_C2 * (xm2 + xp2);
          lx2 = lx2 + TSW_C1 * (xm1 + xp1);
          lx2 = (lx2 + TSW_C0 * pc) * bc.rdx2;
          float lz2 = TSW_C2 * (zm2 + zp2);
          lz2 = lz2 + TSW_C1 * (zm1 + zp1);
          lz2 = (lz2 + TSW_C0 * pc) * bc.rdz2;
          float v = (2.0f * pc - pv[q]) + cv[q] * (lx2 + lz2);
          v = v * spq[q];
          // kept cells (and cells outside the grid) carry the level three back
          const bool interior = col_int && row >= 2 && row < nz - 2;
          n[q] = interior ? v : kv[q];
        }
        if (bc.fs && row0 == 0 && col_in) {  // rows 0, 1 and 2 are in the grid's chunk 0
          n[0] = 0.0f;
          n[1] = -n[2];
        }
        // injection of the next iteration's samples (one row of the chunk)
        if (inj_row >= row0 && inj_row < row0 + R && (injm >> m & 1u)) {
          const int qi = inj_row - row0;
          if (!fwd && dup_block) {
            // columns with several receivers add them in receiver order
            const gfloat* drow = data + (long long)(iter(m) - 1) * nrec;
            static_for<0, R>([&](auto q_c) {
              constexpr int q = decltype(q_c)::value;
              if (qi == q && dup)
                for (int r = r0; r < r1; ++r) n[q] = n[q] + drow[ridx[r]] * inj;
            });
          }
          static_for<0, R>([&](auto q_c) {
            constexpr int q = decltype(q_c)::value;
            if (qi == q) n[q] = inj_ok ? n[q] + sinj[m] : n[q];
          });
        }
        if (m == nsub) {  // kept for the write-back after the stages
#pragma unroll
          for (int q = 0; q < R; ++q) {
            wn[q] = n[q];
            wl[q] = Lc[q];
            wp[q] = pv[q];
          }
        }
        if (m < T) {
          shift_in(lw[m < T ? m : 0], n);
#pragma unroll
          for (int q = 0; q < R; ++q) xs[m < T ? m : 0][jl & 1][q * LS + lx + 2] = n[q];
        }
        // this stage's snapshot rows for its next chunk
        prefetch_snap(m, jl + 1);
      }
    });
    // the last sub-step's chunk: core write-back
    {
      const int jl = u - nsub + 1;
      const int row0 = rbase + jl * R;
      const int ring = ring_of(jl);
      const bool kept_rows = !col_int || row0 + R > nz - 2 || (!bc.fs && row0 < 2);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int row = row0 + q;
        float a = acr[(ring * R + q) * CX + lx];
        a = (fwd && last) ? a * w : a;
        if (core_col && row >= cz0 && row < cz1) {
          const unsigned idx = (unsigned)(row * nx + col);
          acc_g[idx] = a;
          if (!last) {
            wC[idx] = wn[q];
            wB[idx] = wl[q];
            const bool interior = col_int && row >= 2 && row < nz - 2;
            if (kept_rows && !interior && !(bc.fs && row < 2)) wA[idx] = wp[q];
          }
        }
      }
    }
    // end of step: chunk u+2 in, level 0 of chunk u+1 out to LDS, chunk u+4
    // on its way into the register set chunk u+2 just left
    commit(u + 2, e_c);
    put_level0(u + 1, 2);
    load_chunk(u + 4, e_c);
    __syncthreads();
  };
  for (int u = 0; u < nsteps; u += 2) {
    step(u, E0{});
    if (u + 1 < nsteps) step(u + 1, E1{});
  }
}

template <int T, int R, int WX, int H>
void launch_ts(const BatchConst& bc, const TbShot* shots, const TbEntry* ents, const int* pre,
               int nent, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL((k_ts<T, R, WX, H>), dim3(ntiles), dim3(64 * WX), 0, st, bc, shots, ents,
                     pre, nent, ntiles);
}

// The instantiated streaming configurations (T sub-steps, R rows per chunk,
// WX waves across, H rows per depth segment), in table order after k_tb's;
// the kernels are compiled in ts_a.hip / ts_b.hip.
#define TSW_TS_CONFIGS_A(X) X(4, 3, 2, 1 << 20) X(4, 3, 2, 150) X(4, 3, 2, 100) X(4, 4, 2, 152)
#define TSW_TS_CONFIGS_B(X) X(5, 3, 2, 150) X(6, 3, 2, 150) X(4, 3, 2, 75) X(5, 3, 2, 102)

#define TSW_TS_EXTERN(T, R, WX, H)                                                 \
  extern template void launch_ts<T, R, WX, H>(const BatchConst&, const TbShot*,     \
                                              const TbEntry*, const int*, int, int, \
                                              hipStream_t);
#define TSW_TS_INSTANTIATE(T, R, WX, H)                                                \
  template void launch_ts<T, R, WX, H>(const BatchConst&, const TbShot*, const TbEntry*, \
                                       const int*, int, int, hipStream_t);

}  // namespace tsw
