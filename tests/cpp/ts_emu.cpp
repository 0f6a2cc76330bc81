// Runs the streaming blocked kernel (k_ts, das_amd/csrc/ts.h), the tiled
// blocked kernel (k_tb, das_amd/csrc/tb.h) and, built with -DTC_T=.. -DTC_R=..,
// the blocked kernel with its levels in LDS (k_tc, das_amd/csrc/tc.h) on the
// host, through the CPU
// stand-in for the HIP runtime in tests/cpp/hip_emu, over every block of both
// phases of each shot, and compares illumination, image and every snapshot
// bit for bit.  k_tb is bit-identical to the per-step kernel and the oracle on
// the GPU; this pins k_ts (and k_tc) to it on the CPU for shapes that cover one and
// several depth segments and strips, edge and duplicate receivers, sources
// and receivers on rows 0..5, the free surface on and off, and blocks of fewer
// than T sub-steps.  Inputs are seeded random values (models, wavelet, data,
// sponge table): both kernels must agree whatever the values.
//
//   g++ -O1 -std=c++17 -I tests/cpp/hip_emu -I include -I das_amd/csrc \
//       -DTS_T=5 -DTS_WX=2 -DTS_H=100 tests/cpp/ts_emu.cpp -o ts_emu && ./ts_emu [nt_scale]
#include <hip/hip_runtime.h>
#include <cstdint>
#include <random>
#include "tc.h"
#include "ts.h"

using namespace tsw;

struct ShotSpec { int nz, nx, nt, six, siz, riz; };

struct Arrays {
  std::vector<float> set[2][3], snaps, illum, image;
};

static uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

template <typename Launch>
static void run_shot(const ShotSpec& sp, int fs, int T, int W, int H, Launch launch, Arrays& A,
                     std::vector<float>& coef_f, std::vector<float>& coef_b, std::vector<float>& src,
                     std::vector<float>& data, std::vector<int>& roff, std::vector<int>& ridx,
                     std::vector<float>& table, int ns, int nrec) {
  const int nx = sp.nx, nz = sp.nz, nt = sp.nt, k = 2;
  const long long nc = (long long)nx * nz;
  const float qnan = std::nanf("");
  for (int r = 0; r < 2; ++r) for (int j = 0; j < 3; ++j) A.set[r][j].assign(nc, qnan);
  A.snaps.assign(((nt - 1) / k + 1) * nc, qnan);
  A.illum.assign(nc, qnan);
  A.image.assign(nc, qnan);
  TbShot P{};
  P.nx = nx; P.nz = nz; P.nt = nt; P.k = k;
  P.src_cell = sp.siz * nx + sp.six; P.riz = sp.riz; P.nrec = nrec;
  for (int r = 0; r < 2; ++r) for (int j = 0; j < 3; ++j) P.set[r][j] = A.set[r][j].data();
  P.coef_f = coef_f.data(); P.coef_b = coef_b.data(); P.src = src.data(); P.data = data.data();
  P.roff = roff.data(); P.ridx = ridx.data(); P.snaps = A.snaps.data();
  P.illum = A.illum.data(); P.image = A.image.data();
  BatchConst bc{};
  bc.fs = fs; bc.n_sponge = ns; bc.rdx2 = 1.0f / 25.0f; bc.rdz2 = 1.0f / 25.0f;
  bc.inj = 0.25e-6f; bc.w = 1e-3f; bc.table = table.data();
  const int tiles_x = (nx + W - 1) / W, tiles = tiles_x * ((nz + H - 1) / H);
  const int nb = (nt + T - 1) / T;
  std::vector<TbTile> tab(tiles);
  for (int b = 0; b < tiles; ++b) {
    const int t = xcd_tile_tb(b, tiles);
    tab[b] = {0, (t % tiles_x) | (t / tiles_x) << 16};
  }
  for (int ph = 0; ph < 2; ++ph)
    for (int b = 0; b < nb; ++b) {
      TbEntry e{};
      e.shot = 0; e.phase = ph;
      e.it0 = ph == 0 ? b * T : nt - 1 - b * T;
      e.nsub = ph == 0 ? std::min(T, nt - e.it0) : std::min(T, e.it0 + 1);
      e.flags = (b == 0 ? kTbFirst : 0) | (b == nb - 1 ? kTbLast : 0);
      e.rd = b & 1; e.tiles_x = tiles_x;
      tb_schedule(e, nt, k);
      launch(bc, &P, &e, tab.data(), tiles);
    }
}

int main(int argc, char** argv) {
  // nz, nx, nt, source column and row, receiver row; nt scaled by argv[1]
  std::vector<ShotSpec> specs = {{70, 300, 97, 150, 2, 2}, {40, 48, 61, 10, 0, 1}, {130, 190, 140, 5, 3, 0},
                                 {55, 131, 77, 130, 1, 3}, {64, 256, 120, 128, 2, 2}, {230, 140, 33, 70, 5, 1}};
  const double scale = argc > 1 ? std::atof(argv[1]) : 1.0;
  for (auto& sp : specs) sp.nt = std::max(3, (int)(sp.nt * scale));
  int fails = 0;
  std::mt19937 rng(12345);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  for (size_t si = 0; si < specs.size(); ++si)
    for (int fs = 0; fs < 2; ++fs) {
      const ShotSpec& sp = specs[si];
      const int nx = sp.nx, nz = sp.nz, nt = sp.nt;
      const long long nc = (long long)nx * nz;
      const int ns = 8;
      std::vector<float> table(ns + 1);
      table[0] = 1.0f;
      for (int i = 1; i <= ns; ++i) table[i] = 0.6f + 0.4f * U(rng);
      std::vector<float> coef_f(nc), coef_b(nc), src(nt);
      for (auto& v : coef_f) v = 0.5f + U(rng);
      for (auto& v : coef_b) v = 0.5f + U(rng);
      for (auto& v : src) v = U(rng) - 0.5f;
      std::vector<int> rix = {0, 1, nx - 1, nx - 2, 3, 3};
      for (int x = 2; x < nx - 2; x += 3) rix.push_back(x);
      const int nrec = (int)rix.size();
      std::vector<float> data((size_t)nt * nrec);
      for (auto& v : data) v = U(rng) - 0.5f;
      std::vector<int> roff(nx + 1, 0), ridx;
      for (int x = 0; x < nx; ++x) {
        for (int r = 0; r < nrec; ++r) if (rix[r] == x) ridx.push_back(r);
        roff[x + 1] = (int)ridx.size();
      }
      Arrays A, B;
      run_shot(sp, fs, 5, TbShape<5, 4, 1, 16>::W, TbShape<5, 4, 1, 16>::H,
               [](const BatchConst& bc, const TbShot* s, const TbEntry* e, const TbTile* tb, int nti) {
                 hipLaunchKernelGGL((k_tb<5, 4, 1, 16>), dim3(nti), dim3(64, 16), 0, 0, bc, s, e, tb);
               }, A, coef_f, coef_b, src, data, roff, ridx, table, ns, nrec);
      run_shot(sp, fs, TS_T, TsShape<TS_T, TS_WX>::W, TS_H,
               [](const BatchConst& bc, const TbShot* s, const TbEntry* e, const TbTile* tb, int nti) {
                 launch_ts<TS_T, TS_WX, TS_H>(bc, s, e, tb, nti, nullptr);
               }, B, coef_f, coef_b, src, data, roff, ridx, table, ns, nrec);
#ifdef TC_T
      Arrays C;
      run_shot(sp, fs, TC_T, TcShape<TC_T, TC_R>::W, TcShape<TC_T, TC_R>::H,
               [](const BatchConst& bc, const TbShot* s, const TbEntry* e, const TbTile* tb, int nti) {
                 launch_tc<TC_T, TC_R>(bc, s, e, tb, nti, nullptr);
               }, C, coef_f, coef_b, src, data, roff, ridx, table, ns, nrec);
#endif
      auto cmp = [&](const char* what, const std::vector<float>& a, const std::vector<float>& b) {
        long long bad = 0, first = -1;
        for (size_t i = 0; i < a.size(); ++i)
          if (bits(a[i]) != bits(b[i])) { if (first < 0) first = (long long)i; ++bad; }
        if (bad) {
          ++fails;
          std::printf("  %s: %lld differ, first at %lld (row %lld col %lld): %g vs %g\n", what, bad, first,
                      (first % nc) / nx, first % nx, a[first], b[first]);
        }
        return bad;
      };
      std::printf("shot %zu (nz %d nx %d nt %d src %d,%d riz %d) fs=%d\n", si, nz, nx, nt, sp.siz, sp.six, sp.riz, fs);
      cmp("illum", A.illum, B.illum);
      cmp("image", A.image, B.image);
      cmp("snaps", A.snaps, B.snaps);
#ifdef TC_T
      cmp("illum (k_tc)", A.illum, C.illum);
      cmp("image (k_tc)", A.image, C.image);
      cmp("snaps (k_tc)", A.snaps, C.snaps);
#endif
      std::fflush(stdout);
    }
  std::printf("%s (%lld barriers)\n", fails ? "FAIL" : "OK", emu::barriers());
  return fails ? 1 : 0;
}
