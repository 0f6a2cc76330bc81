// Output format: a 2-D matrix as CSV, byte-compatible with save_matrix_csv
// (tsw_rtm/visualizer.py:19-39).  Row 0 holds "row_name/col_name" and the
// column axis (%.6g); every further row holds its row-axis value (%.6g) and
// the matrix values (%.6e), each value formatted from its double value as
// Python's format() does (both are correctly rounded).  Host code only.
#include <cstdio>
#include <string>
#include <vector>

#include "common.h"
#include "tswrtm.h"

namespace tsw {

static void put(std::string& out, const char* fmt, double v) {
  // Python formats every NaN as "nan", whatever its sign bit; printf would
  // write "-nan" for a NaN whose sign bit is set
  if (v != v) {
    out += "nan";
    return;
  }
  char buf[64];
  const int n = std::snprintf(buf, sizeof buf, fmt, v);
  out.append(buf, (size_t)n);
}

}  // namespace tsw

using namespace tsw;

extern "C" {

int tsw_save_matrix_csv(const char* path, const void* matrix, int precision, int nr, int nc,
                        const double* row_axis, const double* col_axis, const char* row_name,
                        const char* col_name) {
  TSW_ARG(path && matrix && row_axis && col_axis && row_name && col_name, "null argument");
  TSW_ARG(nr >= 0 && nc >= 0, "bad shape");
  TSW_ARG(precision == 4 || precision == 8, "precision must be 4 or 8");
  std::FILE* f = std::fopen(path, "wb");
  TSW_ARG(f, "cannot open %s for writing", path);
  std::string line;
  line.reserve((size_t)nc * 16 + 64);
  line += row_name;
  line += '/';
  line += col_name;
  line += ',';
  for (int j = 0; j < nc; ++j) {
    if (j) line += ',';
    put(line, "%.6g", col_axis[j]);
  }
  line += '\n';
  bool ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
  for (int i = 0; i < nr && ok; ++i) {
    line.clear();
    put(line, "%.6g", row_axis[i]);
    line += ',';
    for (int j = 0; j < nc; ++j) {
      if (j) line += ',';
      const size_t e = (size_t)i * nc + j;
      const double v = precision == 4 ? (double)((const float*)matrix)[e]
                                      : ((const double*)matrix)[e];
      put(line, "%.6e", v);
    }
    line += '\n';
    ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
  }
  ok = (std::fclose(f) == 0) && ok;
  TSW_ARG(ok, "writing %s failed", path);
  return kOk;
}

}  // extern "C"
