// Reader of the cooling-function table (ln Lambda over ln D x ln T), the text
// format that files_sys::txt::ReadTableFunc reads
// (lib/files_sys/src/reader_txt.cpp:59-118).
//
// The file is read whole and parsed with strtod, which rounds every decimal
// to the nearest double as the reference's istream extraction does, so the
// values are bitwise the reference's.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>
#include <vector>

#include "mesh_common.h"

namespace {
using rhllc_mesh::fail;

// A count line "<text> = N": the integer after the first '='.
bool count_line(const std::string &line, const char *what, int32_t &n) {
  const size_t eq = line.find('=');
  if (eq == std::string::npos) return fail(std::string("table: no '=' in the ") + what + " line");
  const char *s = line.c_str() + eq + 1;
  char *end = nullptr;
  errno = 0;
  const long v = std::strtol(s, &end, 10);
  if (end == s || errno || v < 2 || v > (1L << 24))
    return fail(std::string("table: bad ") + what + " count in '" + line + "'");
  n = (int32_t)v;
  return true;
}

bool next_line(const std::string &text, size_t &pos, std::string &line) {
  if (pos >= text.size()) return false;
  size_t e = text.find('\n', pos);
  if (e == std::string::npos) e = text.size();
  line.assign(text, pos, e - pos);
  pos = e + 1;
  return true;
}
}  // namespace

extern "C" int rhllc_table_read(const char *path, int32_t *nx, int32_t *ny, double header[6], double *data) {
  if (!path || !nx || !ny || !header) return fail("null argument"), 1;
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(std::string("cannot open ") + path), 1;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();

  size_t pos = 0;
  std::string line;
  int32_t nd = 0, nt = 0;
  if (!next_line(text, pos, line) || !count_line(line, "density", nd)) return 1;
  if (!next_line(text, pos, line) || !count_line(line, "temperature", nt)) return 1;
  if (!next_line(text, pos, line) || !next_line(text, pos, line)) return fail("table: truncated header"), 1;

  // Nd rows of Nt triples; the axes are the first ln D of each row and the
  // ln T column of the last row (the reference overwrites it row by row).
  std::vector<double> dens(nd), temp(nt);
  const char *p = text.c_str() + std::min(pos, text.size());
  for (int32_t i = 0; i < nd; ++i)
    for (int32_t j = 0; j < nt; ++j) {
      double v[3];
      for (int k = 0; k < 3; ++k) {
        char *end = nullptr;
        v[k] = std::strtod(p, &end);
        if (end == p) return fail("table: fewer than Nd*Nt triples"), 1;
        p = end;
      }
      if (j == 0) dens[i] = v[0];
      temp[j] = v[1];
      if (data) data[(int64_t)i * nt + j] = v[2];
    }

  double mnx = std::numeric_limits<double>::max(), mxx = -std::numeric_limits<double>::max();
  double mny = mnx, mxy = mxx;
  for (double x : dens) mxx = std::max(mxx, x), mnx = std::min(mnx, x);
  for (double y : temp) mxy = std::max(mxy, y), mny = std::min(mny, y);
  *nx = nd, *ny = nt;
  header[0] = mnx, header[1] = mxx, header[2] = dens[1] - dens[0];
  header[3] = mny, header[4] = mxy, header[5] = temp[1] - temp[0];
  return 0;
}
