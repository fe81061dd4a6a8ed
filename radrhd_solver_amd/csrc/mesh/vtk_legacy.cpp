// Legacy-format VTK reader for unstructured tetrahedral grids, without a VTK
// library.
//
// Replaces files_sys::vtk::Read (lib/files_sys/include/reader_vtk.h:43-61),
// which hands the file to VTK 9.3's vtkGenericDataObjectReader
// (cmake_modules.cmake:1) and keeps the vtkUnstructuredGrid it returns.  The
// solver's preprocessing uses only the point coordinates and the cell point
// ids, so that is what this reader keeps.
//
// Accepted input:
//   "# vtk DataFile Version 2.0" ... "5.1", a title line, ASCII or BINARY
//   (binary data is big-endian, as the legacy format defines it), and
//   DATASET UNSTRUCTURED_GRID with
//     POINTS n <type>
//     CELLS n size                     (before 5.0: "4 a b c d" per cell)
//     CELLS n_off n_conn, OFFSETS <type> ..., CONNECTIVITY <type> ...  (5.x)
//     CELL_TYPES n                     (every type 10 = VTK_TETRA)
// FIELD and METADATA blocks are skipped; reading ends at CELL_DATA or
// POINT_DATA.  Points declared "float" are rounded to float and widened, as
// VTK keeps them in a vtkFloatArray.
#include "mesh_common.h"

#include <charconv>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

using namespace rhllc_mesh;

struct rhllc_vtk_mesh {
  std::vector<double> points;  // 3 per point
  std::vector<int32_t> tets;   // 4 per cell
};

namespace {

constexpr int kVtkTetra = 10;

class Parser {
 public:
  Parser(std::string &&buf) : b_(std::move(buf)) {}

  // The next non-empty line, without its line end; false at end of file.
  bool line(std::string &out) {
    while (p_ < b_.size()) {
      size_t e = b_.find('\n', p_);
      if (e == std::string::npos) e = b_.size();
      size_t s = p_, t = e;
      p_ = e + 1;
      while (s < t && (b_[s] == ' ' || b_[s] == '\t' || b_[s] == '\r')) ++s;
      while (t > s && (b_[t - 1] == ' ' || b_[t - 1] == '\t' || b_[t - 1] == '\r')) --t;
      if (t > s) {
        out.assign(b_, s, t - s);
        return true;
      }
    }
    return false;
  }

  // One line exactly as it is (used for the title, which may be empty).
  bool raw_line(std::string &out) {
    if (p_ >= b_.size()) return false;
    size_t e = b_.find('\n', p_);
    if (e == std::string::npos) e = b_.size();
    out.assign(b_, p_, e - p_);
    if (!out.empty() && out.back() == '\r') out.pop_back();
    p_ = e + 1;
    return true;
  }

  // ASCII values of the current data block.
  // Bytes left after the read position.
  size_t remaining() const { return p_ < b_.size() ? b_.size() - p_ : 0; }

  // Whether n values can follow: w bytes each (BINARY), or at least one
  // character and a separator each (ASCII).  Checked before a block is sized
  // from its declared count, so a corrupt count is an error, not an allocation.
  bool can_hold(int64_t n, int w) const {
    if (n < 0) return false;
    const size_t r = remaining();
    return w > 0 ? (uint64_t)n <= r / (size_t)w : (uint64_t)n <= r / 2 + 1;
  }

  template <typename T>
  bool ascii(T &v) {
    skip_space();
    const char *s = b_.data() + p_, *e = b_.data() + b_.size();
    auto r = std::from_chars(s, e, v);
    if (r.ec != std::errc()) return false;
    p_ = r.ptr - b_.data();
    return true;
  }

  // Raw bytes of a BINARY data block; a block is followed by a line end.
  const uint8_t *bytes(size_t n) {
    if (n > remaining()) return nullptr;
    const uint8_t *r = reinterpret_cast<const uint8_t *>(b_.data() + p_);
    p_ += n;
    return r;
  }

 private:
  void skip_space() {
    while (p_ < b_.size() && (b_[p_] == ' ' || b_[p_] == '\t' || b_[p_] == '\n' || b_[p_] == '\r')) ++p_;
  }
  std::string b_;
  size_t p_ = 0;
};

std::vector<std::string> words(const std::string &s) {
  std::vector<std::string> w;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
    size_t j = i;
    while (j < s.size() && s[j] != ' ' && s[j] != '\t') ++j;
    if (j > i) w.emplace_back(s, i, j - i);
    i = j;
  }
  return w;
}

bool to_i64(const std::string &s, int64_t &v) {
  auto r = std::from_chars(s.data(), s.data() + s.size(), v);
  return r.ec == std::errc() && r.ptr == s.data() + s.size();
}

// Size in bytes of one BINARY value of a legacy data type, 0 if unsupported.
int binary_size(const std::string &t) {
  if (t == "double" || t == "vtktypeint64" || t == "vtktypeuint64") return 8;
  if (t == "float" || t == "int" || t == "unsigned_int" || t == "vtktypeint32" || t == "vtktypeuint32") return 4;
  return 0;
}
bool is_float_type(const std::string &t) { return t == "float" || t == "double"; }

uint64_t load_be(const uint8_t *p, int n) {
  uint64_t v = 0;
  for (int k = 0; k < n; ++k) v = (v << 8) | p[k];
  return v;
}

// n values of type `type` as doubles (points) -- ASCII or BINARY.
bool read_reals(Parser &P, bool binary, const std::string &type, int64_t n, std::vector<double> &out) {
  if (!P.can_hold(n, binary ? binary_size(type) : 0))
    return fail("POINTS declares " + std::to_string(n) + " values, more than the rest of the file holds");
  out.resize((size_t)n);
  const bool f32 = type == "float";
  if (binary) {
    const int w = binary_size(type);
    if (!is_float_type(type) || w == 0) return fail("unsupported BINARY point type " + type);
    const uint8_t *p = P.bytes((size_t)n * w);
    if (!p) return fail("BINARY POINTS block is truncated");
    for (int64_t i = 0; i < n; ++i) {
      const uint64_t bits = load_be(p + i * w, w);
      if (f32) {
        float f;
        const uint32_t b32 = (uint32_t)bits;
        std::memcpy(&f, &b32, 4);
        out[i] = (double)f;
      } else {
        std::memcpy(&out[i], &bits, 8);
      }
    }
    return true;
  }
  for (int64_t i = 0; i < n; ++i) {
    double v;
    if (!P.ascii(v)) return fail("bad or missing POINTS value");
    out[i] = f32 ? (double)(float)v : v;
  }
  return true;
}

// n integers of type `type` -- ASCII or BINARY.
bool read_ints(Parser &P, bool binary, const std::string &type, int64_t n, std::vector<int64_t> &out,
               const char *what) {
  if (!P.can_hold(n, binary ? binary_size(type) : 0))
    return fail(std::string(what) + " declares " + std::to_string(n) + " values, more than the rest of the file holds");
  out.resize((size_t)n);
  if (binary) {
    const int w = binary_size(type);
    if (w == 0 || is_float_type(type)) return fail(std::string("unsupported BINARY type ") + type + " in " + what);
    const uint8_t *p = P.bytes((size_t)n * w);
    if (!p) return fail(std::string("BINARY ") + what + " block is truncated");
    for (int64_t i = 0; i < n; ++i) {
      const uint64_t v = load_be(p + i * w, w);
      out[i] = w == 4 ? (int64_t)(int32_t)(uint32_t)v : (int64_t)v;
    }
    return true;
  }
  for (int64_t i = 0; i < n; ++i)
    if (!P.ascii(out[i])) return fail(std::string("bad or missing ") + what + " value");
  return true;
}

bool parse(const char *path, rhllc_vtk_mesh &m) {
  std::FILE *f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path);
  std::string buf;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    return fail(std::string("cannot size ") + path);
  }
  buf.resize((size_t)sz);
  const size_t got = sz ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  if (got != (size_t)sz) return fail(std::string("short read of ") + path);

  Parser P(std::move(buf));
  std::string ln;
  if (!P.raw_line(ln) || ln.rfind("# vtk DataFile Version", 0) != 0)
    return fail("not a legacy VTK file (first line is not '# vtk DataFile Version x.y')");
  const double version = std::atof(ln.c_str() + std::strlen("# vtk DataFile Version"));
  if (!P.raw_line(ln)) return fail("missing title line");
  if (!P.line(ln)) return fail("missing ASCII/BINARY line");
  bool binary;
  if (ln == "ASCII")
    binary = false;
  else if (ln == "BINARY")
    binary = true;
  else
    return fail("third line must be ASCII or BINARY, got '" + ln + "'");
  if (!P.line(ln)) return fail("missing DATASET line");
  {
    const auto w = words(ln);
    if (w.size() != 2 || w[0] != "DATASET") return fail("expected 'DATASET <type>', got '" + ln + "'");
    if (w[1] != "UNSTRUCTURED_GRID") return fail("DATASET " + w[1] + " is not UNSTRUCTURED_GRID");
  }

  std::vector<double> pts;
  std::vector<int64_t> conn, offsets, types;
  int64_t n_points = -1, n_cells = -1;
  bool have_cells = false, have_types = false;
  std::vector<int64_t> tmp;

  while (P.line(ln)) {
    const auto w = words(ln);
    const std::string &k = w[0];
    if (k == "POINTS") {
      if (w.size() != 3 || !to_i64(w[1], n_points) || n_points < 0 || n_points > INT64_MAX / 3)
        return fail("bad POINTS line '" + ln + "'");
      if (!read_reals(P, binary, w[2], 3 * n_points, pts)) return false;
    } else if (k == "CELLS") {
      int64_t a, b;
      if (w.size() != 3 || !to_i64(w[1], a) || !to_i64(w[2], b) || a < 0 || b < 0)
        return fail("bad CELLS line '" + ln + "'");
      if (version < 5.0) {
        // n cells, then `size` integers: a count before each cell's ids.
        if (b < a) return fail("CELLS size is smaller than its cell count");
        if (!read_ints(P, binary, "int", b, tmp, "CELLS")) return false;
        n_cells = a;
        offsets.assign((size_t)n_cells + 1, 0);
        conn.clear();
        conn.reserve((size_t)(b - a));
        int64_t at = 0;
        for (int64_t c = 0; c < n_cells; ++c) {
          if (at >= b) return fail("CELLS list ends early");
          const int64_t npts = tmp[at++];
          if (npts < 0 || at + npts > b) return fail("CELLS list ends early");
          for (int64_t q = 0; q < npts; ++q) conn.push_back(tmp[at++]);
          offsets[c + 1] = (int64_t)conn.size();
        }
        if (at != b) return fail("CELLS size does not match its list");
      } else {
        // n_off offsets (n_cells + 1) and n_conn point ids, in two blocks.
        n_cells = a - 1;
        if (n_cells < 0) return fail("CELLS needs at least one offset");
        for (int blk = 0; blk < 2; ++blk) {
          if (!P.line(ln)) return fail("missing OFFSETS/CONNECTIVITY block");
          const auto v = words(ln);
          if (v.size() != 2) return fail("bad line '" + ln + "'");
          if (v[0] == "OFFSETS") {
            if (!read_ints(P, binary, v[1], a, offsets, "OFFSETS")) return false;
          } else if (v[0] == "CONNECTIVITY") {
            if (!read_ints(P, binary, v[1], b, conn, "CONNECTIVITY")) return false;
          } else {
            return fail("expected OFFSETS or CONNECTIVITY, got '" + ln + "'");
          }
        }
        if ((int64_t)offsets.size() != a || (int64_t)conn.size() != b) return fail("CELLS blocks are missing");
        if (offsets[0] != 0 || offsets[n_cells] != b) return fail("OFFSETS do not span CONNECTIVITY");
        for (int64_t c = 0; c < n_cells; ++c)
          if (offsets[c + 1] < offsets[c]) return fail("OFFSETS are not increasing");
      }
      have_cells = true;
    } else if (k == "CELL_TYPES") {
      int64_t n;
      if (w.size() != 2 || !to_i64(w[1], n) || n < 0) return fail("bad CELL_TYPES line '" + ln + "'");
      if (!read_ints(P, binary, "int", n, types, "CELL_TYPES")) return false;
      have_types = true;
    } else if (k == "FIELD") {
      // FIELD name n_arrays; each array: name n_comp n_tuples type, then values.
      int64_t na;
      if (w.size() != 3 || !to_i64(w[2], na) || na < 0) return fail("bad FIELD line '" + ln + "'");
      for (int64_t a = 0; a < na; ++a) {
        if (!P.line(ln)) return fail("FIELD block ends early");
        const auto v = words(ln);
        int64_t nc, nt;
        if (v.size() != 4 || !to_i64(v[1], nc) || !to_i64(v[2], nt) || nc < 0 || nt < 0)
          return fail("bad FIELD array line '" + ln + "'");
        if (nt > 0 && nc > (int64_t)P.remaining() / nt) return fail("FIELD block is truncated");
        if (binary) {
          const int bw = binary_size(v[3]);
          if (bw == 0) return fail("unsupported BINARY FIELD type " + v[3]);
          if (!P.bytes((size_t)(nc * nt) * bw)) return fail("BINARY FIELD block is truncated");
        } else {
          for (int64_t q = 0; q < nc * nt; ++q) {
            double d;
            if (!P.ascii(d)) return fail("FIELD values end early");
          }
        }
      }
    } else if (k == "METADATA") {
      // Information keys up to the next empty line: the line reader skips
      // empty lines, so read raw lines until one is empty.
      std::string r;
      while (P.raw_line(r)) {
        size_t q = 0;
        while (q < r.size() && (r[q] == ' ' || r[q] == '\t' || r[q] == '\r')) ++q;
        if (q == r.size()) break;
      }
    } else if (k == "CELL_DATA" || k == "POINT_DATA") {
      break;
    } else {
      return fail("unsupported section '" + k + "' in an UNSTRUCTURED_GRID");
    }
  }

  if (n_points < 0) return fail("no POINTS section");
  if (!have_cells) return fail("no CELLS section");
  if (!have_types) return fail("no CELL_TYPES section");
  if ((int64_t)types.size() != n_cells) return fail("CELL_TYPES count differs from CELLS count");
  if (n_points > INT32_MAX) return fail("more than 2^31-1 points");
  // The geometry records hold 25-bit cell ids (lib/grids/include/solvers_struct.h:87-102).
  if (n_cells >= (int64_t)1 << 25) return fail("more than 2^25-1 cells do not fit the geo_face_t id fields");

  m.points = std::move(pts);
  m.tets.resize((size_t)(4 * n_cells));
  for (int64_t c = 0; c < n_cells; ++c) {
    if (types[c] != kVtkTetra)
      return fail("cell " + std::to_string(c) + " has VTK type " + std::to_string(types[c]) +
                  "; only tetrahedra (10) are supported");
    if (offsets[c + 1] - offsets[c] != 4)
      return fail("cell " + std::to_string(c) + " does not have 4 points");
    for (int q = 0; q < 4; ++q) {
      const int64_t id = conn[offsets[c] + q];
      if (id < 0 || id >= n_points)
        return fail("cell " + std::to_string(c) + " refers to point " + std::to_string(id) + " out of range");
      m.tets[4 * c + q] = (int32_t)id;
    }
  }
  return true;
}

}  // namespace

extern "C" int rhllc_vtk_open(const char *path, rhllc_vtk_mesh **out, int64_t *n_points, int64_t *n_cells) {
  if (!path || !out) return fail("null argument"), 1;
  *out = nullptr;
  std::unique_ptr<rhllc_vtk_mesh> m;
  try {
    m.reset(new rhllc_vtk_mesh);
    if (!parse(path, *m)) return 1;
  } catch (const std::bad_alloc &) {
    return fail(std::string("out of host memory reading ") + path), 1;
  } catch (const std::exception &e) {
    return fail(std::string("reading ") + path + ": " + e.what()), 1;
  }
  if (n_points) *n_points = (int64_t)(m->points.size() / 3);
  if (n_cells) *n_cells = (int64_t)(m->tets.size() / 4);
  *out = m.release();
  return 0;
}

extern "C" int rhllc_vtk_copy(const rhllc_vtk_mesh *m, double *points, int32_t *tets) {
  if (!m) return fail("null mesh"), 1;
  if (points) std::memcpy(points, m->points.data(), m->points.size() * sizeof(double));
  if (tets) std::memcpy(tets, m->tets.data(), m->tets.size() * sizeof(int32_t));
  return 0;
}

extern "C" void rhllc_vtk_close(rhllc_vtk_mesh *m) { delete m; }
