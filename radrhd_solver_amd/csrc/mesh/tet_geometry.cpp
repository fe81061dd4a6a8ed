// Geometry of a tetrahedral mesh as the reference's preprocessing writes it,
// and the conversion of those arrays into geo_face_t / geo_cell_t records.
//
// rhllc_tet_geometry produces the five arrays of BuildDataFromVTK
// (lib/geometry/src/make_internal_format.cpp:272-336):
//   neighbors.bin  GetNeighborFace3D   (convert_vtk_geo.cpp:122-197), boundary
//                  codes then rewritten by SetTypeOfBound (make_internal_format.cpp:12-108)
//   normals.bin, areas.bin  NormalAndSquareFace3D (convert_vtk_geo.cpp:199-253)
//   volume.bin     GetVolume3D         (convert_vtk_geo.cpp:270-307)
//   centers.bin    CenterOfTetra3D     (convert_vtk_geo.cpp:309-347)
// rhllc_bin_to_geo is ConvertBinToGeo (lib/geometry/src/convert_bin_to_geo.cpp:3-48).
//
// Local face j of a cell is VTK 9.3's vtkTetra face j (the reference calls
// vtkCell::GetFace(j)): the point triples (0,1,3), (1,2,3), (2,0,3), (0,2,1)
// of the cell's own point list, opposite points 2, 0, 1, 3.  The triple order
// fixes the rounding of the face normal and area, so it is kept exactly.
//
// Arithmetic follows the reference operation for operation (3-vector
// cross products by components, sums left to right, division by the norm) and
// this file is compiled with -ffp-contract=off, so the arrays are bitwise the
// reference's.  Neighbour matching is a sort of the 4N face keys instead of
// VTK's point-to-cell links; it finds the same pairs.
#include "mesh_common.h"

#include <algorithm>
#include <cmath>
#include <vector>

#ifdef _OPENMP
#include <parallel/algorithm>
#endif

using namespace rhllc_mesh;

namespace {

// vtkTetra face table (VTK 9.3 Common/DataModel/vtkTetra.cxx) and the point
// of the cell that is not on each face.
constexpr int kFace[4][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {0, 2, 1}};
constexpr int kOpposite[4] = {2, 0, 1, 3};

struct V3 {
  double x, y, z;
};
inline V3 pt(const double *points, int32_t id) { return {points[3 * (int64_t)id], points[3 * (int64_t)id + 1], points[3 * (int64_t)id + 2]}; }
inline V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(const V3 &a, const V3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(const V3 &a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Half of |(q1 - q0) x (q2 - q0)|: the area of triangle q0 q1 q2.
inline double tri_area(const V3 &q0, const V3 &q1, const V3 &q2) { return 0.5 * norm(cross(sub(q1, q0), sub(q2, q0))); }

// Sort record of one half-face: its point ids in increasing order and h = 4*cell + slot.
struct Key {
  uint32_t a, b, c;
  int32_t pad;
  int64_t h;
};
inline bool key_less(const Key &l, const Key &r) {
  if (l.a != r.a) return l.a < r.a;
  if (l.b != r.b) return l.b < r.b;
  if (l.c != r.c) return l.c < r.c;
  return l.h < r.h;
}
inline bool same_face(const Key &l, const Key &r) { return l.a == r.a && l.b == r.b && l.c == r.c; }

bool neighbours(int64_t n_cells, const int32_t *tets, int32_t *nb) {
  const int64_t H = 4 * n_cells;
  std::vector<Key> keys((size_t)H);
#pragma omp parallel for schedule(static)
  for (int64_t h = 0; h < H; ++h) {
    const int32_t *t = tets + 4 * (h >> 2);
    const int *fj = kFace[h & 3];
    uint32_t v[3] = {(uint32_t)t[fj[0]], (uint32_t)t[fj[1]], (uint32_t)t[fj[2]]};
    if (v[0] > v[1]) std::swap(v[0], v[1]);
    if (v[1] > v[2]) std::swap(v[1], v[2]);
    if (v[0] > v[1]) std::swap(v[0], v[1]);
    keys[h] = Key{v[0], v[1], v[2], 0, h};
  }
#ifdef _OPENMP
  __gnu_parallel::sort(keys.begin(), keys.end(), key_less);
#else
  std::sort(keys.begin(), keys.end(), key_less);
#endif
  // Runs of equal keys: 1 = boundary, 2 = a shared face, more = a face shared
  // by three or more cells, which the reference rejects ("More than 1
  // neighbor").
  bool ok = true;
  int64_t bad = -1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int64_t i = 0; i < H; ++i) {
    if (i > 0 && same_face(keys[i - 1], keys[i])) continue;  // not the start of a run
    int64_t e = i + 1;
    while (e < H && same_face(keys[i], keys[e])) ++e;
    if (e - i == 1) {
      nb[keys[i].h] = -1;  // e_neigh_code_out_bound, lib/geometry/include/geo_types.h:145
    } else if (e - i == 2 && (keys[i].h >> 2) != (keys[i + 1].h >> 2)) {
      nb[keys[i].h] = (int32_t)keys[i + 1].h;
      nb[keys[i + 1].h] = (int32_t)keys[i].h;
    } else {
      ok = false;
#pragma omp critical
      bad = keys[i].h >> 2;
    }
  }
  if (!ok) return fail("cell " + std::to_string(bad) + " shares a face with more than one cell");
  return true;
}

}  // namespace

extern "C" int rhllc_tet_geometry(int64_t n_points, const double *points, int64_t n_cells, const int32_t *tets,
                                  int32_t bound_rule, int32_t *neighbors, double *normals, double *areas,
                                  double *volumes, double *centers) {
  if (!points || !tets || !neighbors || !normals || !areas || !volumes || !centers) return fail("null array"), 1;
  if (n_cells < 0 || n_points < 0) return fail("negative size"), 1;
  if (4 * n_cells > INT32_MAX) return fail("4 * n_cells does not fit the int32 neighbour codes"), 1;
  if (bound_rule != RHLLC_GEO_BOUND_CONE && bound_rule != RHLLC_GEO_BOUND_FREE) return fail("unknown bound_rule"), 1;
  bool ok = true;
  int64_t bad = -1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int64_t i = 0; i < n_cells; ++i) {
    const int32_t *t = tets + 4 * i;
    bool good = true;
    for (int q = 0; q < 4; ++q) good = good && t[q] >= 0 && t[q] < n_points;
    good = good && t[0] != t[1] && t[0] != t[2] && t[0] != t[3] && t[1] != t[2] && t[1] != t[3] && t[2] != t[3];
    if (!good) {
      ok = false;
#pragma omp critical
      bad = i;
    }
  }
  if (!ok) return fail("cell " + std::to_string(bad) + " has a point id out of range or repeated"), 1;
  if (!neighbours(n_cells, tets, neighbors)) return 1;

#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    const int32_t *t = tets + 4 * i;
    const V3 P[4] = {pt(points, t[0]), pt(points, t[1]), pt(points, t[2]), pt(points, t[3])};

    // Face normals and areas: the unit normal of (p1 - p0) x (p2 - p0) over
    // the face's points in table order, turned to point away from the
    // fourth point.
    for (int j = 0; j < 4; ++j) {
      const V3 &p0 = P[kFace[j][0]], &p1 = P[kFace[j][1]], &p2 = P[kFace[j][2]];
      const V3 n = cross(sub(p1, p0), sub(p2, p0));
      const double len2 = n.x * n.x + n.y * n.y + n.z * n.z;
      const double len = std::sqrt(len2);
      areas[4 * i + j] = 0.5 * len;
      V3 u = n;
      if (len2 > 0.0) u = {n.x / len, n.y / len, n.z / len};
      if (dot(n, sub(P[kOpposite[j]], p0)) > 0.0) u = {-u.x, -u.y, -u.z};
      double *o = normals + 12 * i + 3 * j;
      o[0] = u.x, o[1] = u.y, o[2] = u.z;
    }

    // Signed volume -(a . (b x c)) / 6 over the cell's point order.
    const V3 a = sub(P[1], P[0]), b = sub(P[2], P[0]), c = sub(P[3], P[0]);
    volumes[i] = -dot(a, cross(b, c)) / 6.0;

    // Centre: the points weighted by the areas of their opposite faces (the
    // insphere centre).
    const double w0 = tri_area(P[1], P[2], P[3]), w1 = tri_area(P[0], P[2], P[3]);
    const double w2 = tri_area(P[0], P[1], P[3]), w3 = tri_area(P[0], P[1], P[2]);
    const double ws = w0 + w1 + w2 + w3;
    centers[3 * i + 0] = (w0 * P[0].x + w1 * P[1].x + w2 * P[2].x + w3 * P[3].x) / ws;
    centers[3 * i + 1] = (w0 * P[0].y + w1 * P[1].y + w2 * P[2].y + w3 * P[3].y) / ws;
    centers[3 * i + 2] = (w0 * P[0].z + w1 * P[1].z + w2 * P[2].z + w3 * P[3].z) / ws;

    // Boundary codes by the face normal (SetTypeOfBound).
    for (int j = 0; j < 4; ++j) {
      int32_t &code = neighbors[4 * i + j];
      if (code >= 0) continue;
      if (bound_rule == RHLLC_GEO_BOUND_FREE) {
        code = RHLLC_BOUND_FREE;
        continue;
      }
      // GEOMETRY_TYPE Cone (include/solvers_config.h:29): a face whose normal
      // is within 1e-7 of -x is the emitting bottom, of +x the outflow
      // surface, anything else a wall.
      const double *n = normals + 12 * i + 3 * j;
      const V3 dm = {n[0] - (-1.0), n[1] - 0.0, n[2] - 0.0};
      const V3 dp = {n[0] - 1.0, n[1] - 0.0, n[2] - 0.0};
      if (norm(dm) < 1e-7)
        code = RHLLC_BOUND_OUT_SOURCE;
      else if (norm(dp) < 1e-7)
        code = RHLLC_BOUND_OUTER_SURFACE;
      else
        code = RHLLC_BOUND_LOCK;
    }
  }
  return 0;
}

extern "C" int rhllc_bin_to_geo(int64_t n_cells, const int32_t *neighbors, const double *normals, const double *areas,
                                const double *volumes, const double *centers, int64_t *n_faces, void *geo_cells,
                                void *geo_faces) {
  if (!neighbors || n_cells < 0) return fail("null array or negative size"), 1;
  if (n_cells >= (int64_t)1 << 25) return fail("more than 2^25-1 cells do not fit the geo_face_t id fields"), 1;
  const int64_t H = 4 * n_cells;
  // A half-face creates a face when ConvertBinToGeo reaches it unmarked: it is
  // a boundary half-face, or the lower-numbered half of a shared face.
  bool ok = true;
  int64_t bad = -1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int64_t h = 0; h < H; ++h) {
    const int32_t k = neighbors[h];
    const bool good = k < 0 ? k >= -5 : (k < H && (k >> 2) != (h >> 2) && neighbors[k] == (int32_t)h);
    if (!good) {
      ok = false;
#pragma omp critical
      bad = h;
    }
  }
  if (!ok) return fail("neighbour entry " + std::to_string(bad) + " is not a boundary code -1..-5 or a mutual pair"), 1;

  std::vector<int32_t> fid((size_t)H);
  int64_t count = 0;
  for (int64_t h = 0; h < H; ++h)
    if (neighbors[h] < 0 || h < neighbors[h]) fid[h] = (int32_t)count++;
  if (n_faces) *n_faces = count;
  if (!geo_cells && !geo_faces) return 0;
  if (!geo_cells || !geo_faces || !normals || !areas || !volumes || !centers) return fail("null array"), 1;

  uint8_t *cells = static_cast<uint8_t *>(geo_cells);
  uint8_t *faces = static_cast<uint8_t *>(geo_faces);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    int32_t ids[4];
    uint8_t sign = 0, nbound = 0;
    for (int j = 0; j < 4; ++j) {
      const int64_t h = 4 * i + j;
      const int32_t k = neighbors[h];
      if (k < 0 || h < k) {
        ids[j] = fid[h];
        sign |= (uint8_t)(1u << j);
        if (k < 0) nbound = (uint8_t)((nbound + 1) & 3);  // 2-bit is_bound counter
        pack_face(faces + (int64_t)fid[h] * RHLLC_GEO_FACE_BYTES, (uint32_t)i, k < 0 ? k : (int32_t)(k >> 2),
                  normals + 3 * h, areas[h]);
      } else {
        ids[j] = fid[k];
      }
    }
    pack_cell(cells + i * RHLLC_GEO_CELL_BYTES, ids, volumes[i], centers + 3 * i, sign, nbound);
  }
  return 0;
}
