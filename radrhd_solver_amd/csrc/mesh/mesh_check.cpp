// Validation of geo_cell_t / geo_face_t records (mesh_common.h validate_geo).
// Shared by the mesh library and the HIP solver library.
#include "mesh_common.h"

namespace rhllc_mesh {

namespace {
// First failing index of a check over [0, n), in parallel; -1 when none fails.
template <typename Pred>
int64_t first_failure(int64_t n, Pred bad) {
  int64_t first = -1;
#pragma omp parallel
  {
    int64_t mine = -1;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      if (mine < 0 && bad(i)) mine = i;
#pragma omp critical
    if (mine >= 0 && (first < 0 || mine < first)) first = mine;
  }
  return first;
}
}  // namespace

bool validate_geo(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
                  bool allow_out_source, std::string &err, int32_t *fl_out, int32_t *fr_out) {
  const int64_t N = n_cells, F = n_faces;
  const uint8_t *gc = static_cast<const uint8_t *>(geo_cells);
  const uint8_t *gf = static_cast<const uint8_t *>(geo_faces);
  RawVec<int32_t> own_l, own_r;
  int32_t *fl = fl_out, *fr = fr_out;
  if (!fl) own_l.resize(F), fl = own_l.data();
  if (!fr) own_r.resize(F), fr = own_r.data();
  auto face_bad = [&](int64_t f) {
    const uint8_t *r = gf + (size_t)f * RHLLC_GEO_FACE_BYTES;
    const int32_t l = (int32_t)face_id_l(r), rr = face_id_r(r);
    fl[f] = l;
    fr[f] = rr;
    return l >= N || rr >= N || rr < RHLLC_BOUND_OUTER_SURFACE || (rr == RHLLC_BOUND_OUT_SOURCE && !allow_out_source);
  };
  if (const int64_t f = first_failure(F, face_bad); f >= 0) {
    const std::string at = "face " + std::to_string(f) + ": ";
    if (fl[f] >= N)
      err = at + "id_l out of range";
    else if (fr[f] >= N)
      err = at + "id_r out of range";
    else if (fr[f] < RHLLC_BOUND_OUTER_SURFACE)
      err = at + "unknown boundary code " + std::to_string(fr[f]);
    else
      err = "mesh has e_bound_out_source (-3) faces but no out_source state was given";
    return false;
  }
  // 0 ok, 1 face id out of range, 2 slot disagrees (slot in *j), 3 V <= 0
  auto cell_check = [&](int64_t c, int *slot) {
    const uint8_t *r = gc + (size_t)c * RHLLC_GEO_CELL_BYTES;
    int32_t ids[4];
    cell_ids(r, ids);
    const uint8_t sign = r[48];
    for (int j = 0; j < 4; ++j) {
      if (ids[j] < 0 || ids[j] >= F) return 1;
      const bool is_left = (sign >> j) & 1;
      if (is_left ? fl[ids[j]] != c : fr[ids[j]] != c) {
        if (slot) *slot = j;
        return 2;
      }
    }
    return cell_V(r) > 0.0 ? 0 : 3;
  };
  if (const int64_t c = first_failure(N, [&](int64_t c) { return cell_check(c, nullptr) != 0; }); c >= 0) {
    int j = 0;
    const int why = cell_check(c, &j);
    const std::string at = "cell " + std::to_string(c) + ": ";
    if (why == 1)
      err = at + "face id out of range";
    else if (why == 2)
      err = at + "slot " + std::to_string(j) + " disagrees with its face's id_l/id_r and sign bit";
    else
      err = at + "V <= 0";
    return false;
  }
  return true;
}

}  // namespace rhllc_mesh
