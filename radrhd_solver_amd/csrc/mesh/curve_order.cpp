// Locality relabelling of a tetrahedral mesh along a Hilbert curve (host).
// Shared by the mesh library and the HIP solver library, which relabels on
// create.
#include "mesh_common.h"

#include <algorithm>
#include <cmath>
#include <vector>
#ifdef _OPENMP
#include <parallel/algorithm>
#endif

namespace rhllc_mesh {

namespace {
inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1FFFFF;
  v = (v | v << 32) & 0x1F00000000FFFFull;
  v = (v | v << 16) & 0x1F0000FF0000FFull;
  v = (v | v << 8) & 0x100F00F00F00F00Full;
  v = (v | v << 4) & 0x10C30C30C30C30C3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

// Hilbert index of 21-bit coordinates: Skilling, "Programming the Hilbert
// curve" (AIP Conf. Proc. 707, 2004), axes -> transposed index, then the bits
// interleaved with x[0] most significant in each triple.  Consecutive indices
// are face-adjacent boxes at every level; the Morton (Z) curve jumps across
// the domain at its level boundaries.
inline uint64_t hilbert3(uint32_t x0, uint32_t x1, uint32_t x2) {
  uint32_t X[3] = {x0, x1, x2};
  const uint32_t M = 1u << 20;
  for (uint32_t Q = M; Q > 1; Q >>= 1) {  // undo the excess work (rotations and reflections)
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      // bit set: invert the low bits of X[0]; clear: exchange them between X[0] and X[i]
      const uint32_t set = 0u - ((X[i] & Q) != 0);
      const uint32_t t = (X[0] ^ X[i]) & P & ~set;
      X[0] ^= (P & set) | t;
      X[i] ^= t;
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];  // Gray encode
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  return spread3(X[2]) | spread3(X[1]) << 1 | spread3(X[0]) << 2;
}
}  // namespace

void curve_order_impl(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
                      int32_t *cell_new_to_old, int32_t *face_new_to_old) {
  const uint8_t *cells = static_cast<const uint8_t *>(geo_cells);
  const uint8_t *faces = static_cast<const uint8_t *>(geo_faces);
  double lo0 = 1e300, lo1 = 1e300, lo2 = 1e300, hi0 = -1e300, hi1 = -1e300, hi2 = -1e300;
#pragma omp parallel for schedule(static) reduction(min : lo0, lo1, lo2) reduction(max : hi0, hi1, hi2)
  for (int64_t i = 0; i < n_cells; ++i) {
    double x[3];
    cell_center(cells + i * RHLLC_GEO_CELL_BYTES, x);
    lo0 = std::min(lo0, x[0]), lo1 = std::min(lo1, x[1]), lo2 = std::min(lo2, x[2]);
    hi0 = std::max(hi0, x[0]), hi1 = std::max(hi1, x[1]), hi2 = std::max(hi2, x[2]);
  }
  const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
  // One common scale for all axes keeps the curve's cells roughly cubic.
  double ext = 0;
  for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[a] - lo[a]);
  const double sc = ext > 0 ? (double)((1u << 21) - 1) / ext : 0.0;
  struct Key {
    uint64_t h;
    int32_t i;
    bool operator<(const Key &o) const { return h < o.h || (h == o.h && i < o.i); }
  };
  RawVec<Key> key(n_cells);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    double x[3];
    cell_center(cells + i * RHLLC_GEO_CELL_BYTES, x);
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = (uint32_t)std::min(2097151.0, std::max(0.0, (x[a] - lo[a]) * sc));
    key[i] = {hilbert3(q[0], q[1], q[2]), (int32_t)i};
  }
#ifdef _OPENMP
  __gnu_parallel::sort(key.begin(), key.end());
#else
  std::sort(key.begin(), key.end());
#endif
  RawVec<int32_t> new_of_old(n_cells);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    cell_new_to_old[i] = key[i].i;
    new_of_old[key[i].i] = (int32_t)i;
  }
  // Faces grouped by the earlier (new) index of their two cells, ties in the
  // old face order; a boundary face goes with its id_l.  The face kernel then
  // writes each flux next to the first cell that gathers it, and the second
  // reader, further along the curve, finds it still cached more often than
  // when faces follow id_l (cell kernel 1.32 against 1.35 ms at 24M tets).
  // Each cell lists the faces it is the earlier cell of (at most 4, from its
  // own slots, by old face index), so the grouping is a prefix sum over the
  // cells.
  RawVec<int32_t> own(4 * (size_t)n_cells);
  RawVec<int64_t> start(n_cells + 1);
  start[0] = 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    const uint8_t *c = cells + (size_t)cell_new_to_old[i] * RHLLC_GEO_CELL_BYTES;
    int32_t ids[4], *out = &own[4 * (size_t)i];
    cell_ids(c, ids);
    int n = 0;
    for (int j = 0; j < 4; ++j) {
      const uint8_t *fp = faces + (size_t)ids[j] * RHLLC_GEO_FACE_BYTES;
      const int32_t r = face_id_r(fp), l = new_of_old[face_id_l(fp)];
      if ((r >= 0 ? std::min(l, new_of_old[r]) : l) == i) {
        int k = n++;
        while (k > 0 && out[k - 1] > ids[j]) out[k] = out[k - 1], --k;
        out[k] = ids[j];
      }
    }
    start[i + 1] = n;
  }
  for (int64_t i = 0; i < n_cells; ++i) start[i + 1] += start[i];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i)
    for (int64_t k = start[i]; k < start[i + 1]; ++k) face_new_to_old[k] = own[4 * (size_t)i + (k - start[i])];
  (void)n_faces;  // every face is listed once: by its earlier cell
}

}  // namespace rhllc_mesh
