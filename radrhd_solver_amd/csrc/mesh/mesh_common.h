// Shared helpers of the host mesh library (libradrhd_mesh.so).
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/rhllc_mesh.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace rhllc_mesh {

// SplitMix64 finaliser: a counter-based hash used for every seeded draw, so
// results do not depend on the thread count.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// Uniform double in [0, 1) from the top 53 bits.
inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

inline std::string &err_storage() {
  static thread_local std::string e;
  return e;
}
inline bool fail(const std::string &msg) {
  err_storage() = msg;
  return false;
}

// Cells sorted by the Hilbert index of their centres; faces by the earlier new
// index of their two cells (id_l for a boundary face), ties in the old face
// order.  The mesh must have passed validate_geo.
void curve_order_impl(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
                      int32_t *cell_new_to_old, int32_t *face_new_to_old);

// Checks what every consumer of the records indexes with: id_l / id_r in
// range, boundary codes known (-3 only when allow_out_source), each cell's
// face ids in range and agreeing with their face's id_l / id_r and sign bit,
// V > 0.  On success fl / fr (n_faces each, may be null) hold id_l and the
// sign-extended id_r; otherwise err names the first offending record.
bool validate_geo(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
                  bool allow_out_source, std::string &err, int32_t *fl = nullptr, int32_t *fr = nullptr);

// Allocator whose value-less construct leaves trivial elements uninitialised,
// so large host arrays are first touched by the parallel loops that fill them.
template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <typename U>
  NoInit(const NoInit<U> &) {}
  template <typename U>
  void construct(U *p) {
    ::new (static_cast<void *>(p)) U;
  }
  template <typename U, typename... Args>
  void construct(U *p, Args &&...args) {
    ::new (static_cast<void *>(p)) U(std::forward<Args>(args)...);
  }
};
template <typename T>
using RawVec = std::vector<T, NoInit<T>>;

inline int num_threads() {
#ifdef _OPENMP
  return omp_get_num_threads();
#else
  return 1;
#endif
}
inline int thread_id() {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

// geo_face_t record (40 B), lib/grids/include/solvers_struct.h:87-102.
// id_l_node = 0, is_regular = 1 (the default constructor's value), id_r_node = 0.
inline void pack_face(uint8_t *f, uint32_t id_l, int32_t id_r, const double n[3], double S) {
  const uint32_t w0 = (id_l & 0x1FFFFFFu) | (1u << 31);
  const uint32_t w1 = (uint32_t)id_r & 0x3FFFFFFu;
  std::memcpy(f + 0, &w0, 4);
  std::memcpy(f + 4, &w1, 4);
  std::memcpy(f + 8, n, 24);
  std::memcpy(f + 32, &S, 8);
}
inline uint32_t face_id_l(const uint8_t *f) {
  uint32_t w0;
  std::memcpy(&w0, f, 4);
  return w0 & 0x1FFFFFFu;
}
// Sign-extends the 26-bit id_r field.
inline int32_t face_id_r(const uint8_t *f) {
  uint32_t w1;
  std::memcpy(&w1, f + 4, 4);
  int32_t v = (int32_t)(w1 << 6);
  return v >> 6;
}
inline double face_S(const uint8_t *f) {
  double S;
  std::memcpy(&S, f + 32, 8);
  return S;
}

// geo_cell_t record (56 B), lib/grids/include/solvers_struct.h:120-154.
inline void pack_cell(uint8_t *c, const int32_t ids[4], double V, const double cen[3], uint8_t sign,
                      uint8_t is_bound) {
  std::memset(c, 0, RHLLC_GEO_CELL_BYTES);
  std::memcpy(c + 0, ids, 16);
  std::memcpy(c + 16, &V, 8);
  std::memcpy(c + 24, cen, 24);
  c[48] = sign;
  c[49] = (uint8_t)((is_bound & 3u) << 6);
}
inline void cell_ids(const uint8_t *c, int32_t ids[4]) { std::memcpy(ids, c, 16); }
inline double cell_V(const uint8_t *c) {
  double V;
  std::memcpy(&V, c + 16, 8);
  return V;
}
inline void cell_center(const uint8_t *c, double x[3]) { std::memcpy(x, c + 24, 24); }

}  // namespace rhllc_mesh
