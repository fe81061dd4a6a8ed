// Host preparation of a solver handle's device layout (rhllc_layout.h).
#include "rhllc_layout.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../mesh/mesh_common.h"

namespace rhllc_layout {

using rhllc_mesh::cell_ids;
using rhllc_mesh::cell_V;
using rhllc_mesh::RawVec;

// S n into the 16-byte form described with Layout::Sn.
void pack_sn(const double e[3], float *o) {
  uint32_t w = 0;
  float h[3];
  for (int k = 0; k < 3; ++k) h[k] = (float)e[k];
  const float m = std::max(std::max(std::fabs(h[0]), std::fabs(h[1])), std::fabs(h[2]));
  uint32_t mb;
  std::memcpy(&mb, &m, 4);
  mb &= 0x7F800000u;  // 2^E (0 for a zero or subnormal vector: no correction)
  float pw;
  std::memcpy(&pw, &mb, 4);
  const double Q = (double)pw * 0x1p-30;
  if (Q > 0.0)
    for (int k = 0; k < 3; ++k) {
      // |e - h| <= ulp(h) / 2 <= 2^(E - 24) = 64 Q; 64 itself is clamped
      const long n = std::min(63L, std::max(-63L, std::lround((e[k] - (double)h[k]) / Q)));
      w |= ((uint32_t)n & 0x7Fu) << (7 * k);
    }
  o[0] = h[0], o[1] = h[1], o[2] = h[2];
  std::memcpy(&o[3], &w, 4);
}

bool build(const Options &opt, int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
           Layout &L, std::string &err) {
  const int32_t N = (int32_t)n_cells, F = (int32_t)n_faces;
  const uint8_t *gc = static_cast<const uint8_t *>(geo_cells);
  const uint8_t *gf = static_cast<const uint8_t *>(geo_faces);
  auto frec = [&](int64_t f) { return gf + (size_t)f * RHLLC_GEO_FACE_BYTES; };
  auto crec = [&](int64_t c) { return gc + (size_t)c * RHLLC_GEO_CELL_BYTES; };

  // Validate so every device gather stays in bounds.
  RawVec<int32_t> fl(F), fr(F);
  if (!rhllc_mesh::validate_geo(N, geo_cells, F, geo_faces, opt.has_out_source || opt.out_source_copy, err,
                                 fl.data(), fr.data()))
    return false;
  if (opt.reject_outer_surface)
    for (int32_t f = 0; f < F; ++f)
      if (fr[f] == RHLLC_BOUND_OUTER_SURFACE) {
        err = "face " + std::to_string(f) +
              ": e_bound_outer_surface (-5) faces abort the reference's step outside GEOMETRY_TYPE Cone";
        return false;
      }

  // Numbering on the device.  Faces are grouped by the earlier new index of
  // their two cells (a boundary face goes with its id_l): the Hilbert order
  // does that itself; without it the tiles need the same grouping.
  L.cell_new_to_old.resize(N);
  L.face_new_to_old.resize(F);
  if (opt.reorder) {
    rhllc_mesh::curve_order_impl(N, geo_cells, F, geo_faces, L.cell_new_to_old.data(), L.face_new_to_old.data());
  } else {
    for (int32_t i = 0; i < N; ++i) L.cell_new_to_old[i] = i;
    if (!opt.tiled) {
      for (int32_t i = 0; i < F; ++i) L.face_new_to_old[i] = i;
    } else {
      std::vector<int64_t> start((size_t)N + 1, 0);  // counting sort, stable in the old face order
      auto owner = [&](int32_t f) { return fr[f] >= 0 ? std::min(fl[f], fr[f]) : fl[f]; };
      for (int32_t f = 0; f < F; ++f) start[owner(f) + 1]++;
      for (int32_t i = 0; i < N; ++i) start[i + 1] += start[i];
      for (int32_t f = 0; f < F; ++f) L.face_new_to_old[start[owner(f)]++] = f;
    }
  }
  RawVec<int32_t> cnew(N), fnew(F);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < N; ++i) cnew[L.cell_new_to_old[i]] = i;
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < F; ++i) fnew[L.face_new_to_old[i]] = i;

  // Faces: cells, geometry.
  L.lr.resize(2 * (size_t)F);
  if (opt.mixed)
    L.Sn.resize(4 * (size_t)F);
  else
    L.nS.resize(4 * (size_t)F);
  RawVec<int32_t> nb(F);  // the neighbour across the face (device numbering), -1 on the boundary
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < F; ++i) {
    const int32_t f = L.face_new_to_old[i];
    const uint8_t *r = frec(f);
    double n[3], S;
    std::memcpy(n, r + 8, 24);
    std::memcpy(&S, r + 32, 8);
    const int32_t l = cnew[fl[f]];
    nb[i] = fr[f] >= 0 ? cnew[fr[f]] : -1;
    // Copy-type boundaries read the cell itself; codes -3 and -4 read a fixed
    // ghost state, held as cells N and N + 1 after the mesh's cells; a wall
    // (-2 with lock_wall) reads the cell itself, mirrored (~l < 0).
    L.lr[2 * (size_t)i] = l;
    int32_t right = l;
    if (fr[f] >= 0)
      right = nb[i];
    else if (fr[f] == RHLLC_BOUND_OUT_SOURCE && opt.has_out_source)
      right = N;
    else if (fr[f] == RHLLC_BOUND_INNER_SOURCE && opt.has_inner_source)
      right = N + 1;
    else if (fr[f] == RHLLC_BOUND_LOCK && opt.lock_wall)
      right = ~l;
    L.lr[2 * (size_t)i + 1] = right;
    if (opt.mixed) {
      const double e[3] = {n[0] * S, n[1] * S, n[2] * S};
      pack_sn(e, &L.Sn[4 * (size_t)i]);
    } else {
      double *q = &L.nS[4 * (size_t)i];
      q[0] = n[0], q[1] = n[1], q[2] = n[2], q[3] = S;
    }
  }

  // Cells: face slots (device face index, bit 31 when the cell is id_r) and V.
  RawVec<int32_t> cf(4 * (size_t)N);
  if (opt.mixed)
    L.Vinv.resize(N);
  else
    L.V.resize(N);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < N; ++i) {
    const uint8_t *r = crec(L.cell_new_to_old[i]);
    int32_t ids[4];
    cell_ids(r, ids);
    for (int j = 0; j < 4; ++j)
      cf[4 * (size_t)i + j] = fnew[ids[j]] | (((r[48] >> j) & 1) ? 0 : (int32_t)0x80000000u);
    const double V = cell_V(r);
    if (!opt.mixed)
      L.V[i] = V;
    else
      L.Vinv[i] = 1.0 / V;
  }
  if (!opt.tiled) {
    L.cf.assign(cf.begin(), cf.end());
    return true;
  }

  // Tiles of T consecutive cells.  A face belongs to the tile of its earlier
  // cell (own faces: contiguous, since faces are grouped by that cell) and is
  // also an extra face of the tile of its later cell when that is another one.
  const int32_t T = opt.tile_cells;
  if (opt.mixed && T > 256) {
    err = "tile layout: the mixed path's tiles hold at most 256 cells";
    return false;
  }
  const int32_t nt = (N + T - 1) / T;
  L.n_tiles = nt;
  // The own faces of tile t are [first[t], first[t + 1]); own_tile is
  // non-decreasing in the face index.
  auto own_tile = [&](int32_t i) {
    const int32_t l = L.lr[2 * (size_t)i], b = nb[i];
    return (b >= 0 ? std::min(l, b) : l) / T;
  };
  std::vector<int32_t> first(nt + 1);
  first[nt] = F;
#pragma omp parallel for schedule(static)
  for (int32_t t = 0; t < nt; ++t) {
    int32_t lo = 0, hi = F;  // first face with own_tile >= t
    while (lo < hi) {
      const int32_t mid = lo + (hi - lo) / 2;
      if (own_tile(mid) < t)
        lo = mid + 1;
      else
        hi = mid;
    }
    first[t] = lo;
  }
  // Extra faces of tile t: the faces its cells reach whose own tile is an
  // earlier one, each met in exactly one slot (that of its later cell), listed
  // by face index.  Per tile, in parallel: counts, then the lists.
  auto extras_of = [&](int32_t t, int32_t *out) {
    int n = 0;
    const int32_t c1 = std::min(N, (t + 1) * T);
    for (int32_t c = t * T; c < c1; ++c)
      for (int j = 0; j < 4; ++j) {
        const int32_t i = cf[4 * (size_t)c + j] & 0x7FFFFFFF;
        if (i < first[t]) out[n++] = i;
      }
    std::sort(out, out + n);
    return n;
  };
  L.tile_meta.resize(4 * (size_t)nt);
#pragma omp parallel
  {
    std::vector<int32_t> buf(4 * (size_t)T);
#pragma omp for schedule(static)
    for (int32_t t = 0; t < nt; ++t) {
      L.tile_meta[4 * (size_t)t + 1] = first[t + 1] - first[t];
      L.tile_meta[4 * (size_t)t + 3] = extras_of(t, buf.data());
    }
  }
  int64_t ne = 0;
  int32_t max_slots = 0;
  for (int32_t t = 0; t < nt; ++t) {
    int32_t *m = &L.tile_meta[4 * (size_t)t];
    m[0] = first[t];
    m[2] = (int32_t)ne;
    ne += m[3];
    max_slots = std::max(max_slots, m[1] + m[3]);
  }
  // A tile's cells touch at most 4 T distinct faces; slots are 15 bits.
  if (max_slots > 4 * T || max_slots > 0x7FFF || ne > 0x7FFFFFFF) {
    err = "tile layout: a tile has " + std::to_string(max_slots) + " faces";
    return false;
  }
  L.max_slots = max_slots;
  L.tile_extra.resize(ne);
  L.cell_slots.resize(2 * (size_t)N);
#pragma omp parallel for schedule(static)
  for (int32_t t = 0; t < nt; ++t) {
    const int32_t *m = &L.tile_meta[4 * (size_t)t];
    int32_t *ex = &L.tile_extra[m[2]];
    extras_of(t, ex);
    const int32_t c1 = std::min(N, (t + 1) * T);
    for (int32_t c = t * T; c < c1; ++c) {
      uint32_t s[4];
      for (int j = 0; j < 4; ++j) {
        const int32_t e = cf[4 * (size_t)c + j], i = e & 0x7FFFFFFF;
        const uint32_t slot = i >= first[t] ? (uint32_t)(i - first[t])
                                            : (uint32_t)(m[1] + (std::lower_bound(ex, ex + m[3], i) - ex));
        s[j] = slot | (e < 0 ? 0x8000u : 0u);
      }
      L.cell_slots[2 * (size_t)c] = s[0] | (s[1] << 16);
      L.cell_slots[2 * (size_t)c + 1] = s[2] | (s[3] << 16);
    }
  }
  // The extra faces' copies, entry k at F + k.
  L.lr.resize(2 * ((size_t)F + ne));
  if (opt.mixed) L.far.resize((size_t)F + ne);
  if (opt.mixed)
    L.Sn.resize(4 * ((size_t)F + ne));
  else
    L.nS.resize(4 * ((size_t)F + ne));
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < ne; ++k) {
    const size_t i = (size_t)L.tile_extra[k], j = (size_t)F + k;
    L.lr[2 * j] = L.lr[2 * i], L.lr[2 * j + 1] = L.lr[2 * i + 1];
    if (opt.mixed)
      std::memcpy(&L.Sn[4 * j], &L.Sn[4 * i], 16);  // the fourth word is bits, copied as such
    else
      for (int a = 0; a < 4; ++a) L.nS[4 * j + a] = L.nS[4 * i + a];
  }
  if (opt.mixed) {
    // Each entry's cell in its tile (bits 21-28 of the fourth word, bit 29 set
    // when it is the right cell) and the other one, or the boundary encoding
    // of the right side, in far: the tile's face work reads 4 bytes of cell
    // indices instead of 8.
#pragma omp parallel for schedule(static)
    for (int32_t t = 0; t < nt; ++t) {
      const int32_t *m = &L.tile_meta[4 * (size_t)t];
      auto put = [&](size_t j) {
        const int32_t x = L.lr[2 * j], y = L.lr[2 * j + 1];
        const bool right = x / T != t;  // else the left cell is the one in the tile
        const uint32_t near = (uint32_t)((right ? y : x) - t * T);
        L.far[j] = right ? x : y;
        uint32_t w;
        std::memcpy(&w, &L.Sn[4 * j + 3], 4);
        w |= (near << 21) | (right ? 1u << 29 : 0u);
        std::memcpy(&L.Sn[4 * j + 3], &w, 4);
      };
      for (int32_t k = 0; k < m[1]; ++k) put((size_t)m[0] + k);
      for (int32_t k = 0; k < m[3]; ++k) put((size_t)F + m[2] + k);
    }
  }
  return true;
}

}  // namespace rhllc_layout
