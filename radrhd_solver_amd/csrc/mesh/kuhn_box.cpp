// Synthetic tetrahedral meshes in the reference's geometry record format.
//
// A box of nx*ny*nz hexahedra is cut into 6 Kuhn tetrahedra per hex; vertices
// get seeded jitter and the cell numbering an optional seeded permutation.
// The output is what the reference reads with ReadGridGeo
// (lib/files_sys/include/reader_bin.h:113-135): geo_cell_t / geo_face_t
// records whose face numbering and orientation follow ConvertBinToGeo
// (lib/geometry/src/convert_bin_to_geo.cpp:3-48).
//
// Kuhn tetrahedron of a hex for an axis permutation (a, b, c):
//   V0 = corner, V1 = V0 + e_a, V2 = V1 + e_b, V3 = corner + (1,1,1).
// Slot s of a cell is the face opposite vertex Vs.  The 6 tets tile the hex
// and conform across hexes (the subdivision is translation invariant), and
// the neighbour of every slot is known in closed form:
//   s=0 {V1,V2,V3} lies in x_a = 1  -> hex + e_a, perm (b,c,a), slot 3
//   s=1 {V0,V2,V3}                  -> same hex, perm (b,a,c), slot 1
//   s=2 {V0,V1,V3}                  -> same hex, perm (a,c,b), slot 2
//   s=3 {V0,V1,V2} lies in x_c = 0  -> hex - e_c, perm (c,a,b), slot 0
#include "mesh_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

int perm_index(int a, int b, int c) {
  for (int t = 0; t < 6; ++t)
    if (kPerm[t][0] == a && kPerm[t][1] == b && kPerm[t][2] == c) return t;
  return -1;
}

struct Box {
  int n[3];
  double org[3], h[3];
  int periodic;
  double jitter;
  uint64_t seed;

  int64_t n_hex() const { return (int64_t)n[0] * n[1] * n[2]; }

  // Coordinates of lattice point p (unwrapped, may equal n on any axis).
  void vertex(const int p[3], double x[3]) const {
    int w[3];
    for (int a = 0; a < 3; ++a) w[a] = (periodic >> a & 1) ? (p[a] % n[a]) : p[a];
    const uint64_t lin =
        ((uint64_t)w[2] * (uint64_t)(n[1] + 1) + (uint64_t)w[1]) * (uint64_t)(n[0] + 1) + (uint64_t)w[0];
    for (int a = 0; a < 3; ++a) {
      double j = 0.0;
      const bool on_wall = !(periodic >> a & 1) && (p[a] == 0 || p[a] == n[a]);
      if (jitter != 0.0 && !on_wall) {
        const double u = rhllc_mesh::u01(rhllc_mesh::mix64(seed ^ (lin * 3u + (uint64_t)a) * 0x9E3779B97F4A7C15ull));
        j = (2.0 * u - 1.0) * jitter * h[a];
      }
      x[a] = org[a] + (double)p[a] * h[a] + j;
    }
  }

  // The four vertices of tet t in hex (i, j, k).
  void tet(int64_t hex, int t, double V[4][3]) const {
    int c[3];
    c[0] = (int)(hex % n[0]);
    c[1] = (int)((hex / n[0]) % n[1]);
    c[2] = (int)(hex / ((int64_t)n[0] * n[1]));
    int p[3] = {c[0], c[1], c[2]};
    vertex(p, V[0]);
    p[kPerm[t][0]] += 1;
    vertex(p, V[1]);
    p[kPerm[t][1]] += 1;
    vertex(p, V[2]);
    p[kPerm[t][2]] += 1;
    vertex(p, V[3]);
  }

  // Neighbour (old cell id * 4 + slot) of (old cell, slot), or -1 at a free wall.
  int64_t neighbour(int64_t cell, int s) const {
    const int64_t hex = cell / 6;
    const int t = (int)(cell % 6);
    const int a = kPerm[t][0], b = kPerm[t][1], c = kPerm[t][2];
    int q[3];
    q[0] = (int)(hex % n[0]);
    q[1] = (int)((hex / n[0]) % n[1]);
    q[2] = (int)(hex / ((int64_t)n[0] * n[1]));
    int t2, s2;
    if (s == 0) {
      q[a] += 1;
      t2 = perm_index(b, c, a);
      s2 = 3;
      if (q[a] == n[a]) {
        if (!(periodic >> a & 1)) return -1;
        q[a] = 0;
      }
    } else if (s == 1) {
      t2 = perm_index(b, a, c);
      s2 = 1;
    } else if (s == 2) {
      t2 = perm_index(a, c, b);
      s2 = 2;
    } else {
      q[c] -= 1;
      t2 = perm_index(c, a, b);
      s2 = 0;
      if (q[c] < 0) {
        if (!(periodic >> c & 1)) return -1;
        q[c] = n[c] - 1;
      }
    }
    const int64_t hex2 = ((int64_t)q[2] * n[1] + q[1]) * n[0] + q[0];
    return (hex2 * 6 + t2) * 4 + s2;
  }
};

bool make_box(const rhllc_box_desc *d, Box &B) {
  if (!d) return rhllc_mesh::fail("null box descriptor");
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return rhllc_mesh::fail("box sizes must be >= 1");
  const int nn[3] = {d->nx, d->ny, d->nz};
  for (int a = 0; a < 3; ++a) {
    if ((d->periodic >> a & 1) && nn[a] < 2) return rhllc_mesh::fail("a periodic axis needs >= 2 hexes");
    if (!(d->length[a] > 0.0)) return rhllc_mesh::fail("box lengths must be > 0");
    B.n[a] = nn[a];
    B.org[a] = d->origin[a];
    B.h[a] = d->length[a] / nn[a];
  }
  if ((int64_t)nn[0] * nn[1] * nn[2] * 6 >= (int64_t)1 << 25)
    return rhllc_mesh::fail("more than 2^25 cells: geo_face_t::id_l is a 25-bit field");
  if (d->jitter < 0.0 || d->jitter >= 0.25) return rhllc_mesh::fail("jitter must be in [0, 0.25)");
  B.periodic = d->periodic & 7;
  B.jitter = d->jitter;
  B.seed = d->seed;
  return true;
}

inline void sub3(const double *a, const double *b, double *r) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
inline void cross3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Outward unit normal and area of the face opposite vertex s.
void face_geometry(const double V[4][3], int s, double nrm[3], double *area) {
  int idx[3], m = 0;
  for (int v = 0; v < 4; ++v)
    if (v != s) idx[m++] = v;
  double e1[3], e2[3], cr[3], to_opp[3];
  sub3(V[idx[1]], V[idx[0]], e1);
  sub3(V[idx[2]], V[idx[0]], e2);
  cross3(e1, e2, cr);
  const double len = std::sqrt(dot3(cr, cr));
  sub3(V[s], V[idx[0]], to_opp);
  const double sg = dot3(cr, to_opp) > 0.0 ? -1.0 : 1.0;
  for (int a = 0; a < 3; ++a) nrm[a] = sg * cr[a] / len;
  *area = 0.5 * len;
}

}  // namespace

extern "C" int rhllc_box_sizes(const rhllc_box_desc *d, int64_t *n_cells, int64_t *n_faces) {
  Box B;
  if (!make_box(d, B)) return 1;
  const int64_t N = B.n_hex() * 6;
  int64_t bnd = 0;
  const int64_t nx = B.n[0], ny = B.n[1], nz = B.n[2];
  if (!(B.periodic & 1)) bnd += 4 * ny * nz;
  if (!(B.periodic & 2)) bnd += 4 * nx * nz;
  if (!(B.periodic & 4)) bnd += 4 * nx * ny;
  if (n_cells) *n_cells = N;
  if (n_faces) *n_faces = (4 * N + bnd) / 2;
  return 0;
}

extern "C" int rhllc_box_build(const rhllc_box_desc *d, void *geo_cells, void *geo_faces) {
  Box B;
  if (!make_box(d, B)) return 1;
  if (!geo_cells || !geo_faces) return rhllc_mesh::fail("null output array"), 1;
  int64_t N = 0, F = 0;
  rhllc_box_sizes(d, &N, &F);

  // Cell permutation: old_of_new via a seeded Fisher-Yates shuffle.
  std::vector<int32_t> old_of_new(N), new_of_old(N);
  std::iota(old_of_new.begin(), old_of_new.end(), 0);
  if (d->permute) {
    uint64_t st = d->seed + 1;
    for (int64_t i = N - 1; i > 0; --i) {
      st += 0x9E3779B97F4A7C15ull;
      const uint64_t r = rhllc_mesh::mix64(st);
      const int64_t j = (int64_t)(r % (uint64_t)(i + 1));
      std::swap(old_of_new[i], old_of_new[j]);
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) new_of_old[old_of_new[i]] = (int32_t)i;

  // Neighbour table in the new numbering: new cell * 4 + slot, or -1.
  std::vector<int64_t> nb(4 * N);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int64_t c = old_of_new[i];
    for (int s = 0; s < 4; ++s) {
      const int64_t o = B.neighbour(c, s);
      nb[4 * i + s] = o < 0 ? -1 : (int64_t)new_of_old[o / 4] * 4 + (o % 4);
    }
  }

  // ConvertBinToGeo numbering: half-face h creates a face iff it is visited
  // first, i.e. it is a boundary half-face or h < its partner.  Face ids are
  // the running count of creating half-faces (a parallel exclusive scan).
  std::vector<int32_t> fid(4 * N);
  const int64_t H = 4 * N;
  int nth = 1;
#pragma omp parallel
  {
#pragma omp single
    nth = rhllc_mesh::num_threads();
  }
  std::vector<int64_t> part(nth + 1, 0);
#pragma omp parallel num_threads(nth)
  {
    const int t = rhllc_mesh::thread_id();
    const int64_t lo = H * t / nth, hi = H * (t + 1) / nth;
    int64_t cnt = 0;
    for (int64_t h = lo; h < hi; ++h) cnt += (nb[h] < 0 || h < nb[h]) ? 1 : 0;
    part[t + 1] = cnt;
#pragma omp barrier
#pragma omp single
    for (int k = 0; k < nth; ++k) part[k + 1] += part[k];
    int64_t run = part[t];
    for (int64_t h = lo; h < hi; ++h)
      if (nb[h] < 0 || h < nb[h]) fid[h] = (int32_t)run++;
  }
  if (part[nth] != F) return rhllc_mesh::fail("face count mismatch (internal)"), 1;

  uint8_t *cells = static_cast<uint8_t *>(geo_cells);
  uint8_t *faces = static_cast<uint8_t *>(geo_faces);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int64_t c = old_of_new[i];
    double V[4][3];
    B.tet(c / 6, (int)(c % 6), V);
    uint8_t *rec = cells + i * RHLLC_GEO_CELL_BYTES;
    std::memset(rec, 0, RHLLC_GEO_CELL_BYTES);
    int32_t ids[4];
    uint8_t sign = 0, nbound = 0;
    for (int s = 0; s < 4; ++s) {
      const int64_t h = 4 * i + s;
      const bool creates = nb[h] < 0 || h < nb[h];
      if (creates) {
        ids[s] = fid[h];
        sign |= (uint8_t)(1u << s);
        double nrm[3], area;
        face_geometry(V, s, nrm, &area);
        uint8_t *f = faces + (int64_t)fid[h] * RHLLC_GEO_FACE_BYTES;
        int32_t id_r;
        if (nb[h] < 0) {
          id_r = RHLLC_BOUND_FREE;
          nbound = (uint8_t)((nbound + 1) & 3);
        } else {
          id_r = (int32_t)(nb[h] / 4);
        }
        rhllc_mesh::pack_face(f, (uint32_t)i, id_r, nrm, area);
      } else {
        ids[s] = fid[nb[h]];
      }
    }
    double vol, cen[3];
    {
      double e1[3], e2[3], e3[3], cr[3];
      sub3(V[1], V[0], e1);
      sub3(V[2], V[0], e2);
      sub3(V[3], V[0], e3);
      cross3(e1, e2, cr);
      vol = std::fabs(dot3(cr, e3)) / 6.0;
      for (int a = 0; a < 3; ++a) cen[a] = (V[0][a] + V[1][a] + V[2][a] + V[3][a]) / 4.0;
    }
    rhllc_mesh::pack_cell(rec, ids, vol, cen, sign, nbound);
  }
  return 0;
}
