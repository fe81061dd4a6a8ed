// Host mesh tools: h_min, initial states, radiation fields, Hilbert relabelling
// and the "int32 n + n records" binary files of the reference.
#include "mesh_common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

using namespace rhllc_mesh;

extern "C" const char *rhllc_mesh_last_error(void) { return err_storage().c_str(); }

// utils/src/get_average_volume.cpp:31-46: r = (3/4) V / sum_j S_j, min over cells.
extern "C" double rhllc_mesh_h_min(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces) {
  if (!geo_cells || !geo_faces || n_cells < 1) return fail("null array or empty mesh"), -1.0;
  {
    std::string err;
    if (!validate_geo(n_cells, geo_cells, n_faces, geo_faces, true, err)) return fail(err), -1.0;
  }
  const uint8_t *cells = static_cast<const uint8_t *>(geo_cells);
  const uint8_t *faces = static_cast<const uint8_t *>(geo_faces);
  double best = 1e10;
#pragma omp parallel for reduction(min : best) schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    const uint8_t *c = cells + i * RHLLC_GEO_CELL_BYTES;
    int32_t ids[4];
    cell_ids(c, ids);
    double surface = 0;
    for (int j = 0; j < 4; ++j) surface += face_S(faces + (int64_t)ids[j] * RHLLC_GEO_FACE_BYTES);
    const double r = (3. / 4.) * cell_V(c) / surface;
    best = std::min(best, r);
  }
  return best;
}

extern "C" int rhllc_init_state(int32_t kind, int64_t n_cells, const void *geo_cells, const double *params,
                                uint64_t seed, void *W_out) {
  if (!geo_cells || !W_out) return fail("null array"), 1;
  const uint8_t *cells = static_cast<const uint8_t *>(geo_cells);
  double *W = static_cast<double *>(W_out);
  double phase[4] = {0, 0, 0, 0};
  if (kind == RHLLC_INIT_SHEAR)
    for (int k = 0; k < 4; ++k) phase[k] = 2.0 * M_PI * u01(mix64((seed + 2) * 0x2545F4914F6CDD1Dull + (uint64_t)k));
  if (kind < 0 || kind > RHLLC_INIT_UNIFORM) return fail("unknown initial-state kind"), 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    double x[3];
    cell_center(cells + i * RHLLC_GEO_CELL_BYTES, x);
    double *w = W + 5 * i;
    w[1] = w[2] = w[3] = 0.0;
    switch (kind) {
      case RHLLC_INIT_SHOCK_TUBE: {
        // params: x0, rhoL, pL, rhoR, pR
        const double x0 = params ? params[0] : 0.5;
        const bool left = x[0] < x0;
        w[0] = left ? (params ? params[1] : 10.0) : (params ? params[3] : 1.0);
        w[4] = left ? (params ? params[2] : 13.33) : (params ? params[4] : 1e-6);
        break;
      }
      case RHLLC_INIT_BLAST: {
        // params: cx, cy, cz, radius, rho, p_in, p_out
        const double cx = params ? params[0] : 0.5, cy = params ? params[1] : 0.5, cz = params ? params[2] : 0.5;
        const double r0 = params ? params[3] : 0.1;
        const double dx = x[0] - cx, dy = x[1] - cy, dz = x[2] - cz;
        const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
        w[0] = params ? params[4] : 1.0;
        w[4] = r < r0 ? (params ? params[5] : 1000.0) : (params ? params[6] : 1.0);
        break;
      }
      case RHLLC_INIT_SHEAR: {
        // vx = 0.995 tanh(y / 0.01); rho = 1 (y > 0) / 0.1 (y < 0); p = 1;
        // vy = 1e-3 sum_{k=1..4} sin(2 pi k x + phi_k) + U(-1e-4, 1e-4).
        const double y = x[1];
        w[0] = y > 0.0 ? 1.0 : 0.1;
        w[1] = 0.995 * std::tanh(y / 0.01);
        double vy = 0.0;
        for (int k = 1; k <= 4; ++k) vy += std::sin(2.0 * M_PI * k * x[0] + phase[k - 1]);
        vy *= 1e-3;
        vy += 1e-4 * (2.0 * u01(mix64((seed + 2) ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull))) - 1.0);
        w[2] = vy;
        w[4] = 1.0;
        break;
      }
      default: {
        for (int k = 0; k < 5; ++k) w[k] = params ? params[k] : (k == 0 || k == 4 ? 1.0 : 0.0);
        break;
      }
    }
  }
  return 0;
}

extern "C" int rhllc_init_radiation(int64_t n_cells, uint64_t seed, double ur_lo, double ur_hi, double kmin,
                                    double kmax, float *Ur, float *Fr3, float *ka, float *ks) {
  if (!Ur || !Fr3 || !ka || !ks) return fail("null array"), 1;
  if (!(ur_lo > 0.0) || !(ur_hi >= ur_lo)) return fail("bad Ur range"), 1;
  const double l0 = std::log(ur_lo), l1 = std::log(ur_hi);
  const uint64_t base = (seed + 2) * 0xD1B54A32D192ED03ull;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_cells; ++i) {
    const uint64_t k = base ^ ((uint64_t)i * 8u);
    const double u = std::exp(l0 + (l1 - l0) * u01(mix64(k + 0)));
    const double f = u01(mix64(k + 1));
    const double z = 2.0 * u01(mix64(k + 2)) - 1.0;
    const double ph = 2.0 * M_PI * u01(mix64(k + 3));
    const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
    Ur[i] = (float)u;
    Fr3[3 * i + 0] = (float)(u * f * s * std::cos(ph));
    Fr3[3 * i + 1] = (float)(u * f * s * std::sin(ph));
    Fr3[3 * i + 2] = (float)(u * f * z);
    ka[i] = (float)(kmin + (kmax - kmin) * u01(mix64(k + 4)));
    ks[i] = (float)(kmin + (kmax - kmin) * u01(mix64(k + 5)));
  }
  return 0;
}

extern "C" int rhllc_curve_order(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
                                 int32_t *cell_new_to_old, int32_t *face_new_to_old) {
  if (!geo_cells || !geo_faces || !cell_new_to_old || !face_new_to_old) return fail("null array"), 1;
  {
    std::string err;
    if (!validate_geo(n_cells, geo_cells, n_faces, geo_faces, true, err)) return fail(err), 1;
  }
  curve_order_impl(n_cells, geo_cells, n_faces, geo_faces, cell_new_to_old, face_new_to_old);
  return 0;
}

extern "C" int64_t rhllc_file_count(const char *path) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path), -1;
  int32_t n = 0;
  const size_t got = std::fread(&n, sizeof(int32_t), 1, f);
  std::fclose(f);
  if (got != 1) return fail(std::string("cannot read count from ") + path), -1;
  return n;
}

extern "C" int rhllc_file_read_records(const char *path, size_t rec_bytes, int64_t n, void *out) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path), 1;
  int32_t cnt = 0;
  if (std::fread(&cnt, sizeof(int32_t), 1, f) != 1 || cnt != n) {
    std::fclose(f);
    return fail(std::string("record count mismatch in ") + path), 1;
  }
  const size_t got = std::fread(out, rec_bytes, (size_t)n, f);
  std::fclose(f);
  if (got != (size_t)n) return fail(std::string("short read in ") + path), 1;
  return 0;
}

extern "C" int rhllc_file_write_records(const char *path, size_t rec_bytes, int64_t n, const void *recs) {
  if (n > 0x7FFFFFFF) return fail("too many records for an int32 header"), 1;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(std::string("cannot create ") + path), 1;
  const int32_t cnt = (int32_t)n;
  bool ok = std::fwrite(&cnt, sizeof(int32_t), 1, f) == 1;
  ok = ok && std::fwrite(recs, rec_bytes, (size_t)n, f) == (size_t)n;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(std::string("write failed for ") + path), 1;
  return 0;
}
