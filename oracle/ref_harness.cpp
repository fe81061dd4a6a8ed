// ref_harness.cpp -- C entry points around the reference's own RHLLC code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's hot-path sources where they lie under /root/reference
// (solvers/rhllc/src/rhllc_flux_stab.cpp, rhllc_3d_stab.cpp, rhllc_utils.cpp,
// lib/geometry/src/linear_alg.cpp, lib/grids/src/solvers_struct.cpp,
// lib/json/src/json_struct.cpp, solvers/rhllc/src/rhllc_flux.cpp,
// lib/geometry/src/convert_bin_to_geo.cpp, lib/files_sys/src/reader_txt.cpp
// and the serial lib/mpi_extension/src/*.cpp it calls) into
// oracle/_ref/libref_rhllc.so.  Tests use it to pin the C restatement in
// oracle/rhllc_oracle.c; it is never part of the product.
//
// The harness mirrors the driver of solvers/rhllc/src/rhllc_main.cpp:28-56:
// fill grid_t from geometry records and an initial W, HllcPhysToConv,
// max_signal_speed = 1, tau = GetTimeStep, then {Hllc3dStab; GetTimeStep}.
// EOS is compiled as IDEAL into libref_rhllc.so and as TAUB (the reference's
// shipped setting) into libref_rhllc_taub.so (see the Makefile); gamma is the reference's
// constexpr kGamma1 = 5/3.
#include <cstring>
#include <string>
#include <vector>

#include "convert_bin_to_geo.h"
#include "gas_state.h"
#include "global_value.h"
#include "rhllc_calc.h"
#include "rhllc_flux.h"
#include "rhllc_flux_stab.h"
#include "rhllc_utils.h"
#include "linear_alg.h"
#include "reader_txt.h"
#include "solvers_struct.h"

extern "C" {

int ref_layout(int *out) {
  out[0] = (int)sizeof(flux_t);
  out[1] = (int)sizeof(face_t);
  out[2] = (int)sizeof(elem_t);
  out[3] = (int)sizeof(geo_face_t);
  out[4] = (int)sizeof(geo_cell_t);
  out[5] = (int)offsetof(elem_t, conv_val);
  out[6] = (int)offsetof(elem_t, geo);
  out[7] = (int)offsetof(face_t, geo);
  out[8] = EOS;
  return 0;
}

double ref_gamma(void) { return kGamma1; }

void ref_set_log_dir(const char *dir) { glb_files.base_address = dir; }

void ref_p2c_n(int64_t n, void *W, void *U) {
  flux_t *w = static_cast<flux_t *>(W);
  flux_t *u = static_cast<flux_t *>(U);
  for (int64_t i = 0; i < n; ++i) rhllc::GetConvValueStab(w[i], u[i]);
}

void ref_c2p_n(int64_t n, const void *U, void *W, int32_t *codes) {
  const flux_t *u = static_cast<const flux_t *>(U);
  flux_t *w = static_cast<flux_t *>(W);
  for (int64_t i = 0; i < n; ++i) codes[i] = rhllc::GetPhysValueStab(u[i], w[i]);
}

void ref_pressure_fix_n(int64_t n, void *U, void *W, int32_t *codes) {
  flux_t *u = static_cast<flux_t *>(U);
  flux_t *w = static_cast<flux_t *>(W);
  for (int64_t i = 0; i < n; ++i) codes[i] = rhllc::PhysPressureFix(u[i], w[i]);
}

// T row-major per normal (9 doubles).
void ref_rotation_n(int64_t n, const double *normals, double *T) {
  for (int64_t i = 0; i < n; ++i) {
    Vector3 nv(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
    Matrix3 M;
    GetRotationMatrix(nv, M);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) T[9 * i + 3 * r + c] = M(r, c);
  }
}

// nS4 = (nx, ny, nz, S) per face.
void ref_flux_n(int64_t n, const void *UL, const void *UR, const void *WL, const void *WR, const double *nS4,
                void *F, double *speed) {
  const flux_t *ul = static_cast<const flux_t *>(UL);
  const flux_t *ur = static_cast<const flux_t *>(UR);
  const flux_t *wl = static_cast<const flux_t *>(WL);
  const flux_t *wr = static_cast<const flux_t *>(WR);
  flux_t *f = static_cast<flux_t *>(F);
  for (int64_t i = 0; i < n; ++i) {
    face_t face;
    face.geo.n = Vector3(nS4[4 * i], nS4[4 * i + 1], nS4[4 * i + 2]);
    face.geo.S = nS4[4 * i + 3];
    speed[i] = rhllc::GetFluxStab(ul[i], ur[i], wl[i], wr[i], face);
    f[i] = face.f;
  }
}

// Runs `steps` steps; W_out/U_out are n x 5 (d, v, p), taus/speeds per step.
int ref_run(int64_t n_cells, const void *geo_cells, const double *W0, int64_t n_faces, const void *geo_faces,
            double cfl, double h_min, int steps, double *W_out, double *U_out, double *taus, double *speeds) {
  grid_t grid;
  grid.cells.resize(n_cells);
  grid.faces.resize(n_faces);
  grid.size = n_cells;
  const uint8_t *gc = static_cast<const uint8_t *>(geo_cells);
  const uint8_t *gf = static_cast<const uint8_t *>(geo_faces);
  for (int64_t i = 0; i < n_cells; ++i) {
    std::memcpy((void *)&grid.cells[i].geo, gc + 56 * i, 56);
    flux_t &w = grid.cells[i].phys_val;
    w.d = W0[5 * i];
    w.v = Vector3(W0[5 * i + 1], W0[5 * i + 2], W0[5 * i + 3]);
    w.p = W0[5 * i + 4];
  }
  for (int64_t i = 0; i < n_faces; ++i) std::memcpy((void *)&grid.faces[i].geo, gf + 40 * i, 40);

  _hllc_cfg.CFL = cfl;
  _hllc_cfg.h_min = h_min;
  rhllc::HllcPhysToConv(grid.cells);
  rhllc::max_signal_speed = 1;
  _hllc_cfg.tau = rhllc::GetTimeStep(_hllc_cfg, grid.cells);
  for (int k = 0; k < steps; ++k) {
    rhllc::Hllc3dStab(_hllc_cfg.tau, grid);
    if (taus) taus[k] = _hllc_cfg.tau;
    if (speeds) speeds[k] = rhllc::max_signal_speed;
    _hllc_cfg.tau = rhllc::GetTimeStep(_hllc_cfg, grid.cells);
  }
  for (int64_t i = 0; i < n_cells; ++i) {
    const flux_t &w = grid.cells[i].phys_val;
    const flux_t &u = grid.cells[i].conv_val;
    double *wo = W_out + 5 * i, *uo = U_out + 5 * i;
    wo[0] = w.d, wo[1] = w.v[0], wo[2] = w.v[1], wo[3] = w.v[2], wo[4] = w.p;
    uo[0] = u.d, uo[1] = u.v[0], uo[2] = u.v[1], uo[3] = u.v[2], uo[4] = u.p;
  }
  return 0;
}

// ConvertBinToGeo (lib/geometry/src/convert_bin_to_geo.cpp:3-48) on the five
// preprocessing arrays; writes up to max_faces face records and returns the
// face count.
int64_t ref_convert_bin_to_geo(int64_t n_cells, const int32_t *nb, const double *normals, const double *areas,
                               const double *vol, const double *cen, int64_t max_faces, void *geo_cells,
                               void *geo_faces) {
  std::vector<IntId> neigh(nb, nb + 4 * n_cells);
  std::vector<Normals> nrm(n_cells);
  std::vector<Type> ar(areas, areas + 4 * n_cells), v(vol, vol + n_cells);
  std::vector<Vector3> c(n_cells);
  for (int64_t i = 0; i < n_cells; ++i) {
    for (int j = 0; j < 4; ++j)
      nrm[i].n[j] = Vector3(normals[12 * i + 3 * j], normals[12 * i + 3 * j + 1], normals[12 * i + 3 * j + 2]);
    c[i] = Vector3(cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]);
  }
  std::vector<face_t> faces;
  std::vector<elem_t> cells;
  ConvertBinToGeo(neigh, nrm, ar, v, c, faces, cells);
  uint8_t *gc = static_cast<uint8_t *>(geo_cells);
  uint8_t *gf = static_cast<uint8_t *>(geo_faces);
  for (int64_t i = 0; i < n_cells; ++i) std::memcpy(gc + 56 * i, (const void *)&cells[i].geo, 56);
  for (int64_t i = 0; i < (int64_t)faces.size() && i < max_faces; ++i)
    std::memcpy(gf + 40 * i, (const void *)&faces[i].geo, 40);
  return (int64_t)faces.size();
}

// The GEOMETRY_TYPE Cone states of the compiled reference: the default initial W of
// rhllc_init.cpp:130-132 and the e_bound_out_source ghost of
// rhllc_utils.cpp:80-84 (W, and U from rhllc::GetConvValue), from the
// reference's own constants and GetPressure (lib/physics/include/gas_state.h:38-43).
void ref_cone_states(double *W_init, double *W_src, double *U_src) {
  flux_t w0;
  w0.d = kM_hydrogen * 1e14 / kDensity;
  w0.p = GetPressure(w0.d, kEv);
  w0.v = Vector3(0, 0, 0);
  flux_t ws, us;
  ws.d = kM_hydrogen * 1e14 / kDensity;
  ws.p = GetPressure(ws.d, 10 * kEv);
  ws.v = Vector3(1e3 / kVelocity, 0, 0);
  rhllc::GetConvValue(ws, us);
  const flux_t *src[3] = {&w0, &ws, &us};
  double *dst[3] = {W_init, W_src, U_src};
  for (int k = 0; k < 3; ++k) {
    dst[k][0] = src[k]->d, dst[k][1] = src[k]->v[0], dst[k][2] = src[k]->v[1], dst[k][3] = src[k]->v[2];
    dst[k][4] = src[k]->p;
  }
}

// files_sys::txt::ReadTableFunc (lib/files_sys/src/reader_txt.cpp:59-118):
// hdr = {min_x, max_x, step_x, min_y, max_y, step_y}; data (nx*ny) when not NULL.
int ref_table_read(const char *path, int32_t *nx, int32_t *ny, double *hdr, double *data) {
  TableFunc t;
  if (files_sys::txt::ReadTableFunc(path, t) != e_completion_success) return 1;
  *nx = t.Nx, *ny = t.Ny;
  hdr[0] = t.min_x, hdr[1] = t.max_x, hdr[2] = t.step_x;
  hdr[3] = t.min_y, hdr[4] = t.max_y, hdr[5] = t.step_y;
  if (data) std::memcpy(data, t.data.data(), sizeof(double) * t.data.size());
  return 0;
}

// TableFunc::operator() (lib/grids/src/solvers_struct.cpp:10-58) at n points.
void ref_table_lookup(int32_t nx, int32_t ny, const double *hdr, const double *data, int64_t n, const double *x,
                      const double *y, double *out) {
  TableFunc t(nx, ny);
  t.min_x = hdr[0], t.max_x = hdr[1], t.step_x = hdr[2];
  t.min_y = hdr[3], t.max_y = hdr[4], t.step_y = hdr[5];
  std::memcpy(t.data.data(), data, sizeof(double) * nx * ny);
  for (int64_t i = 0; i < n; ++i) out[i] = t(x[i], y[i]);
}

}  // extern "C"
