// Drop-in RHLLC interface on top of librhllc_hip.so (see rhllc_dropin.h).
#include "rhllc_dropin.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "../../include/rhllc_hip.h"
#include "../../include/rhllc_mesh.h"
#include "json_min.h"

hllc_value_t _hllc_cfg;
solve_mode_t _solve_mode;
global_files_t glb_files;

namespace {

[[noreturn]] void die(const char *what, const std::string &detail = "") {
  // DIE_IF of lib/dbgdef.h:50-65 logs the condition and aborts.
  std::fprintf(stderr, "DIE: %s %s\n", what, detail.c_str());
  std::fflush(stderr);
  std::abort();
}

// lib/global_consts.h:34-82 and lib/physics/include/gas_state.h:38-43
const double kM_hydrogen = 1.6735575 * 1e-24, k_boltzmann = 1.3807 * 1e-16, kDist = 1e13, kC_Light = 3 * 1e10,
             kGravity = 6.6726 * 1e-8, kEv = 11604525.0061657;
const double kMass = (kC_Light * kC_Light * kDist) / kGravity;
const double kTime = kDist / kC_Light;
const double kDensity = kMass / (kDist * kDist * kDist);
const double kPressure = kMass / (kDist * kTime * kTime);
const double kVelocity = kC_Light;
double GetPressure(double rho, double tem) {
  const double g_idealGasConst = (kM_hydrogen / k_boltzmann);
  const double g_idealGasConst_inv = kDensity / g_idealGasConst;
  return (g_idealGasConst_inv / kPressure) * tem * rho;
}

// rhllc::GetConvValue (rhllc_flux.cpp:7-16) of a fixed ghost state W.
void ghost_conv(double gamma, const double W[5], double U[5]) {
  const double Gg = gamma / (gamma - 1.0);
  const double Gamma = 1. / std::sqrt(1 - (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]));
  const double h = 1 + Gg * W[4] / W[0];
  const double dhGG = W[0] * h * Gamma * Gamma;
  U[0] = Gamma * W[0];
  U[1] = dhGG * W[1], U[2] = dhGG * W[2], U[3] = dhGG * W[3];
  U[4] = dhGG - W[4];
}

// Ghost state of e_bound_out_source faces under GEOMETRY_TYPE Cone
// (rhllc_utils.cpp:80-84): W = (rho0, 1e3 / kVelocity, 0, 0, GetPressure(rho0, 10 kEv)),
// U from W by rhllc::GetConvValue (rhllc_flux.cpp:7-16), which is not the
// clamped GetConvValueStab of the cells.
void cone_out_source(double gamma, double W[5], double U[5]) {
  W[0] = kM_hydrogen * 1e14 / kDensity;
  W[1] = 1e3 / kVelocity, W[2] = 0.0, W[3] = 0.0;
  W[4] = GetPressure(W[0], 10 * kEv);
  ghost_conv(gamma, W, U);
}

// The boundary branches of BoundConditions for the configured GEOMETRY_TYPE
// (rhllc_utils.cpp:45-133; the reference fixes it at compile time,
// include/solvers_config.h:29, default Cone).  Cone: -3 is the fixed inflow
// state, every other code copies the cell.  Every other type: -2 is a wall
// and -5 aborts; Sphere and Cone_JET give -4 a fixed state, Cylinder gives -3
// one; the source codes copy the cell otherwise.
void geometry_bounds(const std::string &g, double gamma, bool has_code3, rhllc_params &p) {
  static const char *kTypes[] = {"cube", "step", "cone", "cone_jet", "sphere", "cylinder", "test_ellipse",
                                 "main_ellipse"};
  bool known = false;
  for (const char *t : kTypes) known = known || g == t;
  if (!known) die("hllc_cfg.json: unknown \"geometry\"", g);
  if (g == "cone") {
    if (has_code3) {
      p.has_out_source = 1;
      cone_out_source(gamma, p.out_source_W, p.out_source_U);
    }
    return;
  }
  p.lock_wall = 1;
  p.reject_outer_surface = 1;
  if (g == "sphere" || g == "cone_jet") {
    const double Ws[5] = {0.1, 0.0, 0.0, 0.0, 1.0}, Wj[5] = {0.1, 0.99, 0.0, 0.0, 0.01};
    const double *W = g == "sphere" ? Ws : Wj;
    p.has_inner_source = 1;
    for (int k = 0; k < 5; ++k) p.inner_source_W[k] = W[k];
    ghost_conv(gamma, p.inner_source_W, p.inner_source_U);
  }
  if (g == "cylinder") {
    const double W[5] = {0.1, 0.99, 0.0, 0.0, 0.01};
    p.has_out_source = 1;
    for (int k = 0; k < 5; ++k) p.out_source_W[k] = W[k];
    ghost_conv(gamma, p.out_source_W, p.out_source_U);
  } else {
    p.out_source_copy = 1;
  }
}

struct DeviceGrid {
  rhllc_solver *h = nullptr;
  const elem_t *cells_ptr = nullptr;
  size_t n_cells = 0, n_faces = 0;
  bool resident = false;
  bool uploaded = false;
  void release() {
    if (h) rhllc_destroy(h);
    h = nullptr;
    cells_ptr = nullptr;
    uploaded = false;
  }
};
DeviceGrid g_dev;

int precision_code(const std::string &s) { return s == "mixed" ? RHLLC_PREC_MIXED : RHLLC_PREC_FP64; }
int eos_code(const std::string &s) {
  if (s == "taub") return RHLLC_EOS_TAUB;
  if (s == "ideal") return RHLLC_EOS_IDEAL;
  die("hllc_cfg.json: \"eos\" must be \"taub\" or \"ideal\", got", s);
}
int guess_code(const std::string &s) {
  if (s == "previous") return RHLLC_GUESS_PREVIOUS;
  if (s == "mlp") return RHLLC_GUESS_MLP;
  return RHLLC_GUESS_REFERENCE;
}

void ensure_device(grid_t &grid) {
  if (g_dev.h && g_dev.cells_ptr == grid.cells.data() && g_dev.n_cells == grid.cells.size() &&
      g_dev.n_faces == grid.faces.size())
    return;
  g_dev.release();
  const size_t N = grid.cells.size(), F = grid.faces.size();
  std::vector<uint8_t> gc(56 * N), gf(40 * F);
  for (size_t i = 0; i < N; ++i) std::memcpy(&gc[56 * i], &grid.cells[i].geo, 56);
  for (size_t i = 0; i < F; ++i) std::memcpy(&gf[40 * i], &grid.faces[i].geo, 40);
  rhllc_params p;
  std::memset(&p, 0, sizeof(p));
  p.eos = eos_code(_hllc_cfg.eos);
  p.gamma = _hllc_cfg.gamma;
  p.cfl = _hllc_cfg.CFL;
  p.h_min = _hllc_cfg.h_min;
  p.precision = precision_code(_hllc_cfg.precision);
  p.c2p_guess = guess_code(_hllc_cfg.c2p_start);
  p.reorder = 1;
  p.device = 0;
  bool has_code3 = false;
  for (size_t i = 0; i < F && !has_code3; ++i) has_code3 = grid.faces[i].geo.id_r == RHLLC_BOUND_OUT_SOURCE;
  geometry_bounds(_hllc_cfg.geometry, _hllc_cfg.gamma, has_code3, p);
  if (rhllc_create(&p, (int64_t)N, gc.data(), (int64_t)F, gf.data(), &g_dev.h) != RHLLC_OK)
    die("rhllc_create", rhllc_last_error());
  if (p.c2p_guess == RHLLC_GUESS_MLP) {
    std::vector<float> model(RHLLC_MLP_MODEL);
    std::FILE *f = std::fopen(_hllc_cfg.mlp_model.c_str(), "rb");
    if (!f || std::fread(model.data(), sizeof(float), model.size(), f) != model.size()) die("mlp_model", _hllc_cfg.mlp_model);
    std::fclose(f);
    if (rhllc_set_mlp(g_dev.h, model.data()) != RHLLC_OK) die("rhllc_set_mlp", rhllc_last_error());
  }
  g_dev.cells_ptr = grid.cells.data();
  g_dev.n_cells = N;
  g_dev.n_faces = F;
}

void download(grid_t &grid) {
  if (rhllc_download_state(g_dev.h, grid.cells.data(), sizeof(elem_t), 0, 40) != RHLLC_OK)
    die("rhllc_download_state", rhllc_last_error());
  std::vector<double> F(5 * grid.faces.size());
  if (rhllc_download_face_flux(g_dev.h, F.data()) != RHLLC_OK) die("rhllc_download_face_flux", rhllc_last_error());
  for (size_t i = 0; i < grid.faces.size(); ++i) {
    flux_t &f = grid.faces[i].f;
    f.d = F[5 * i], f.v = Vector3(F[5 * i + 1], F[5 * i + 2], F[5 * i + 3]), f.p = F[5 * i + 4];
  }
}

template <typename T>
int read_records(const std::string &name, std::vector<T> &data, size_t rec_bytes, size_t offset) {
  std::FILE *f = std::fopen(name.c_str(), "rb");
  if (!f) {
    std::fprintf(stderr, "Error open file %s\n", name.c_str());
    return e_completion_fail;
  }
  int n = 0;
  if (std::fread(&n, sizeof(int), 1, f) != 1 || n < 0) {
    std::fclose(f);
    return e_completion_fail;
  }
  data.resize(n);
  std::vector<uint8_t> rec(rec_bytes);
  for (int i = 0; i < n; ++i) {
    if (std::fread(rec.data(), rec_bytes, 1, f) != 1) {
      std::fclose(f);
      return e_completion_fail;
    }
    std::memcpy(reinterpret_cast<uint8_t *>(&data[i]) + offset, rec.data(), rec_bytes);
  }
  std::fclose(f);
  return e_completion_success;
}

}  // namespace

namespace rhllc {

Type max_signal_speed = 1;  // rhllc_utils.cpp:13

void SetDeviceResident(bool resident) { g_dev.resident = resident; }
void ReleaseDevice() { g_dev.release(); }
void SyncGridFromDevice(grid_t &grid) {
  if (g_dev.h && g_dev.uploaded) download(grid);
}

// rhllc_3d_stab.cpp:10-65 on the device.
void Hllc3dStab(const Type tau, grid_t &grid) {
  ensure_device(grid);
  if (!g_dev.resident || !g_dev.uploaded) {
    if (rhllc_upload_state(g_dev.h, grid.cells.data(), sizeof(elem_t), 0, 40) != RHLLC_OK)
      die("rhllc_upload_state", rhllc_last_error());
    g_dev.uploaded = true;
  }
  double speed = 0;
  const int rc = rhllc_step(g_dev.h, tau, &speed);
  if (rc == RHLLC_ERR_C2P) die("PhysPressureFix(el.conv_val, el.phys_val)", rhllc_last_error());
  if (rc != RHLLC_OK) die("Hllc3dStab", rhllc_last_error());
  max_signal_speed = speed;
  if (!g_dev.resident) download(grid);
}

// rhllc_utils.cpp:15-23
Type GetTimeStep(const hllc_value_t &hllc_set, const std::vector<elem_t> &cells) {
  (void)cells;
  const Type t = hllc_set.CFL * hllc_set.h_min / max_signal_speed;
  if (t < 0) die("t < 0");
  return t;
}

// rhllc_utils.cpp:25-33: GetConvValueStab per cell (on the device, same arithmetic).
void HllcPhysToConv(std::vector<elem_t> &cells) {
  const size_t n = cells.size();
  std::vector<double> W(5 * n), U(5 * n);
  for (size_t i = 0; i < n; ++i) std::memcpy(&W[5 * i], &cells[i].phys_val, 40);
  if (rhllc_eval_p2c(eos_code(_hllc_cfg.eos), _hllc_cfg.gamma, (int64_t)n, W.data(), U.data()) != RHLLC_OK)
    die("HllcPhysToConv", rhllc_last_error());
  for (size_t i = 0; i < n; ++i) {
    std::memcpy(static_cast<void *>(&cells[i].phys_val), &W[5 * i], 40);
    std::memcpy(static_cast<void *>(&cells[i].conv_val), &U[5 * i], 40);
  }
}

// rhllc_init.cpp:180-194.  Without an init file the reference's default for
// GEOMETRY_TYPE Cone (rhllc_init.cpp:129-132) is used; like the reference it
// requires <base_address>centers.bin to exist.
int Init(std::string &file_init_value, std::vector<elem_t> &cells) {
  if (files_sys::bin::ReadHllcInit(file_init_value, cells) != e_completion_success) {
    std::FILE *f = std::fopen((glb_files.base_address + "centers.bin").c_str(), "rb");
    if (!f) {
      std::fprintf(stderr, "Default rhllc value not set\n");
      return e_completion_fail;
    }
    std::fclose(f);
    for (auto &el : cells) {
      el.phys_val.d = kM_hydrogen * 1e14 / kDensity;
      el.phys_val.p = GetPressure(el.phys_val.d, kEv);
      el.phys_val.v = Vector3(0, 0, 0);
    }
  }
  HllcPhysToConv(cells);
  max_signal_speed = 1;
  _hllc_cfg.tau = GetTimeStep(_hllc_cfg, cells);
  return e_completion_success;
}

// rhllc_main.cpp:15-62.  The state stays on the device between steps and is
// brought back only to write a solution.
int RunRhllcModule() {
  grid_t grid;
  uint32_t err = 0;
  err |= files_sys::bin::ReadGridGeo(glb_files.name_file_geometry_faces, grid.faces);
  err |= files_sys::bin::ReadGridGeo(glb_files.name_file_geometry_cells, grid.cells);
  if (err) {
    std::fprintf(stderr, "Error reading \n");
    return e_completion_fail;
  }
  grid.size = (IdType)grid.cells.size();
  if (rhllc::Init(glb_files.hllc_init_value, grid.cells)) die("rhllc::Init(glb_files.hllc_init_value, grid.cells)");
  Type t = 0.0;
  Type cur_timer = 0;
  int res_count = _solve_mode.start_point;
  _hllc_cfg.tau = GetTimeStep(_hllc_cfg, grid.cells);
  files_sys::bin::WriteSolution(glb_files.solve_address + std::to_string(res_count++), grid);
  const bool was_resident = g_dev.resident;
  SetDeviceResident(true);
  while (t < _hllc_cfg.T) {
    Hllc3dStab(_hllc_cfg.tau, grid);
    t += _hllc_cfg.tau;
    cur_timer += _hllc_cfg.tau;
    if (cur_timer >= _hllc_cfg.save_timer) {
      SyncGridFromDevice(grid);
      if (files_sys::bin::WriteSolution(glb_files.solve_address + std::to_string(res_count++), grid) !=
          e_completion_success)
        die("WriteSolution");
      cur_timer = 0;
    }
    _hllc_cfg.tau = GetTimeStep(_hllc_cfg, grid.cells);
  }
  SyncGridFromDevice(grid);
  files_sys::bin::WriteSolution(glb_files.solve_address + std::to_string(res_count++), grid);
  SetDeviceResident(was_resident);
  ReleaseDevice();
  return e_completion_success;
}

// rhllc_main_mpi.cpp:18-84 on one GPU, i.e. one rank: the partition file
// (glb_files.base_address + "metis", ReadSimple text: count, then one rank id
// per cell) must put every cell on rank 0.  rhllc_mpi::Hllc3dStab at one rank
// is rhllc::Hllc3dStab (rhllc_3d_mpi.cpp:371-486: the same BoundConditions,
// GetFluxStab, slot-order flux sum, c2p and PhysPressureFix), and the
// solution is written by WriteSolutionMPI (flux.bin).
int RunRhllcMpiModule() {
  grid_t grid;
  uint32_t err = 0;
  err |= files_sys::bin::ReadGridGeo(glb_files.name_file_geometry_faces, grid.faces);
  err |= files_sys::bin::ReadGridGeo(glb_files.name_file_geometry_cells, grid.cells);
  if (err) {
    std::fprintf(stderr, "Error reading \n");
    return e_completion_fail;
  }
  grid.size = (IdType)grid.cells.size();
  {
    // Read before any device work; a partition over several ranks is an error return.
    std::ifstream in(glb_files.base_address + "metis");
    long n = -1;
    if (!in || !(in >> n)) {
      std::fprintf(stderr, "Error reading metis \n");
      return e_completion_fail;
    }
    if (n != (long)grid.size) {
      std::fprintf(stderr, "metis: %ld entries for %ld cells\n", n, (long)grid.size);
      return e_completion_fail;
    }
    for (long i = 0; i < n; ++i) {
      long r = -1;
      if (!(in >> r) || r != 0) {
        std::fprintf(stderr, "metis: cell %ld is not on rank 0; this step runs one rank\n", i);
        return e_completion_fail;
      }
    }
  }
  if (rhllc::Init(glb_files.hllc_init_value, grid.cells)) die("rhllc::Init(glb_files.hllc_init_value, grid.cells)");
  Type t = 0.0;
  Type cur_timer = 0;
  int res_count = _solve_mode.start_point;
  _hllc_cfg.tau = GetTimeStep(_hllc_cfg, grid.cells);
  files_sys::bin::WriteSolutionMPI(glb_files.solve_address + std::to_string(res_count++), grid);
  const bool was_resident = g_dev.resident;
  SetDeviceResident(true);
  while (t < _hllc_cfg.T) {
    Hllc3dStab(_hllc_cfg.tau, grid);
    t += _hllc_cfg.tau;
    cur_timer += _hllc_cfg.tau;
    if (cur_timer >= _hllc_cfg.save_timer) {
      SyncGridFromDevice(grid);
      if (files_sys::bin::WriteSolutionMPI(glb_files.solve_address + std::to_string(res_count++), grid) !=
          e_completion_success)
        die("WriteSolutionMPI");
      cur_timer = 0;
    }
    _hllc_cfg.tau = GetTimeStep(_hllc_cfg, grid.cells);
  }
  SyncGridFromDevice(grid);
  files_sys::bin::WriteSolutionMPI(glb_files.solve_address + std::to_string(res_count++), grid);
  SetDeviceResident(was_resident);
  ReleaseDevice();
  return e_completion_success;
}

}  // namespace rhllc

namespace files_sys {
namespace json {

namespace {
int load(const std::string &file, std::map<std::string, jsonmin::Value> &kv) {
  std::ifstream in(file);
  if (!in) {
    std::fprintf(stderr, "Error open file %s\n", file.c_str());
    return e_completion_fail;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  std::string err;
  if (!jsonmin::parse_object(ss.str(), kv, err)) {
    std::fprintf(stderr, "Error : %s: %s\n", file.c_str(), err.c_str());
    return e_completion_fail;
  }
  return e_completion_success;
}
// j.at(name) throws on a missing key in the reference (json_struct.h:33,62,167).
template <typename T>
bool need(const std::map<std::string, jsonmin::Value> &kv, const char *name, T &out, const std::string &file) {
  auto it = kv.find(name);
  if (it == kv.end()) {
    std::fprintf(stderr, "Error : key '%s' missing in %s\n", name, file.c_str());
    return false;
  }
  return jsonmin::get(it->second, out);
}
template <typename T>
void opt(const std::map<std::string, jsonmin::Value> &kv, const char *name, T &out) {
  auto it = kv.find(name);
  if (it != kv.end()) jsonmin::get(it->second, out);
}
}  // namespace

int Read(const std::string &file, global_files_t &st) {
  std::map<std::string, jsonmin::Value> kv;
  if (load(file, kv)) return e_completion_fail;
  bool ok = need(kv, "name_file_vtk", st.name_file_vtk, file) &&
            need(kv, "name_file_sphere_direction", st.name_file_sphere_direction, file) &&
            need(kv, "base_address", st.base_address, file) &&
            need(kv, "illum_geo_address", st.illum_geo_address, file) &&
            need(kv, "graph_address", st.graph_address, file) && need(kv, "trace_address", st.trace_address, file) &&
            need(kv, "solve_address", st.solve_address, file) && need(kv, "add_dir_address", st.add_dir_address, file) &&
            need(kv, "tab_func_address", st.tab_func_address, file) &&
            need(kv, "name_file_hllc_set", st.name_file_hllc_set, file) &&
            need(kv, "hllc_init_value", st.hllc_init_value, file) &&
            need(kv, "solve_configuration", st.solve_configuration, file);
  return ok ? e_completion_success : e_completion_fail;
}

int Read(const std::string &file, solve_mode_t &st) {
  std::map<std::string, jsonmin::Value> kv;
  if (load(file, kv)) return e_completion_fail;
  bool ok = need(kv, "class_vtk", st.class_vtk, file) && need(kv, "max_number_of_iter", st.max_number_of_iter, file) &&
            need(kv, "accuracy", st.accuracy, file) && need(kv, "start_point", st.start_point, file);
  return ok ? e_completion_success : e_completion_fail;
}

int Read(const std::string &file, hllc_value_t &st) {
  std::map<std::string, jsonmin::Value> kv;
  if (load(file, kv)) return e_completion_fail;
  bool ok = need(kv, "T", st.T, file) && need(kv, "CFL", st.CFL, file) && need(kv, "h_min", st.h_min, file) &&
            need(kv, "save_timer", st.save_timer, file) && need(kv, "tau", st.tau, file);
  opt(kv, "eos", st.eos);
  opt(kv, "gamma", st.gamma);
  opt(kv, "precision", st.precision);
  opt(kv, "c2p_start", st.c2p_start);
  opt(kv, "mlp_model", st.mlp_model);
  opt(kv, "geometry", st.geometry);
  return ok ? e_completion_success : e_completion_fail;
}

// lib/files_sys/include/reader_json.h:63-89
int ReadStartSettings(const std::string &file_set, global_files_t &glb, solve_mode_t *solve_mode,
                      hllc_value_t *hllc_conf) {
  if (Read(file_set, glb)) return e_completion_fail;
  glb.name_file_settings = file_set;
  glb.Build();
  if (solve_mode != nullptr && Read(glb.solve_configuration, *solve_mode)) return e_completion_fail;
  if (hllc_conf != nullptr && Read(glb.name_file_hllc_set, *hllc_conf)) return e_completion_fail;
  return e_completion_success;
}

}  // namespace json

namespace bin {

int ReadGridGeo(const std::string &name_file, std::vector<face_t> &faces) {
  return read_records(name_file, faces, 40, offsetof(face_t, geo));
}
int ReadGridGeo(const std::string &name_file, std::vector<elem_t> &cells) {
  return read_records(name_file, cells, 56, offsetof(elem_t, geo));
}
int ReadHllcInit(const std::string &file_init_value, std::vector<elem_t> &cells) {
  return read_records(file_init_value, cells, 40, offsetof(elem_t, phys_val));
}

// lib/files_sys/src/writer_bin.cpp:104-112 (WRITE_FILE_ELEM of writer_bin.h:24-35)
int WriteSolution(const std::string &main_dir, const grid_t &grid) {
  auto write = [&](const std::string &name, size_t off, size_t bytes) {
    std::FILE *f = std::fopen((main_dir + name).c_str(), "wb");
    if (!f) return false;
    const int n = (int)grid.cells.size();
    bool ok = std::fwrite(&n, sizeof(int), 1, f) == 1;
    for (const auto &el : grid.cells) ok = ok && std::fwrite(reinterpret_cast<const uint8_t *>(&el) + off, bytes, 1, f) == 1;
    return (std::fclose(f) == 0) && ok;
  };
  const size_t pv = offsetof(elem_t, phys_val);
  bool ok = write("density.bin", pv + offsetof(flux_t, d), sizeof(Type)) &&
            write("pressure.bin", pv + offsetof(flux_t, p), sizeof(Type)) &&
            write("velocity.bin", pv + offsetof(flux_t, v), sizeof(Vector3));
  return ok ? e_completion_success : e_completion_fail;
}

// lib/files_sys/src/writer_bin.cpp:120-170 at one rank: WriteSimpleMPI of
// flux.bin (writer_bin.h:240-263) writes the int count, then each cell's
// phys_val (MPI_phys_val_t: the 40-byte flux_t of an elem_t, mpi_ext.cpp:90).
int WriteSolutionMPI(const std::string &main_dir, const grid_t &grid) {
  std::FILE *f = std::fopen((main_dir + "flux.bin").c_str(), "wb");
  if (!f) return e_completion_fail;
  const int n = (int)grid.cells.size();
  bool ok = std::fwrite(&n, sizeof(int), 1, f) == 1;
  for (const auto &el : grid.cells) ok = ok && std::fwrite(&el.phys_val, sizeof(flux_t), 1, f) == 1;
  return (std::fclose(f) == 0) && ok ? e_completion_success : e_completion_fail;
}

}  // namespace bin
}  // namespace files_sys

// lib/geometry/src/make_internal_format.cpp:272-336, the part the RHLLC module
// reads: neighbors.bin, normals.bin, areas.bin, volume.bin and centers.bin
// under glb_files.base_address, from the tetrahedra of glb_files.name_file_vtk,
// with the boundary codes of GEOMETRY_TYPE Cone.  The radiation-transfer files
// the reference also writes there are not produced.
int BuildDataFromVTK(const global_files_t &glb_files) {
  rhllc_vtk_mesh *vm = nullptr;
  int64_t np = 0, nc = 0;
  if (rhllc_vtk_open(glb_files.name_file_vtk.c_str(), &vm, &np, &nc)) {
    std::fprintf(stderr, "Error reading file vtk: %s\n", rhllc_mesh_last_error());
    return e_completion_fail;
  }
  std::vector<double> points(3 * np), normals(12 * nc), areas(4 * nc), volumes(nc), centers(3 * nc);
  std::vector<int32_t> tets(4 * nc), neighbors(4 * nc);
  rhllc_vtk_copy(vm, points.data(), tets.data());
  rhllc_vtk_close(vm);
  std::printf("Grid has Cell: %lld\n", (long long)nc);
  if (rhllc_tet_geometry(np, points.data(), nc, tets.data(), RHLLC_GEO_BOUND_CONE, neighbors.data(), normals.data(),
                         areas.data(), volumes.data(), centers.data())) {
    std::fprintf(stderr, "Neighbor wasn't found: %s\n", rhllc_mesh_last_error());
    return e_completion_fail;
  }
  const std::string &b = glb_files.base_address;
  const bool ok = !rhllc_file_write_records((b + "neighbors.bin").c_str(), 4, 4 * nc, neighbors.data()) &&
                  !rhllc_file_write_records((b + "normals.bin").c_str(), 96, nc, normals.data()) &&
                  !rhllc_file_write_records((b + "areas.bin").c_str(), 8, 4 * nc, areas.data()) &&
                  !rhllc_file_write_records((b + "volume.bin").c_str(), 8, nc, volumes.data()) &&
                  !rhllc_file_write_records((b + "centers.bin").c_str(), 24, nc, centers.data());
  if (!ok) std::fprintf(stderr, "Error writing geometry files under %s\n", b.c_str());
  return ok ? e_completion_success : e_completion_fail;
}

// utils/template_functions/bin_files_to_geo.h:11-38: geo_faces.bin and
// geo_cells.bin from the five files above.
int BinToGeo(const std::string &base_address) {
  const std::string &b = base_address;
  const int64_t nc = rhllc_file_count((b + "volume.bin").c_str());
  if (nc < 0) {
    std::fprintf(stderr, "Error reading file volume\n");
    return e_completion_fail;
  }
  std::vector<double> normals(12 * nc), areas(4 * nc), volumes(nc), centers(3 * nc);
  std::vector<int32_t> neighbors(4 * nc);
  const char *what = nullptr;
  if (rhllc_file_count((b + "neighbors.bin").c_str()) != 4 * nc ||
      rhllc_file_read_records((b + "neighbors.bin").c_str(), 4, 4 * nc, neighbors.data()))
    what = "neighbours";
  else if (rhllc_file_count((b + "normals.bin").c_str()) != nc ||
           rhllc_file_read_records((b + "normals.bin").c_str(), 96, nc, normals.data()))
    what = "normals";
  else if (rhllc_file_count((b + "areas.bin").c_str()) != 4 * nc ||
           rhllc_file_read_records((b + "areas.bin").c_str(), 8, 4 * nc, areas.data()))
    what = "areas_faces";
  else if (rhllc_file_read_records((b + "volume.bin").c_str(), 8, nc, volumes.data()))
    what = "volume";
  else if (rhllc_file_count((b + "centers.bin").c_str()) != nc ||
           rhllc_file_read_records((b + "centers.bin").c_str(), 24, nc, centers.data()))
    what = "centers";
  if (what) {
    std::fprintf(stderr, "Error reading file %s\n", what);
    return e_completion_fail;
  }
  int64_t nf = 0;
  if (rhllc_bin_to_geo(nc, neighbors.data(), normals.data(), areas.data(), volumes.data(), centers.data(), &nf,
                       nullptr, nullptr)) {
    std::fprintf(stderr, "ConvertBinToGeo: %s\n", rhllc_mesh_last_error());
    return e_completion_fail;
  }
  std::vector<uint8_t> gc(56 * nc), gf(40 * nf);
  rhllc_bin_to_geo(nc, neighbors.data(), normals.data(), areas.data(), volumes.data(), centers.data(), &nf, gc.data(),
                   gf.data());
  const bool ok = !rhllc_file_write_records((b + "geo_cells.bin").c_str(), 56, nc, gc.data()) &&
                  !rhllc_file_write_records((b + "geo_faces.bin").c_str(), 40, nf, gf.data());
  return ok ? e_completion_success : e_completion_fail;
}
