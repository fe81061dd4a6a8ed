// Record types of the reference's RHLLC interface, layout-compatible with
// lib/grids/include/solvers_struct.h (flux_t :51-70, geo_face_t :87-102,
// face_t :108-113, geo_cell_t :120-154, elem_t :280-294, grid_t :297-315) and
// the JSON config structs of lib/json/include/json_struct.h.  Vector3 is a
// plain 3-double record instead of Eigen's (same 24-byte layout), so code
// written against the reference's field names compiles unchanged for the
// fields the RHLLC step uses.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

typedef double Type;  // lib/global_types.h:12
typedef int64_t IdType;

enum e_type_completion { e_completion_success = 0, e_completion_fail = 1 };  // lib/global_types.h:48-51
enum e_boundary_types_t {                                                    // lib/global_types.h:25-31
  e_bound_free = -1,
  e_bound_lock = -2,
  e_bound_out_source = -3,
  e_bound_inner_source = -4,
  e_bound_outer_surface = -5
};

struct Vector3 {
  Type x[3];
  Vector3() : x{0, 0, 0} {}
  Vector3(Type a, Type b, Type c) : x{a, b, c} {}
  Type &operator[](int i) { return x[i]; }
  Type operator[](int i) const { return x[i]; }
  Type dot(const Vector3 &o) const { return (x[0] * o.x[0] + x[1] * o.x[1]) + x[2] * o.x[2]; }
};

struct flux_t {
  Type d;    // density (W) or D (U)
  Vector3 v; // velocity (W) or momentum (U)
  Type p;    // pressure (W) or energy E (U)
  flux_t() : d(0), v(0, 0, 0), p(0) {}
  flux_t(Type a, const Vector3 &b, Type c) : d(a), v(b), p(c) {}
};

struct flux_all_t {
  flux_t phys_val;
  flux_t conv_val;
};

struct geo_face_t {
  uint32_t id_l : 25;
  uint32_t id_l_node : 6;
  uint32_t is_regular : 1;
  int id_r : 26;
  uint32_t id_r_node : 6;
  Vector3 n;
  Type S;
  geo_face_t() : id_l(0), id_l_node(0), is_regular(1), id_r(0), id_r_node(0), n(0, 0, 0), S(0) {}
};

struct face_t {
  flux_t f;
  geo_face_t geo;
};

struct geo_cell_t {
  struct bits_flag {
    uint8_t bits;
    bits_flag(uint8_t a = 0) : bits(a) {}
    bool operator[](int i) const { return (bits >> i) & 1; }
    void set_sign(int idx, bool sign) { bits = sign ? (uint8_t)(bits | (1u << idx)) : (uint8_t)(bits & ~(1u << idx)); }
  };
  int id_faces[4];
  Type V;
  Vector3 center;
  bits_flag sign_n;
  uint8_t node : 6;
  uint8_t is_bound : 2;
  geo_cell_t() : id_faces{0, 0, 0, 0}, V(0), center(0, 0, 0), sign_n(0x7), node(0), is_bound(0) {}
};

struct elem_t {
  flux_t phys_val;
  flux_t conv_val;
  geo_cell_t geo;
};

struct grid_t {
  IdType size = 0;
  std::vector<elem_t> cells;
  std::vector<face_t> faces;
};

static_assert(sizeof(flux_t) == 40, "flux_t layout");
static_assert(sizeof(geo_face_t) == 40, "geo_face_t layout");
static_assert(sizeof(face_t) == 80, "face_t layout");
static_assert(sizeof(geo_cell_t) == 56, "geo_cell_t layout");
static_assert(sizeof(elem_t) == 136, "elem_t layout");

// lib/json/include/json_struct.h:15-26 (all keys required).  The drop-in also
// reads optional keys the reference ignores: "eos" ("taub", the reference's
// shipped EOS, include/hd_config.h:20, or "ideal"), "gamma" (index of the
// ideal gas, default 5/3 = the reference's kGamma1), "precision" ("fp64" |
// "mixed") and "c2p_start" ("reference" | "previous" | "mlp") with
// "mlp_model".
struct hllc_value_t {
  Type T = 0, CFL = 0, h_min = 0, save_timer = 0, tau = 0;
  std::string eos = "taub";
  Type gamma = 5.0 / 3.0;
  std::string precision = "fp64";
  std::string c2p_start = "reference";
  std::string mlp_model;
  std::string geometry = "cone";  // the reference's GEOMETRY_TYPE (include/solvers_config.h:19-29)
};

// lib/json/include/json_struct.h:39-55
struct solve_mode_t {
  int class_vtk = 0;
  int max_number_of_iter = 1;
  double accuracy = 1e-5;
  int start_point = 0;
};

// lib/json/include/json_struct.h:103-158; Build() derives the geometry file
// names from base_address (lib/json/src/json_struct.cpp:22-47).
struct global_files_t {
  std::string name_file_vtk, name_file_sphere_direction, base_address, illum_geo_address, graph_address,
      trace_address, solve_address, add_dir_address, tab_func_address, name_file_hllc_set, hllc_init_value,
      solve_configuration;
  std::string name_file_settings;
  std::string name_file_geometry_faces, name_file_geometry_cells;
  void Build() {
    name_file_geometry_faces = base_address + "geo_faces.bin";
    name_file_geometry_cells = base_address + "geo_cells.bin";
  }
};
