/*
 * rhllc_oracle.h -- CPU restatement of the reference's RHLLC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check and the CPU-baseline leg of bench.py call it to check the HIP path.
 * The product (radrhd_solver_amd, librhllc_hip.so) never links or calls it.
 *
 * Record layouts are the reference's AoS records so that the CPU timing sees
 * the reference's memory behaviour:
 *   flux_t 40 B, face_t 80 B (flux + geo_face_t), elem_t 136 B
 *   (phys_val, conv_val, geo_cell_t)  -- lib/grids/include/solvers_struct.h:51-294.
 *
 * Parity status of every function is listed in the header of rhllc_oracle.c.
 */
#ifndef RHLLC_ORACLE_H
#define RHLLC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double d;
  double v[3];
  double p;
} orc_flux_t; /* 40 B */

typedef struct {
  uint32_t w0; /* id_l:25 | id_l_node:6 | is_regular:1 */
  uint32_t w1; /* id_r:26 (signed) | id_r_node:6 */
  double n[3];
  double S;
} orc_geo_face_t; /* 40 B */

typedef struct {
  orc_flux_t f;
  orc_geo_face_t geo;
} orc_face_t; /* 80 B */

typedef struct {
  int32_t id_faces[4];
  double V;
  double center[3];
  uint8_t sign_n;
  uint8_t bits; /* node:6 | is_bound:2 */
} orc_geo_cell_t; /* 56 B */

typedef struct {
  orc_flux_t phys_val;
  orc_flux_t conv_val;
  orc_geo_cell_t geo;
} orc_elem_t; /* 136 B */

/* Options of the physics that the reference fixes at compile time. */
typedef struct {
  double gamma;      /* kGamma1, lib/global_consts.h:20 (constexpr 5/3 there); EOS IDEAL only */
  int has_out_source;   /* e_bound_out_source ghost state below is valid */
  orc_flux_t out_source_W, out_source_U;
  int eos;              /* ORC_EOS_IDEAL or ORC_EOS_TAUB (include/hd_config.h:16-20) */
  /* The boundary branches of the other GEOMETRY_TYPE builds (rhllc_utils.cpp:58-110): */
  int lock_wall;        /* e_bound_lock (-2) mirrors the normal velocity (every type but Cone) */
  int has_inner_source; /* e_bound_inner_source (-4) takes the fixed state below (Sphere, Cone_JET) */
  orc_flux_t inner_source_W, inner_source_U;
} orc_physics_t;

#define ORC_EOS_IDEAL 0
#define ORC_EOS_TAUB 1

typedef struct {
  int64_t c2p_fail[5];   /* count by GetPhysValueStab return code 1..4 (index 0 unused) */
  int64_t fix_fail;      /* PhysPressureFix returned 2: the reference aborts here */
  int64_t lor_die;       /* the DIE_IF(lor2 < 1.0) condition inside Newton */
  int64_t newton_iters;  /* total Newton iterations */
} orc_stats_t;

/* Pointwise functions (one call = one state / one face). */
void orc_p2c(const orc_physics_t *ph, orc_flux_t *W, orc_flux_t *U);
int orc_c2p(const orc_physics_t *ph, const orc_flux_t *U, orc_flux_t *W, int *iters);
int orc_pressure_fix(const orc_physics_t *ph, orc_flux_t *U, orc_flux_t *W);
void orc_rotation(const double n[3], double T[9]); /* row-major T[3*r+c] */
void orc_wall_mirror(const double n[3], double v[3]); /* the e_bound_lock reflection */
double orc_flux(const orc_physics_t *ph, const orc_flux_t *UL, const orc_flux_t *UR, const orc_flux_t *WL,
                const orc_flux_t *WR, const double n[3], double S, orc_flux_t *F);
void orc_bound(const orc_physics_t *ph, const orc_face_t *f, const orc_elem_t *cells, orc_flux_t *W_ghost,
               orc_flux_t *U_ghost);

/* Vector forms over arrays of flux_t records (n states). */
void orc_p2c_n(const orc_physics_t *ph, int64_t n, orc_flux_t *W, orc_flux_t *U);
void orc_c2p_n(const orc_physics_t *ph, int64_t n, const orc_flux_t *U, orc_flux_t *W, int32_t *codes,
               int32_t *iters);
void orc_pressure_fix_n(const orc_physics_t *ph, int64_t n, orc_flux_t *U, orc_flux_t *W, int32_t *codes);
void orc_flux_n(const orc_physics_t *ph, int64_t n, const orc_flux_t *UL, const orc_flux_t *UR,
                const orc_flux_t *WL, const orc_flux_t *WR, const double *nS4, orc_flux_t *F, double *speed);

/* The step: rhllc::Hllc3dStab (solvers/rhllc/src/rhllc_3d_stab.cpp:10-65).
 * Returns the max signal speed; updates cells and face fluxes in place. */
double orc_step(const orc_physics_t *ph, int64_t n_cells, orc_elem_t *cells, int64_t n_faces,
                orc_face_t *faces, double tau, orc_stats_t *st);

/* Init + time loop of rhllc_main.cpp:28-56 for a fixed number of steps:
 * HllcPhysToConv, speed = 1, tau = CFL*h_min/speed, then {step; tau = ...}.
 * taus[k] / speeds[k] receive the tau used by and the speed returned by step k
 * (either may be NULL).  Returns the total simulated time. */
double orc_run(const orc_physics_t *ph, int64_t n_cells, orc_elem_t *cells, int64_t n_faces, orc_face_t *faces,
               double cfl, double h_min, int steps, double *taus, double *speeds, orc_stats_t *st);

/* Fill AoS records from geometry records (56 B cells / 40 B faces) and W (n x 5). */
void orc_load(int64_t n_cells, const void *geo_cells, const double *W, int64_t n_faces, const void *geo_faces,
              orc_elem_t *cells, orc_face_t *faces);

/* Radiation 4-momentum source (solvers/RadRHD/src/radRHD_utils.cpp:8-101).
 * D(j, i) is the Eddington tensor by the M1 closure (the ISOTROPIC_INTENSITY
 * == 1 branch) when Tr9 is NULL, else the pressure tensor Tr(j, i) =
 * Tr9[3 j + i] (ISOTROPIC_INTENSITY 0, the reference's compiled setting,
 * radRHD_config.h:8).  emission != 0 adds the thermal term
 * q = -ka*rho*B(T)/kRadiation with the reference's CGS constants; G[4]
 * receives (G0, G1, G2, G3). */
double orc_edd_tensor(int i, int j, double U, const double F[3]);
void orc_rad_source(const orc_flux_t *W, double Ur, const double Fr[3], double ka, double ks, int emission,
                    const float *Tr9, double G[4]);
/* TableFunc (lib/grids/include/solvers_struct.h:24-44): ln Lambda over
 * (ln D, ln T), data[i*ny + j]; ReadTableFunc (lib/files_sys/src/reader_txt.cpp:59-118)
 * fills it from tab_func/rcf_H.dat. */
typedef struct {
  int32_t nx, ny;
  double min_x, max_x, step_x, min_y, max_y, step_y;
  const double *data;
} orc_table_t;
/* ReadTableFunc: returns 0 and fills *nx, *ny, hdr = {min_x, max_x, step_x,
 * min_y, max_y, step_y} and data (when not NULL, nx*ny values); nonzero on a
 * missing or malformed file. */
int orc_table_read(const char *path, int32_t *nx, int32_t *ny, double hdr[6], double *data);
/* TableFunc::operator() (lib/grids/src/solvers_struct.cpp:10-58). */
double orc_table_lookup(const orc_table_t *t, double x, double y);
/* full_phys_data_t::Init(const flux_t *) (lib/grids/src/solvers_struct.cpp:264-281):
 * cell[5] = (T, vel, lorenz, alpha, betta). */
void orc_cell_rad(const orc_table_t *t, const orc_flux_t *W, double cell[5]);
/* GetRadSourceOpt (radRHD_utils.cpp:103-158); Tr9 as for orc_rad_source. */
void orc_rad_source_opt(const orc_flux_t *W, const double cell[5], double Ur, const double Fr[3], const float *Tr9,
                        double G[4]);
/* Coupled step of radRHD_main.cpp:99-132: hydro step, then U += tau*G(W),
 * then c2p (return code ignored, as there).  Ur..ks are per cell.  With a
 * table (tab != NULL; ka and ks unused) the source step is AddRadFlux
 * (rhllc_3d_mpi.cpp:509-533: GetRadSourceOpt of each cell's Init) followed by
 * HllcConvToPhys (rhllc_3d_mpi.cpp:488-505: c2p, PhysPressureFix on failure). */
double orc_step_rad(const orc_physics_t *ph, int64_t n_cells, orc_elem_t *cells, int64_t n_faces,
                    orc_face_t *faces, double tau, const float *Ur, const float *Fr3, const float *ka,
                    const float *ks, int emission, const orc_table_t *tab, const float *Tr9, orc_stats_t *st);
double orc_run_rad(const orc_physics_t *ph, int64_t n_cells, orc_elem_t *cells, int64_t n_faces,
                   orc_face_t *faces, double cfl, double h_min, int steps, const float *Ur, const float *Fr3,
                   const float *ka, const float *ks, int emission, const orc_table_t *tab, const float *Tr9,
                   double *taus, double *speeds, orc_stats_t *st);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
