/*
 * rhllc_hip.h -- C ABI of the MI355X (gfx950) RHLLC step, librhllc_hip.so.
 *
 * The boundary replaces the reference's step entry points (no plugin system
 * exists there; these are plain C++ free functions in namespace rhllc):
 *   rhllc::Hllc3dStab(const Type tau, grid_t &grid)  solvers/rhllc/include/rhllc_calc.h:32
 *     -> rhllc_step (one step at the caller's tau) / rhllc_run (device-side dt loop)
 *   rhllc::max_signal_speed                          solvers/rhllc/include/rhllc_utils.h:17
 *     -> the max_speed out-argument of rhllc_step / rhllc_stats.max_speed
 *   rhllc::GetTimeStep(hllc_set, cells)             solvers/rhllc/include/rhllc_utils.h:35
 *     -> computed on the device after every step (rhllc_stats.tau)
 *   rhllc::Init / HllcPhysToConv                    solvers/rhllc/include/rhllc_init.h:32,
 *                                                   solvers/rhllc/include/rhllc_utils.h:43
 *     -> rhllc_upload_state with off_conv < 0
 * The mesh enters as the reference's geo_cell_t / geo_face_t records
 * (include/rhllc_mesh.h), the state as any array of records holding flux_t
 * W and/or U at given byte offsets (elem_t: stride 136, W at 0, U at 40).
 *
 * Conventions: every call returns 0 on success or an RHLLC_ERR_* code, with a
 * message in rhllc_last_error().  Nothing aborts across the ABI: where the
 * reference calls abort() (DIE_IF after a failed PhysPressureFix,
 * rhllc_3d_stab.cpp:49-52), the step counts the cell in rhllc_stats.fix_fail
 * and carries on.  A handle owns its device memory and one HIP stream; all
 * calls on a handle come from one host thread.
 */
#ifndef RHLLC_HIP_H
#define RHLLC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RHLLC_OK 0
#define RHLLC_ERR_ARG 1     /* invalid argument or mesh */
#define RHLLC_ERR_HIP 2     /* HIP runtime failure (no device, out of memory, launch) */
#define RHLLC_ERR_STATE 3   /* call out of order (e.g. step before upload) */
#define RHLLC_ERR_NUMERIC 4 /* non-finite max signal speed */
#define RHLLC_ERR_C2P 5     /* rhllc_step: a PhysPressureFix failed in this step (the reference
                               aborts there, rhllc_3d_stab.cpp:49-52); the state was updated */

/* Arithmetic of the step. */
#define RHLLC_PREC_FP64 0  /* all FP64, bitwise equal to the reference compiled strict */
#define RHLLC_PREC_MIXED 1 /* FP32 storage of the primitives v and p (per-cell records, which
                              also hold s - D and D in FP64); face vectors S n in 16 bytes
                              (FP32 plus 7-bit residuals, exact to 2^-30); the two-kernel
                              step also stores its face flux in FP32 with small extensions;
                              FP64 U, 1/V and arithmetic; conserves to rounding */

/* Equation of state: the reference's compile-time EOS (include/hd_config.h:16-20). */
#define RHLLC_EOS_IDEAL 0 /* h = 1 + gamma/(gamma-1) p/rho, runtime gamma (EOS IDEAL) */
#define RHLLC_EOS_TAUB 1  /* Taub-Mathews h = 2.5 theta + sqrt(2.25 theta^2 + 1), theta = p/rho
                             (EOS TAUB, the reference's shipped setting); gamma unused */

/* Initial pressure of the c2p Newton iteration. */
#define RHLLC_GUESS_REFERENCE 0 /* max(|m|-E, 0, pmin) as rhllc_flux_stab.cpp:139-151 */
#define RHLLC_GUESS_PREVIOUS 1  /* the previous step's pressure */
#define RHLLC_GUESS_MLP 2       /* the MLP surrogate (rhllc_set_mlp) in cells whose tau = E - D
                                   changed by more than 5% in the step (the previous pressure
                                   elsewhere); then the previous pressure, the reference start
                                   and PhysPressureFix as fallbacks */

typedef struct rhllc_params {
  double gamma;       /* adiabatic index of the ideal-gas EOS (reference: 5/3 constexpr) */
  double cfl;         /* hllc_cfg.json "CFL" */
  double h_min;       /* hllc_cfg.json "h_min" */
  int32_t precision;  /* RHLLC_PREC_* */
  int32_t c2p_guess;  /* RHLLC_GUESS_*; must be REFERENCE for RHLLC_PREC_FP64 parity */
  int32_t reorder;    /* 1: relabel cells/faces in Hilbert order on the device (results are
                         returned in the caller's numbering either way) */
  int32_t device;     /* HIP device ordinal */
  int32_t collect_stats; /* 1: Newton histogram + iteration totals each step */
  int32_t has_out_source; /* ghost state for boundary code -3 (e_bound_out_source) */
  double out_source_W[5], out_source_U[5];
  int32_t eos;            /* RHLLC_EOS_*; 0 (zero-initialised params) is the ideal gas.
                             RHLLC_GUESS_MLP needs RHLLC_EOS_IDEAL */
  int32_t schedule;       /* RHLLC_SCHED_*; 0 (zero-initialised params) is the tiled step */
  /* The boundary branches of GEOMETRY_TYPE builds other than the default
   * Cone (rhllc_utils.cpp:58-110, include/solvers_config.h:19-29); all zero
   * is Cone: */
  int32_t lock_wall;        /* 1: boundary code -2 (e_bound_lock) mirrors the normal velocity,
                               as every type but Cone does; 0: it copies the cell */
  int32_t has_inner_source; /* 1: code -4 (e_bound_inner_source) takes the fixed ghost state
                               below (Sphere, Cone_JET); 0: it copies the cell */
  double inner_source_W[5], inner_source_U[5];
  int32_t reject_outer_surface; /* 1: a mesh with code -5 faces is RHLLC_ERR_ARG (every type but
                                   Cone aborts on them, D_LD at rhllc_utils.cpp:121-129) */
  int32_t out_source_copy;      /* 1 (with has_out_source 0): code -3 copies the cell, as every
                                   type but Cone and Cylinder does; 0: a mesh with code -3 faces
                                   needs the out_source state */
} rhllc_params;

/* Kernel schedule of the step; the results are the same bits either way on
 * the FP64 path.  The mixed path differs between the two in the face flux
 * only: the tiled step keeps it in FP64, the two-kernel step stores it packed
 * (FP32 with 8-bit extensions). */
#define RHLLC_SCHED_TILED 0      /* one kernel: each workgroup computes the faces of a tile of 256
                                    cells (Hilbert order) into LDS, then updates the tile's cells
                                    from there; faces shared with another tile are computed by
                                    both; the state the faces read is double-buffered */
#define RHLLC_SCHED_TWO_KERNEL 1 /* a face kernel writes every face flux to device memory, a cell
                                    kernel gathers them (the reference's face loop / cell loop) */

#define RHLLC_HIST_BINS 41 /* Newton iterations 0..40 */

typedef struct rhllc_stats {
  int64_t steps;
  int64_t c2p_fail[6];     /* cells whose first c2p returned code 1..5 (5: lor2 < 1) */
  int64_t fix_calls;       /* PhysPressureFix invocations */
  int64_t fix_fail;        /* PhysPressureFix failures (the reference aborts) */
  int64_t fallback_cells;  /* GUESS_MLP/PREVIOUS: cells that needed a fallback start */
  int64_t mlp_starts;      /* GUESS_MLP: cells started from the surrogate */
  int64_t newton_iters;    /* sum of Newton iterations (collect_stats) */
  int64_t newton_hist[RHLLC_HIST_BINS]; /* histogram of per-cell iterations (collect_stats) */
  double max_speed;        /* max signal speed of the last step (rhllc::max_signal_speed) */
  double tau;              /* tau of the next step: CFL*h_min/max_speed */
  double t;                /* simulated time accumulated by the device loop */
  double max_residual;     /* max |f(p)|/(E+p) over cells of the last step (collect_stats) */
} rhllc_stats;

typedef struct rhllc_solver rhllc_solver;

/* Upload the mesh (reference records) and allocate the state. */
int rhllc_create(const rhllc_params *params, int64_t n_cells, const void *geo_cells, int64_t n_faces,
                 const void *geo_faces, rhllc_solver **out);
int rhllc_destroy(rhllc_solver *h);

/* State transfer from/to host records.  recs points at n_cells records of
 * `stride` bytes; W (flux_t: rho, v, p) sits at off_phys, U (flux_t: D, m, E)
 * at off_conv.  Upload with off_conv < 0 computes U = p2c(W) on the device
 * (HllcPhysToConv, which clamps |v| >= 1 and so may rewrite W) and resets the
 * time loop as rhllc::Init does: max_speed = 1, tau = CFL*h_min, t = 0.
 * Download with off_phys or off_conv < 0 skips that part. */
int rhllc_upload_state(rhllc_solver *h, const void *recs, int64_t stride, int64_t off_phys, int64_t off_conv);
int rhllc_download_state(rhllc_solver *h, void *recs, int64_t stride, int64_t off_phys, int64_t off_conv);

/* One step at the given tau (rhllc::Hllc3dStab); *max_speed receives the new
 * rhllc::max_signal_speed.  Blocks until the step is done. */
int rhllc_step(rhllc_solver *h, double tau, double *max_speed);

/* nsteps steps with tau computed on the device between steps
 * (tau_{k+1} = CFL*h_min/max_speed_k, rhllc_utils.cpp:15-23), starting from
 * the device's current tau.  use_graph = 1 replays a captured hipGraph of
 * the step.  Asynchronous on the handle's stream; rhllc_sync waits. */
int rhllc_run(rhllc_solver *h, int32_t nsteps, int32_t use_graph);
int rhllc_sync(rhllc_solver *h);

/* Switch rhllc_params.collect_stats (Newton histogram, iteration total and
 * residual maximum of every step) on or off for the following steps; the
 * statistics kernels differ, so a captured graph is rebuilt.  No counterpart
 * in the reference. */
int rhllc_set_collect_stats(rhllc_solver *h, int32_t on);

/* Set the device's next tau and simulated time (e.g. from hllc_cfg.json). */
int rhllc_set_time(rhllc_solver *h, double tau, double t);
int rhllc_get_stats(rhllc_solver *h, rhllc_stats *out);
int rhllc_reset_stats(rhllc_solver *h);

/* Face fluxes of the last step in the caller's face numbering, n_faces x 5
 * doubles (S * F in the global frame, face_t::f). */
int rhllc_download_face_flux(rhllc_solver *h, double *F);

/* Radiation fields for the explicit 4-momentum source, applied after the
 * hydro update as radRHD_main.cpp:117-132 does (M1 Eddington tensor).  Per
 * cell in the caller's numbering, FP32.  NULL Ur removes the source.  With an
 * opacity table set (below) kappa_a and kappa_s are not used and may be NULL. */
int rhllc_set_radiation(rhllc_solver *h, const float *Ur, const float *Fr3, const float *kappa_a,
                        const float *kappa_s);

/* The thermal emission term of GetRadSource (radRHD_utils.cpp:89-90),
 * q = -kappa_a rho B_Plank(T) / kRadiation with T = GetTemperature(rho, p)
 * (lib/physics/include/gas_state.h:22-27) in the reference's CGS unit system
 * (lib/global_consts.h).  Off by default: with code-unit states of order one
 * (T ~ 1e13 K) the term dwarfs every other. */
int rhllc_set_radiation_emission(rhllc_solver *h, int32_t on);

/* Opacities from the reference's cooling-function table (TableFunc over
 * ln D x ln T, tab_func/rcf_H.dat; read with rhllc_table_read of
 * include/rhllc_mesh.h).  header = {min_x, max_x, step_x, min_y, max_y,
 * step_y}; data = nx*ny values of ln Lambda at [i*ny + j]; nx >= 1, ny >= 2.
 * With a table the source is the coupling path of the multi-rank step:
 * AddRadFlux (rhllc_3d_mpi.cpp:509-533) with GetRadSourceOpt
 * (radRHD_utils.cpp:103-158) and each cell's full_phys_data_t::Init
 * (lib/grids/src/solvers_struct.cpp:264-281) of the updated state, thermal
 * emission included.  NULL data removes the table; that is refused
 * (RHLLC_ERR_STATE) while the source is on without kappa_a/kappa_s. */
int rhllc_set_opacity_table(rhllc_solver *h, int32_t nx, int32_t ny, const double header[6], const double *data);

/* The table-opacity source on its own, for a caller that runs the transport
 * between the hydro update and the source as radRHD_mpi_main.cpp:112-122
 * does: AddRadFlux (rhllc_3d_mpi.cpp:509-533) with the moments Ur, Fr3 (per
 * cell, caller's numbering) and tau, then HllcConvToPhys
 * (rhllc_3d_mpi.cpp:488-505), on the handle's current state.  Needs an opacity
 * table; replaces the moments of rhllc_set_radiation and leaves rhllc_step's
 * own source on or off as it was (run the step without it: Ur = NULL there).
 * RHLLC_ERR_C2P when a PhysPressureFix failed. */
int rhllc_add_rad_flux(rhllc_solver *h, double tau, const float *Ur, const float *Fr3);

/* The radiation pressure tensor Tr of each cell (9 FP32 values, row-major
 * Tr(j, i) at [9 c + 3 j + i], caller's numbering), the illum module's
 * grid.impuls.  With it the source uses D(j, i) = Tr(j, i) instead of the M1
 * Eddington tensor, as the reference is compiled (radRHD_config.h:8,
 * ISOTROPIC_INTENSITY 0; radRHD_utils.cpp:57-63, 80-86, 117-123, 137-143),
 * including the product Tr(j, i) * Ur that it forms there.  NULL returns to
 * the M1 closure.  The fused source of rhllc_step / rhllc_run uses it. */
int rhllc_set_radiation_tensor(rhllc_solver *h, const float *Tr9);

/* rhllc_add_rad_flux with the moments' pressure tensor Tr (9 FP32 per cell,
 * as rhllc_set_radiation_tensor): AddRadFlux of the reference as compiled,
 * ISOTROPIC_INTENSITY 0 (radRHD_utils.cpp:137-143). */
int rhllc_add_rad_flux_tensor(rhllc_solver *h, double tau, const float *Ur, const float *Fr3, const float *Tr9);

/* Average device time (ms) of each kernel of the step, measured with HIP
 * events around each launch on the handle's stream over nsteps eager steps
 * (1..4096, after twenty untimed ones while the clock settles), enqueued
 * back to back with one wait at the end; the steps advance the state:
 * out[0] face kernel (tiled step: the tile kernel, faces and cells),
 * out[1] cell kernel and surrogate-list kernel (tiled step: the list kernel
 * only), out[2] dt kernel, out[3] whole step. */
int rhllc_time_kernels(rhllc_solver *h, int32_t nsteps, double out[4]);

/* The tiled step's layout of a mesh, computed on the host (no device is
 * touched), for tests and tools; no counterpart in the reference.  Returns
 * the tile count, the total of extra faces and the largest face count of a
 * tile; each non-NULL array receives: cell_new_to_old[n_cells] and
 * face_new_to_old[n_faces] (device numbering -> caller's), tile_meta
 * [4 n_tiles] (first own face, own count, first extra, extra count),
 * tile_extra[n_extra] (device face indices) and cell_slots[2 n_cells]
 * (four 16-bit tile-local slots per cell in slot order, bit 15 when the cell
 * is the face's id_r).  Call once with NULL arrays for the sizes. */
int rhllc_tile_layout(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces, int32_t reorder,
                      int32_t *n_tiles, int64_t *n_extra, int32_t *max_slots, int32_t *cell_new_to_old,
                      int32_t *face_new_to_old, int32_t *tile_meta, int32_t *tile_extra, uint32_t *cell_slots);

/* The mixed path's 16-byte form of n face vectors S n (3 doubles each) into
 * out (4 floats each), as rhllc_create stores them: the FP32 vector, and in
 * the fourth word (as bits) three signed 7-bit multiples k (bits 0-6,
 * 7-13, 14-20) of Q = 2^(E - 30), 2^E the power of two at or below the
 * largest |component| of the FP32 vector; fp32 + k Q is the vector to 2^-30
 * of its largest component.  Bits 21-31 are 0 here; the tiled step's layout
 * keeps each face's tile-local cell there.  No counterpart in the reference
 * (its faces carry FP64 n and S); for callers that inspect or prepare the
 * layout. */
int rhllc_pack_face_vectors(int64_t n, const double *Sn, float *out);

/* Sizes, and the algorithmic bytes of one step for roofline reporting:
 * the two-kernel structure's compulsory traffic at the variant's storage
 * precision (FP64 233 N + 120 F, mixed 161 N + 64 F, plus 24 N for the
 * radiation fields), whatever the schedule moves. */
int rhllc_info(rhllc_solver *h, int64_t *n_cells, int64_t *n_faces, int64_t *device_bytes);

/* Pointwise device evaluations over host arrays (n states, flux_t layout),
 * for per-function parity checks.  precision selects the FP64 (reference
 * arithmetic) or the mixed path's arithmetic of the flux and of c2p; eos is
 * RHLLC_EOS_* (gamma is used by RHLLC_EOS_IDEAL only); nS4 holds
 * (nx, ny, nz, S) per face. */
int rhllc_eval_flux(int32_t precision, int32_t eos, double gamma, int64_t n, const double *UL, const double *UR,
                    const double *WL, const double *WR, const double *nS4, double *F, double *speed);
int rhllc_eval_c2p(int32_t precision, int32_t eos, double gamma, int64_t n, const double *U, const double *guess,
                   double *W, int32_t *codes, int32_t *iters);
int rhllc_eval_p2c(int32_t eos, double gamma, int64_t n, double *W, double *U);
int rhllc_eval_pressure_fix(int32_t eos, double gamma, int64_t n, double *U, double *W, int32_t *codes);
int rhllc_eval_rad_source(int32_t precision, int64_t n, const double *W, const double *Ur, const double *Fr3,
                          const double *ka, const double *ks, int32_t emission, const float *Tr9, double *G);
/* GetRadSourceOpt with the table's opacities: G (4 per state) and the cell
 * data (T, vel, lorenz, alpha, betta: 5 per state).  For both: Tr9 (9 FP32
 * per state, row-major) selects the pressure-tensor branch, NULL the M1
 * closure. */
int rhllc_eval_rad_source_opt(int32_t nx, int32_t ny, const double header[6], const double *table, int64_t n,
                              const double *W, const double *Ur, const double *Fr3, const float *Tr9, double *G,
                              double *cell);

/* Calibration copy: reps launches of a plain vectorised copy kernel over
 * nbytes (read nbytes + write nbytes per launch); *gbps receives the mean
 * (read + write) bandwidth.  Used to scale rocprofv3 FETCH_SIZE/WRITE_SIZE
 * counters and as the measured HBM ceiling.  No counterpart in the reference. */
int rhllc_copy_probe(int32_t device, int64_t nbytes, int32_t reps, double *gbps);

/* ---- c2p surrogate (no counterpart in the reference) --------------------
 * A 3-64-64-1 SiLU MLP on log-scaled functions of (D, |m|, tau = E - D):
 * inputs (log D, log1p(|m|/D), log(p0/D)), output log(p/p0), where p0 is the
 * c2p equation's root with the Lorentz factor frozen at its p = 0 value,
 * W0 = E/sqrt(E^2 - m^2): p0 = (E - D W0)/(Gg W0^2 - 1), Gg = gamma/(gamma-1).
 * The model is RHLLC_MLP_MODEL floats: the parameters W1[64][3] b1[64]
 * W2[64][64] b2[64] W3[64] b3, then the normalisation mu_x[3] sd_x[3] mu_y
 * sd_y, then the gamma the model was trained for. */
#define RHLLC_MLP_PARAMS 4481
#define RHLLC_MLP_MODEL 4490

typedef struct rhllc_mlp_config {
  int32_t device;
  double gamma;          /* EOS index of the training states */
  int32_t steps;         /* Adam steps */
  int32_t batch;         /* samples per step, multiple of 256 */
  double lr;             /* peak learning rate (cosine schedule) */
  uint64_t seed;
  int32_t norm_samples;  /* samples used for the feature/target normalisation */
} rhllc_mlp_config;

/* Train on the device with the library's own forward/backward kernels (BF16
 * MFMA, FP32 accumulation, FP32 master weights, Adam).  Training states are
 * sampled log-uniformly (rho in [1e-4, 1e2], p/rho in [1e-6, 1e2], Lorentz
 * factor in [1, 50], isotropic velocity), converted by the FP64 p2c and
 * labelled by the FP64 c2p.  loss_hist (may be NULL) receives the mean
 * squared error of the normalised target per step. */
int rhllc_mlp_train(const rhllc_mlp_config *cfg, float *model_out, float *loss_hist);

/* Pressure predicted by the model for n conserved states (U: n x 5). */
int rhllc_mlp_eval(int32_t device, const float *model, int64_t n, const double *U, double *p_out);

/* Gradient of sum_i ok_i (f(x_i) - y_i)^2 / n over one batch of normalised
 * inputs X (n x 3), targets y and masks ok (n a multiple of 256), summed over
 * waves in order; *loss receives sum_i ok_i (f(x_i) - y_i)^2. */
int rhllc_mlp_grad(int32_t device, const float *model, int32_t n, const float *X, const float *y, const float *ok,
                   float *grad, float *loss);

/* rhllc_mlp_grad on device arrays, allocating nothing: model (RHLLC_MLP_MODEL
 * floats), X (n x 3), y, ok (n each), grad (RHLLC_MLP_PARAMS floats) and loss
 * (1 float) are device pointers; the work is enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) and the caller synchronizes.  The
 * packed weights, per-wave partials and losses live in a scratch array the
 * library holds on each GPU (18.4 MB, for n <= RHLLC_MLP_NOALLOC_MAX_N; larger
 * n is RHLLC_ERR_ARG); grad and loss are the same bits as rhllc_mlp_grad's.
 * Calls that overlap on one GPU share that scratch and race.  No counterpart
 * in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight training loops). */
#define RHLLC_MLP_NOALLOC_MAX_N 65536
int rhllc_mlp_grad_noalloc(int32_t device, const float *model, int32_t n, const float *X, const float *y,
                           const float *ok, float *grad, float *loss, void *stream);

/* Runs one v_mfma_f32_16x16x32_bf16 on known operands, with the operand
 * layout the MLP kernels use; out[64*4] = the accumulator registers per lane,
 * to check that layout on the device. */
int rhllc_mlp_selftest(int32_t device, float *out);

/* Installs a model for RHLLC_GUESS_MLP (mixed path); its gamma must be the
 * handle's (RHLLC_ERR_ARG otherwise). */
int rhllc_set_mlp(rhllc_solver *h, const float *model);

/* Device name and compute units of the current device (for reports). */
int rhllc_device_query(int32_t device, char *name, int32_t name_len, int32_t *cus, double *hbm_gb);

const char *rhllc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RHLLC_HIP_H */
