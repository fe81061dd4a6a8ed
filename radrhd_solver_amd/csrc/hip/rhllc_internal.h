// Device layout and kernel launchers of the RHLLC step (shared by the kernel
// and C-ABI translation units of librhllc_hip.so).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhllc_dev {

constexpr int kBlock = 256;         // threads per workgroup of the face and cell kernels
constexpr int kFacesPerThread = 1;  // faces per k_face thread (2 measured slower: fewer waves to hide latency)
constexpr int kSpeedSlots = 1024;  // per-step max-speed partials (one atomic slot per block group)
constexpr int kSlotStride = 16;    // doubles between slots: one 128-byte line each
constexpr int kHistBins = 41;
constexpr int kTileCells = 256;     // cells per tile of the tiled step (one workgroup, one cell per thread)
constexpr int kTileMaxSlots = 4 * kTileCells;  // faces a tile can touch (4 per cell)

// Device-resident scalars of the time loop and the step counters.
struct Scalars {
  double tau;         // tau of the step about to run
  double t;           // simulated time
  double last_speed;  // max signal speed of the last finished step
  unsigned long long steps;
  unsigned long long c2p_fail[6];
  unsigned long long fix_calls, fix_fail, fallback, newton_iters;
  unsigned long long hist[kHistBins];
  unsigned long long max_residual_bits;
  unsigned long long mlp_starts;
  unsigned int mlp_count;  // cells appended to Dev::mlp_list in the current step
};

// Everything a kernel needs, passed by value.  Cells and faces are in the
// device (Hilbert) numbering.  SoA: each per-cell / per-face quantity is a
// 16/32-byte vector array plus a scalar array so that gathers are single
// wide loads.
struct Dev {
  int32_t n_cells, n_faces;
  int32_t eos;  // Eos (rhllc_physics.h): kIdeal or kTaub
  double gamma, Gg, cfl_hmin;
  // faces
  const int2 *face_lr;      // (left, right): right n_cells / n_cells + 1 = ghost states, ~left = wall
  const double4 *face_nS;   // (nx, ny, nz, S), FP64 path
  const float4 *face_Sn;    // S * (nx, ny, nz), mixed path (16-byte form, unpack_sn)
  const int32_t *face_far;  // tiled mixed path: the cell outside the entry's tile (tile_face_in)
  double4 *F4;              // flux (d, mx, my, mz) * S, FP64 path
  double *FE;               // flux E * S
  float4 *F4f;              // mixed path: flux * S, FP32 parts (flux_ext.h)
  uint2 *FXf;               // mixed path: bits of the E part, extension counts
  // cells
  const int4 *cell_faces;  // face index | 0x80000000 when the cell is id_r (sign bit 0)
  const double *cell_V;    // FP64 path
  const double *cell_Vinv;  // mixed path: 1/V
  // State at the start of the step (read by the face work and by each cell's
  // own update) and where the cell update writes it.  The two-kernel step
  // updates in place (out == in); the tiled step gathers neighbours' state
  // while other tiles already write theirs, so whatever the face work reads
  // there is double-buffered: U and W on the FP64 path, the cell records on
  // the mixed path (whose U is read by its own cell only and stays in place).
  double4 *U4;   // (D, mx, my, mz), FP64 on both paths (c2p amplifies errors in U)
  double *UE;    // E
  double4 *W4;   // (rho, vx, vy, vz), FP64 path
  double *Wp;    // p (both paths: the FP64 pressure)
  float4 *Wva;   // mixed path: (vx, vy, vz, p) in FP32 (mixed_record)
  double2 *Wsg;  // mixed path: (s - D, D) in FP64, s = rho h W^2
  double4 *U4o;
  double *UEo;
  double4 *W4o;
  double *Wpo;
  float4 *Wvao;
  double2 *Wsgo;
  // Tiled step: tile t owns cells [t kTileCells, (t + 1) kTileCells) and
  // computes every face of them: its own faces (those whose earlier cell is in
  // the tile, a contiguous range since faces are grouped by their earlier
  // cell) and the extra faces it shares with an earlier tile (recomputed, same
  // bits).  tile_meta = (first own face, own count, first extra, extra count);
  // cell_slots = four 16-bit tile-local face slots in slot order, bit 15 set
  // when the cell is the face's id_r (the flux is subtracted).
  // face_lr and face_nS / face_Sn hold n_faces + (extra faces of all tiles)
  // entries: extra face j of tile t is entry n_faces + first extra + j, a copy
  // of the face's own entry, so a tile reads its extra faces contiguously.
  const int4 *tile_meta;
  const uint2 *cell_slots;
  int32_t tile_reverse;  // walk the tiles last to first (steps from buffer 1)
  // radiation source (optional): (Fx, Fy, Fz, Ur) and (kappa_a, kappa_s)
  const float4 *rad_FU;
  const float2 *rad_k;
  // Radiation pressure tensor Tr per cell (9 FP32, row-major), the reference's
  // compiled ISOTROPIC_INTENSITY 0 branch; null: the M1 closure.
  const float *rad_T;
  int32_t rad_emission;  // thermal emission term of the source (0: off)
  // Opacity table (ln Lambda over ln D x ln T, TableFunc of the reference):
  // when set, each cell's opacities come from its state and rad_k is unused.
  const double *tab;
  int32_t tab_nx, tab_ny;
  double tab_min_x, tab_max_x, tab_step_x, tab_min_y, tab_max_y, tab_step_y;
  // Fixed ghost states for boundary codes -3 and -4.  The state arrays hold
  // them as two more cells, indices n_cells and n_cells + 1 (written by
  // launch_ghost_cell), which no step updates; faces of those codes gather
  // them as their right cell.
  double ghostW[5], ghostU[5], ghost2W[5], ghost2U[5];
  // surrogate c2p guess (RHLLC_GUESS_MLP): FP32 model and bf16 W2
  const float *mlpP;
  const short *mlpW2b;
  int32_t *mlp_list;  // kMixedMlp: cells deferred to k_cell_mlp (n_cells entries)
  Scalars *sc;
  double *speed_slots;  // kSpeedSlots * kSlotStride
};

enum Variant : int {
  kFp64Ref = 0,   // FP64, reference c2p start (bitwise parity)
  kFp64Prev = 1,  // FP64, previous-pressure start
  kMixedPrev = 2, // mixed, previous-pressure start
  kMixedMlp = 3,  // mixed, MLP start (previous pressure, then reference start as fallbacks)
};

hipError_t launch_p2c_init(const Dev &d, hipStream_t s);
hipError_t launch_face(const Dev &d, bool mixed, hipStream_t s);
hipError_t launch_cell(const Dev &d, int variant, bool rad, bool stats, hipStream_t s);
hipError_t launch_cell_mlp(const Dev &d, bool rad, bool stats, hipStream_t s);
// The tiled step: faces into LDS, then the cells of the tile from LDS.
// lds_slots = the largest face count of any tile.
hipError_t launch_tile(const Dev &d, bool mixed, int variant, bool rad, bool stats, int n_tiles, int lds_slots,
                       hipStream_t s);
// Face fluxes of the state in d's input buffers as the step computes them
// (FP64 S * F, energy slot = E flux), into F4/FE: for rhllc_download_face_flux.
hipError_t launch_face_fp64_out(const Dev &d, bool mixed, double4 *F4, double *FE, hipStream_t s);
hipError_t launch_dt(const Dev &d, hipStream_t s);
hipError_t launch_clear_scalars(const Dev &d, hipStream_t s);
// Mixed path: W (FP64) recovered from U by c2p (start from the FP32 pressure).
hipError_t launch_seed_mixed(const Dev &d, hipStream_t s);
// Writes the ghost states into entries n_cells and n_cells + 1 of the input and output buffers
// of d (U and W on the FP64 path, the cell record on the mixed path).
hipError_t launch_ghost_cell(const Dev &d, bool mixed, hipStream_t s);
hipError_t launch_recover_W(const Dev &d, hipStream_t s);
// The radiation source alone (d.rad_FU set): U += tau G(W), then c2p.
hipError_t launch_add_rad(const Dev &d, bool mixed, double tau, hipStream_t s);

hipError_t launch_copy(const void *src, void *dst, int64_t n16, hipStream_t s);

// Pointwise evaluations (n states) for parity tests; eos is an Eos value.
hipError_t launch_eval_flux(bool fp32, int eos, double gamma, int64_t n, const double *UL, const double *UR,
                            const double *WL, const double *WR, const double *nS4, double *F, double *speed,
                            hipStream_t s);
hipError_t launch_eval_c2p(bool mixed, int eos, double gamma, int64_t n, const double *U, const double *guess,
                           double *W, int32_t *codes, int32_t *iters, hipStream_t s);
hipError_t launch_eval_p2c(int eos, double gamma, int64_t n, double *W, double *U, hipStream_t s);
hipError_t launch_eval_fix(int eos, double gamma, int64_t n, double *U, double *W, int32_t *codes, hipStream_t s);
hipError_t launch_eval_rad(bool mixed, int64_t n, const double *W, const double *Ur, const double *Fr3,
                           const double *ka, const double *ks, int emission, const float *Tr9, double *G,
                           hipStream_t s);
// hdr = {min_x, max_x, step_x, min_y, max_y, step_y}; host pointer, read at launch.
hipError_t launch_eval_rad_opt(int nx, int ny, const double *hdr, const double *table, int64_t n, const double *W,
                               const double *Ur, const double *Fr3, const float *Tr9, double *G, double *cell,
                               hipStream_t s);

}  // namespace rhllc_dev
