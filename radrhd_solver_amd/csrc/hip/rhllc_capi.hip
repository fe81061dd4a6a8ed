// C ABI of librhllc_hip.so (declarations and the reference interfaces each
// entry point replaces: include/rhllc_hip.h).
//
// A handle owns the mesh and state in device memory (SoA, Hilbert-relabelled
// when asked), one HIP stream, and optionally a captured hipGraph of the step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/rhllc_hip.h"
#include "../mesh/mesh_common.h"
#include "mlp_internal.h"
#include "rhllc_internal.h"
#include "flux_ext.h"
#include "rhllc_layout.h"

using namespace rhllc_dev;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char *what) {
  return fail(RHLLC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #expr);   \
  } while (0)

}  // namespace

namespace rhllc_capi_detail {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace rhllc_capi_detail

struct rhllc_solver {
  rhllc_params prm{};
  int32_t n_cells = 0, n_faces = 0;
  bool mixed = false;
  bool tiled = true;
  int variant = kFp64Ref;
  bool has_state = false;
  bool rad = false;
  bool rad_kappa = false;  // rhllc_set_radiation was given kappa_a / kappa_s
  hipStream_t stream = nullptr;
  Dev d{};
  // Buffers of the state.  [cur] holds the state at the start of the next
  // step.  The tiled step reads [cur] and writes [cur ^ 1] for whatever the
  // face work gathers (FP64: U, W; mixed: the cell records); arrays a cell
  // only updates in place have [1] == [0], and so has every array of the
  // two-kernel step.
  int cur = 0;
  int last_in = -1;  // buffer the last step read (the tiled step's face fluxes are recomputed from it)
  double4 *U4b[2] = {nullptr, nullptr};
  double *UEb[2] = {nullptr, nullptr};
  double4 *W4b[2] = {nullptr, nullptr};
  double *Wpb[2] = {nullptr, nullptr};
  float4 *Wvab[2] = {nullptr, nullptr};
  double2 *Wsgb[2] = {nullptr, nullptr};
  int32_t n_tiles = 0, tile_slots = 0;
  std::vector<void *> allocs;
  std::vector<int32_t> cell_new_to_old, face_new_to_old;
  Scalars *host_sc = nullptr;  // pinned staging of the device scalars
  unsigned long long fix_fail_seen = 0;  // fix_fail already reported by rhllc_step / rhllc_sync
  float4 *rad_FU = nullptr;
  float4 *rad_FU_add = nullptr;  // moments of rhllc_add_rad_flux
  float *rad_T = nullptr;        // pressure tensor of the fused source (rhllc_set_radiation_tensor)
  float *rad_T_add = nullptr;    // pressure tensor of rhllc_add_rad_flux_tensor
  float2 *rad_k = nullptr;
  double *tab = nullptr;  // opacity table (nx*ny), owned
  // Captured step graphs: the two-kernel step keeps one step in [0]; the
  // tiled step keeps two steps starting from buffer k in [k], so a replay
  // leaves the buffer parity as it was.
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipGraph_t graph_src[2] = {nullptr, nullptr};
  float *mlp_P = nullptr;
  short *mlp_W2b = nullptr;

  template <typename T>
  hipError_t alloc(T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 64);
    if (e == hipSuccess) allocs.push_back(q);
    *p = static_cast<T *>(q);
    return e;
  }

  void drop_graph() {
    for (int k = 0; k < 2; ++k) {
      if (graph[k]) (void)hipGraphExecDestroy(graph[k]);
      if (graph_src[k]) (void)hipGraphDestroy(graph_src[k]);
      graph[k] = nullptr;
      graph_src[k] = nullptr;
    }
  }

  ~rhllc_solver() {
    drop_graph();
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : allocs) (void)hipFree(p);
    if (rad_FU) (void)hipFree(rad_FU);
    if (rad_FU_add) (void)hipFree(rad_FU_add);
    if (rad_T) (void)hipFree(rad_T);
    if (rad_T_add) (void)hipFree(rad_T_add);
    if (rad_k) (void)hipFree(rad_k);
    if (tab) (void)hipFree(tab);
    if (mlp_P) (void)hipFree(mlp_P);
    if (mlp_W2b) (void)hipFree(mlp_W2b);
    if (host_sc) (void)hipHostFree(host_sc);
    if (stream) (void)hipStreamDestroy(stream);
  }

  bool ready() const { return has_state && (variant != kMixedMlp || mlp_P != nullptr); }

  // The kernel arguments with input buffers [in] and output buffers [out].
  Dev dev(int in, int out) const {
    Dev e = d;
    e.U4 = U4b[in], e.UE = UEb[in], e.W4 = W4b[in], e.Wp = Wpb[in], e.Wva = Wvab[in], e.Wsg = Wsgb[in];
    e.U4o = U4b[out], e.UEo = UEb[out], e.W4o = W4b[out], e.Wpo = Wpb[out], e.Wvao = Wvab[out], e.Wsgo = Wsgb[out];
    e.tile_reverse = in;
    return e;
  }
  // The current state, updated in place.
  Dev now() const { return dev(cur, cur); }

  // Cell kernel, plus the list kernel of the surrogate-start variant.
  hipError_t enqueue_cells(const Dev &e) {
    hipError_t r = launch_cell(e, variant, rad, prm.collect_stats != 0, stream);
    if (r == hipSuccess && variant == kMixedMlp) r = launch_cell_mlp(e, rad, prm.collect_stats != 0, stream);
    return r;
  }
  // One step from buffer cur; the tiled step flips cur.
  hipError_t enqueue_step() {
    hipError_t e;
    if (tiled) {
      const Dev x = dev(cur, cur ^ 1);
      e = launch_tile(x, mixed, variant, rad, prm.collect_stats != 0, n_tiles, tile_slots, stream);
      if (e == hipSuccess && variant == kMixedMlp) e = launch_cell_mlp(x, rad, prm.collect_stats != 0, stream);
      if (e == hipSuccess) e = launch_dt(x, stream);
      if (e == hipSuccess) {
        last_in = cur;
        cur ^= 1;
      }
      return e;
    }
    const Dev x = now();
    e = launch_face(x, mixed, stream);
    if (e == hipSuccess) e = enqueue_cells(x);
    if (e == hipSuccess) e = launch_dt(x, stream);
    if (e == hipSuccess) last_in = cur;
    return e;
  }
  // Captures graph[cur]: two steps (tiled) or one (two-kernel).
  hipError_t capture() {
    const int k = cur, saved_in = last_in;
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    for (int s = 0; s < (tiled ? 2 : 1) && e == hipSuccess; ++s) e = enqueue_step();
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(stream, &g);
    cur = k;  // capture enqueues nothing; replays set the parity
    last_in = saved_in;
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&graph[k], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      graph[k] = nullptr;
      return e;
    }
    graph_src[k] = g;
    return hipSuccess;
  }
};

extern "C" const char *rhllc_last_error(void) { return g_err.c_str(); }

extern "C" int rhllc_create(const rhllc_params *params, int64_t n_cells, const void *geo_cells, int64_t n_faces,
                            const void *geo_faces, rhllc_solver **out) {
  if (!params || !out || !geo_cells || !geo_faces) return fail(RHLLC_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_cells < 1 || n_cells >= ((int64_t)1 << 25))
    return fail(RHLLC_ERR_ARG, "n_cells must be in [1, 2^25) (geo_face_t::id_l is 25 bits)");
  // Faces below 2^27 (a mesh of 2^25 tets has at most 4 * 2^25 faces): the
  // kernels address the face arrays (at most 32 bytes an element) with 32-bit
  // byte offsets.
  if (n_faces < 1 || n_faces > ((int64_t)1 << 27)) return fail(RHLLC_ERR_ARG, "n_faces out of range [1, 2^27]");
  const rhllc_params &p = *params;
  if (p.eos != RHLLC_EOS_IDEAL && p.eos != RHLLC_EOS_TAUB) return fail(RHLLC_ERR_ARG, "unknown eos");
  if (p.eos == RHLLC_EOS_IDEAL && !(p.gamma > 1.0)) return fail(RHLLC_ERR_ARG, "gamma must be > 1");
  if (p.eos == RHLLC_EOS_TAUB && p.c2p_guess == RHLLC_GUESS_MLP)
    return fail(RHLLC_ERR_ARG, "the surrogate start is trained for the ideal gas; use RHLLC_EOS_IDEAL");
  if (!(p.cfl > 0.0) || !(p.h_min > 0.0)) return fail(RHLLC_ERR_ARG, "CFL and h_min must be > 0");
  if (p.precision != RHLLC_PREC_FP64 && p.precision != RHLLC_PREC_MIXED)
    return fail(RHLLC_ERR_ARG, "unknown precision");
  if (p.schedule != RHLLC_SCHED_TILED && p.schedule != RHLLC_SCHED_TWO_KERNEL)
    return fail(RHLLC_ERR_ARG, "unknown schedule");
  int variant;
  if (p.precision == RHLLC_PREC_FP64) {
    if (p.c2p_guess == RHLLC_GUESS_REFERENCE)
      variant = kFp64Ref;
    else if (p.c2p_guess == RHLLC_GUESS_PREVIOUS)
      variant = kFp64Prev;
    else
      return fail(RHLLC_ERR_ARG, "the FP64 path supports the reference and previous-pressure starts");
  } else {
    if (p.c2p_guess == RHLLC_GUESS_PREVIOUS)
      variant = kMixedPrev;
    else if (p.c2p_guess == RHLLC_GUESS_MLP)
      variant = kMixedMlp;
    else
      return fail(RHLLC_ERR_ARG, "the mixed path supports the previous-pressure and MLP starts");
  }
  const bool mixed = p.precision == RHLLC_PREC_MIXED, tiled = p.schedule == RHLLC_SCHED_TILED;

  // Host layout (validation, relabelling, SoA arrays, tiles) before the device is touched.
  rhllc_layout::Options opt;
  opt.mixed = mixed;
  opt.reorder = p.reorder != 0;
  opt.tiled = tiled;
  opt.has_out_source = p.has_out_source != 0;
  opt.has_inner_source = p.has_inner_source != 0;
  opt.lock_wall = p.lock_wall != 0;
  opt.reject_outer_surface = p.reject_outer_surface != 0;
  opt.out_source_copy = p.out_source_copy != 0;
  opt.tile_cells = kTileCells;
  rhllc_layout::Layout L;
  {
    std::string err;
    if (!rhllc_layout::build(opt, n_cells, geo_cells, n_faces, geo_faces, L, err)) return fail(RHLLC_ERR_ARG, err);
  }
  // The kernels address the face arrays with 32-bit byte offsets (32-byte
  // FP64 entries).
  if (n_faces + (int64_t)L.tile_extra.size() > ((int64_t)1 << 27))
    return fail(RHLLC_ERR_ARG, "faces plus the tiles' extra faces exceed 2^27; use RHLLC_SCHED_TWO_KERNEL");
  const int32_t N = (int32_t)n_cells, F = (int32_t)n_faces;

  std::unique_ptr<rhllc_solver> h(new rhllc_solver());
  h->prm = p;
  h->n_cells = N;
  h->n_faces = F;
  h->mixed = mixed;
  h->tiled = tiled;
  h->variant = variant;
  HIP_TRY(hipSetDevice(p.device));
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipHostMalloc((void **)&h->host_sc, sizeof(Scalars), hipHostMallocDefault));
  h->cell_new_to_old = std::move(L.cell_new_to_old);
  h->face_new_to_old = std::move(L.face_new_to_old);

  Dev &d = h->d;
  d.n_cells = N;
  d.n_faces = F;
  d.eos = p.eos;
  d.gamma = p.eos == RHLLC_EOS_IDEAL ? p.gamma : 0.0;
  d.Gg = p.eos == RHLLC_EOS_IDEAL ? p.gamma / (p.gamma - 1.0) : 0.0;
  d.cfl_hmin = p.cfl * p.h_min;
  for (int k = 0; k < 5; ++k) {
    d.ghostW[k] = p.out_source_W[k], d.ghostU[k] = p.out_source_U[k];
    d.ghost2W[k] = p.inner_source_W[k], d.ghost2U[k] = p.inner_source_U[k];
  }

  // Uploads count elements of a host array into a new device array (at least
  // one element is allocated; an empty array copies nothing).
  auto upload = [&](auto **dst, const void *src, size_t count) -> hipError_t {
    hipError_t e = h->alloc(dst, count ? count : 1);
    if (e == hipSuccess && count) e = hipMemcpy(*dst, src, sizeof(**dst) * count, hipMemcpyHostToDevice);
    return e;
  };
  // The face arrays, followed on the tiled step by the extra faces' copies
  // (the tiled mixed step reads face_far and the tile-local cells in face_Sn
  // in place of face_lr, which keeps the F faces for the flux download).
  const size_t nfe = (size_t)F + L.tile_extra.size();
  int2 *lr_d;
  HIP_TRY(upload(&lr_d, L.lr.data(), mixed && tiled ? (size_t)F : nfe));
  d.face_lr = lr_d;
  if (mixed && tiled) {
    int32_t *far_d;
    HIP_TRY(upload(&far_d, L.far.data(), nfe));
    d.face_far = far_d;
  }
  if (!mixed) {
    double *V_d;
    double4 *nS_d;
    HIP_TRY(upload(&V_d, L.V.data(), N));
    HIP_TRY(upload(&nS_d, L.nS.data(), nfe));
    d.cell_V = V_d;
    d.face_nS = nS_d;
  } else {
    double *vi_d;
    float4 *sn_d;
    HIP_TRY(upload(&vi_d, L.Vinv.data(), N));
    HIP_TRY(upload(&sn_d, L.Sn.data(), nfe));
    d.cell_Vinv = vi_d;
    d.face_Sn = sn_d;
  }
  if (tiled) {
    int4 *tm;
    uint2 *cs;
    HIP_TRY(upload(&tm, L.tile_meta.data(), L.n_tiles));
    HIP_TRY(upload(&cs, L.cell_slots.data(), N));
    d.tile_meta = tm;
    d.cell_slots = cs;
    h->n_tiles = L.n_tiles;
    h->tile_slots = L.max_slots > 0 ? L.max_slots : 1;
  } else {
    int4 *cf_d;
    HIP_TRY(upload(&cf_d, L.cf.data(), N));
    d.cell_faces = cf_d;
    if (mixed) {
      HIP_TRY(h->alloc(&d.F4f, F));
      HIP_TRY(h->alloc(&d.FXf, F));
    } else {
      HIP_TRY(h->alloc(&d.F4, F));
      HIP_TRY(h->alloc(&d.FE, F));
    }
  }
  // State buffers: double-buffered where the tiled step gathers them.
  // Two entries more than the cells: N and N + 1 hold the ghost states of
  // codes -3 and -4.
  const bool twoU = tiled && !mixed, twoRec = tiled && mixed;
  for (int k = 0; k < 2; ++k) {
    if (k == 0 || twoU) {
      HIP_TRY(h->alloc(&h->U4b[k], N + 2));
      HIP_TRY(h->alloc(&h->UEb[k], N + 2));
      HIP_TRY(h->alloc(&h->W4b[k], N + 2));
      HIP_TRY(h->alloc(&h->Wpb[k], N + 2));
    } else {
      h->U4b[k] = h->U4b[0], h->UEb[k] = h->UEb[0], h->W4b[k] = h->W4b[0], h->Wpb[k] = h->Wpb[0];
    }
    if (mixed && (k == 0 || twoRec)) {
      HIP_TRY(h->alloc(&h->Wvab[k], N + 2));
      HIP_TRY(h->alloc(&h->Wsgb[k], N + 2));
    } else {
      h->Wvab[k] = h->Wvab[0], h->Wsgb[k] = h->Wsgb[0];
    }
  }
  if (variant == kMixedMlp) HIP_TRY(h->alloc(&d.mlp_list, N));
  HIP_TRY(h->alloc(&d.sc, 1));
  HIP_TRY(h->alloc(&d.speed_slots, (size_t)kSpeedSlots * kSlotStride));
  std::memset(h->host_sc, 0, sizeof(Scalars));
  h->host_sc->tau = d.cfl_hmin / 1.0;
  h->host_sc->last_speed = 1.0;
  HIP_TRY(hipMemcpy(d.sc, h->host_sc, sizeof(Scalars), hipMemcpyHostToDevice));
  HIP_TRY(launch_clear_scalars(h->now(), h->stream));
  HIP_TRY(launch_ghost_cell(h->dev(0, 1), mixed, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *out = h.release();
  return RHLLC_OK;
}

extern "C" int rhllc_destroy(rhllc_solver *h) {
  delete h;
  return RHLLC_OK;
}

extern "C" int rhllc_upload_state(rhllc_solver *h, const void *recs, int64_t stride, int64_t off_phys,
                                  int64_t off_conv) {
  if (!h || !recs) return fail(RHLLC_ERR_ARG, "null argument");
  if (off_phys < 0 || stride < 40 || off_phys + 40 > stride || (off_conv >= 0 && off_conv + 40 > stride))
    return fail(RHLLC_ERR_ARG, "bad record stride/offsets");
  HIP_TRY(hipSetDevice(h->prm.device));
  // The copies below are synchronous with the host but not with the
  // handle's (non-blocking) stream: finish what rhllc_run left in flight.
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int32_t N = h->n_cells;
  const uint8_t *base = static_cast<const uint8_t *>(recs);
  std::vector<double4> w4(N), u4(off_conv >= 0 ? N : 0);
  std::vector<double> wp(N), ue(off_conv >= 0 ? N : 0);
  const int32_t *order = h->cell_new_to_old.data();
  for (int32_t i = 0; i < N; ++i) {
    const uint8_t *r = base + (size_t)order[i] * (size_t)stride;
    double w[5];
    std::memcpy(w, r + off_phys, 40);
    w4[i] = make_double4(w[0], w[1], w[2], w[3]);
    wp[i] = w[4];
    if (off_conv >= 0) {
      double u[5];
      std::memcpy(u, r + off_conv, 40);
      u4[i] = make_double4(u[0], u[1], u[2], u[3]);
      ue[i] = u[4];
    }
  }
  const Dev d = h->now();
  HIP_TRY(hipMemcpy(d.W4, w4.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d.Wp, wp.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  if (off_conv >= 0) {
    HIP_TRY(hipMemcpy(d.U4, u4.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.UE, ue.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    if (h->mixed) HIP_TRY(launch_seed_mixed(d, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  } else {
    // rhllc::Init (rhllc_init.cpp:190-192): HllcPhysToConv, max_signal_speed = 1,
    // tau = GetTimeStep.
    HIP_TRY(launch_p2c_init(d, h->stream));
    if (h->mixed) HIP_TRY(launch_seed_mixed(d, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(h->host_sc, d.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    h->host_sc->tau = d.cfl_hmin / 1.0;
    h->host_sc->t = 0.0;
    h->host_sc->last_speed = 1.0;
    HIP_TRY(hipMemcpy(d.sc, h->host_sc, sizeof(Scalars), hipMemcpyHostToDevice));
  }
  h->has_state = true;
  h->last_in = -1;
  return RHLLC_OK;
}

extern "C" int rhllc_download_state(rhllc_solver *h, void *recs, int64_t stride, int64_t off_phys,
                                    int64_t off_conv) {
  if (!h || !recs) return fail(RHLLC_ERR_ARG, "null argument");
  if (stride < 40 || (off_phys >= 0 && off_phys + 40 > stride) || (off_conv >= 0 && off_conv + 40 > stride))
    return fail(RHLLC_ERR_ARG, "bad record stride/offsets");
  if (!h->has_state) return fail(RHLLC_ERR_STATE, "no state uploaded");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int32_t N = h->n_cells;
  const Dev d = h->now();
  uint8_t *base = static_cast<uint8_t *>(recs);
  const int32_t *order = h->cell_new_to_old.data();
  if (off_phys >= 0) {
    if (h->mixed) {
      // The mixed step keeps W in FP32; rebuild FP64 W from the FP64 U.
      HIP_TRY(launch_recover_W(d, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
    }
    std::vector<double> wp(N);
    std::vector<double4> w4(N);
    HIP_TRY(hipMemcpy(wp.data(), d.Wp, sizeof(double) * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(w4.data(), d.W4, sizeof(double4) * N, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < N; ++i) {
      const double w[5] = {w4[i].x, w4[i].y, w4[i].z, w4[i].w, wp[i]};
      std::memcpy(base + (size_t)order[i] * stride + off_phys, w, 40);
    }
  }
  if (off_conv >= 0) {
    std::vector<double4> u4(N);
    std::vector<double> ue(N);
    HIP_TRY(hipMemcpy(u4.data(), d.U4, sizeof(double4) * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ue.data(), d.UE, sizeof(double) * N, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < N; ++i) {
      const double u[5] = {u4[i].x, u4[i].y, u4[i].z, u4[i].w, ue[i]};
      std::memcpy(base + (size_t)order[i] * stride + off_conv, u, 40);
    }
  }
  return RHLLC_OK;
}

// PhysPressureFix failures and a non-finite speed since the last report.
static int report_step_errors(rhllc_solver *h) {
  const unsigned long long ff = h->host_sc->fix_fail;
  const unsigned long long before = h->fix_fail_seen;
  h->fix_fail_seen = ff;
  if (ff > before)
    return fail(RHLLC_ERR_C2P, std::to_string(ff - before) + " cells failed PhysPressureFix");
  if (!std::isfinite(h->host_sc->last_speed) || !std::isfinite(h->host_sc->tau))
    return fail(RHLLC_ERR_NUMERIC, "non-finite max signal speed");
  return RHLLC_OK;
}

extern "C" int rhllc_step(rhllc_solver *h, double tau, double *max_speed) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  if (!h->ready()) return fail(RHLLC_ERR_STATE, "no state uploaded (or no MLP model set for RHLLC_GUESS_MLP)");
  HIP_TRY(hipSetDevice(h->prm.device));
  h->host_sc->tau = tau;
  HIP_TRY(hipMemcpyAsync(&h->d.sc->tau, &h->host_sc->tau, sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h->enqueue_step());
  HIP_TRY(hipMemcpyAsync(h->host_sc, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const double s = h->host_sc->last_speed;
  if (max_speed) *max_speed = s;
  return report_step_errors(h);
}

extern "C" int rhllc_run(rhllc_solver *h, int32_t nsteps, int32_t use_graph) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  if (!h->ready()) return fail(RHLLC_ERR_STATE, "no state uploaded (or no MLP model set for RHLLC_GUESS_MLP)");
  if (nsteps < 0) return fail(RHLLC_ERR_ARG, "nsteps < 0");
  HIP_TRY(hipSetDevice(h->prm.device));
  if (!use_graph) {
    for (int32_t k = 0; k < nsteps; ++k) HIP_TRY(h->enqueue_step());
    return RHLLC_OK;
  }
  const int per = h->tiled ? 2 : 1;
  const int32_t reps = nsteps / per;
  if (reps > 0) {
    if (!h->graph[h->cur]) HIP_TRY(h->capture());
    for (int32_t k = 0; k < reps; ++k) HIP_TRY(hipGraphLaunch(h->graph[h->cur], h->stream));
    h->last_in = h->tiled ? h->cur ^ 1 : h->cur;
  }
  for (int32_t k = reps * per; k < nsteps; ++k) HIP_TRY(h->enqueue_step());
  return RHLLC_OK;
}

extern "C" int rhllc_sync(rhllc_solver *h) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(h->host_sc, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
  return report_step_errors(h);
}

extern "C" int rhllc_set_collect_stats(rhllc_solver *h, int32_t on) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if ((h->prm.collect_stats != 0) != (on != 0)) h->drop_graph();  // other kernels
  h->prm.collect_stats = on != 0;
  return RHLLC_OK;
}

extern "C" int rhllc_set_time(rhllc_solver *h, double tau, double t) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->host_sc->tau = tau;
  h->host_sc->t = t;
  HIP_TRY(hipMemcpy(&h->d.sc->tau, &h->host_sc->tau, 2 * sizeof(double), hipMemcpyHostToDevice));
  return RHLLC_OK;
}

extern "C" int rhllc_get_stats(rhllc_solver *h, rhllc_stats *out) {
  if (!h || !out) return fail(RHLLC_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  Scalars s;
  HIP_TRY(hipMemcpy(&s, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
  std::memset(out, 0, sizeof(*out));
  out->steps = (int64_t)s.steps;
  for (int k = 0; k < 6; ++k) out->c2p_fail[k] = (int64_t)s.c2p_fail[k];
  out->fix_calls = (int64_t)s.fix_calls;
  out->fix_fail = (int64_t)s.fix_fail;
  out->fallback_cells = (int64_t)s.fallback;
  out->mlp_starts = (int64_t)s.mlp_starts;
  out->newton_iters = (int64_t)s.newton_iters;
  for (int k = 0; k < RHLLC_HIST_BINS; ++k) out->newton_hist[k] = (int64_t)s.hist[k];
  out->max_speed = s.last_speed;
  out->tau = s.tau;
  out->t = s.t;
  double r;
  std::memcpy(&r, &s.max_residual_bits, 8);
  out->max_residual = r;
  return RHLLC_OK;
}

extern "C" int rhllc_reset_stats(rhllc_solver *h) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(launch_clear_scalars(h->now(), h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->fix_fail_seen = 0;
  return RHLLC_OK;
}

extern "C" int rhllc_download_face_flux(rhllc_solver *h, double *F) {
  if (!h || !F) return fail(RHLLC_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int32_t nf = h->n_faces;
  if (h->last_in < 0) {  // no step since the state was set
    std::memset(F, 0, sizeof(double) * 5 * (size_t)nf);
    return RHLLC_OK;
  }
  std::vector<double4> a(nf);
  std::vector<double> e(nf);
  if (h->tiled) {
    // The tiled step keeps its fluxes in LDS; they are recomputed, to the same
    // bits, from the state the last step read (its input buffer is intact).
    double4 *F4 = nullptr;
    double *FE = nullptr;
    HIP_TRY(hipMalloc((void **)&F4, sizeof(double4) * nf));
    hipError_t r = hipMalloc((void **)&FE, sizeof(double) * nf);
    if (r == hipSuccess) r = launch_face_fp64_out(h->dev(h->last_in, h->last_in), h->mixed, F4, FE, h->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(h->stream);
    if (r == hipSuccess) r = hipMemcpy(a.data(), F4, sizeof(double4) * nf, hipMemcpyDeviceToHost);
    if (r == hipSuccess) r = hipMemcpy(e.data(), FE, sizeof(double) * nf, hipMemcpyDeviceToHost);
    (void)hipFree(F4);
    if (FE) (void)hipFree(FE);
    HIP_TRY(r);
  } else if (!h->mixed) {
    HIP_TRY(hipMemcpy(a.data(), h->d.F4, sizeof(double4) * nf, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(e.data(), h->d.FE, sizeof(double) * nf, hipMemcpyDeviceToHost));
  } else {
    std::vector<float4> pa(nf);
    std::vector<uint2> pe(nf);
    HIP_TRY(hipMemcpy(pa.data(), h->d.F4f, sizeof(float4) * nf, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pe.data(), h->d.FXf, sizeof(uint2) * nf, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < nf; ++i) rhllc_ext::unpack(pa[i], pe[i], F + 5 * (size_t)h->face_new_to_old[i]);
    return RHLLC_OK;
  }
  for (int32_t i = 0; i < nf; ++i) {
    double *o = F + 5 * (size_t)h->face_new_to_old[i];
    o[0] = a[i].x, o[1] = a[i].y, o[2] = a[i].z, o[3] = a[i].w, o[4] = e[i];
  }
  return RHLLC_OK;
}

extern "C" int rhllc_set_radiation(rhllc_solver *h, const float *Ur, const float *Fr3, const float *kappa_a,
                                   const float *kappa_s) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->drop_graph();
  if (!Ur) {
    h->rad = false;
    return RHLLC_OK;
  }
  if (!Fr3) return fail(RHLLC_ERR_ARG, "null radiation array");
  if (!kappa_a != !kappa_s) return fail(RHLLC_ERR_ARG, "kappa_a and kappa_s must both be given or both be NULL");
  if (!kappa_a && !h->tab) return fail(RHLLC_ERR_ARG, "no opacities: give kappa_a/kappa_s or set an opacity table");
  const int32_t N = h->n_cells;
  std::vector<float4> fu(N);
  std::vector<float2> kk(N);
  for (int32_t i = 0; i < N; ++i) {
    const int32_t c = h->cell_new_to_old[i];
    fu[i] = make_float4(Fr3[3 * (size_t)c], Fr3[3 * (size_t)c + 1], Fr3[3 * (size_t)c + 2], Ur[c]);
    kk[i] = kappa_a ? make_float2(kappa_a[c], kappa_s[c]) : make_float2(0.f, 0.f);
  }
  if (!h->rad_FU) HIP_TRY(hipMalloc((void **)&h->rad_FU, sizeof(float4) * N));
  if (!h->rad_k) HIP_TRY(hipMalloc((void **)&h->rad_k, sizeof(float2) * N));
  HIP_TRY(hipMemcpy(h->rad_FU, fu.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->rad_k, kk.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  h->d.rad_FU = h->rad_FU;
  h->d.rad_k = h->rad_k;
  h->rad = true;
  h->rad_kappa = kappa_a != nullptr;
  return RHLLC_OK;
}

// Per-cell FP32 tensors (9 values, caller's numbering) into device order.
static int upload_tensor(rhllc_solver *h, const float *Tr9, float **dst) {
  const int32_t N = h->n_cells;
  std::vector<float> t(9 * (size_t)N);
  for (int32_t i = 0; i < N; ++i)
    std::memcpy(&t[9 * (size_t)i], Tr9 + 9 * (size_t)h->cell_new_to_old[i], 9 * sizeof(float));
  if (!*dst) HIP_TRY(hipMalloc((void **)dst, sizeof(float) * 9 * N));
  HIP_TRY(hipMemcpy(*dst, t.data(), sizeof(float) * 9 * N, hipMemcpyHostToDevice));
  return RHLLC_OK;
}

extern "C" int rhllc_set_radiation_tensor(rhllc_solver *h, const float *Tr9) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->drop_graph();  // the graph holds the kernel arguments by value
  if (!Tr9) {
    h->d.rad_T = nullptr;
    return RHLLC_OK;
  }
  if (int e = upload_tensor(h, Tr9, &h->rad_T)) return e;
  h->d.rad_T = h->rad_T;
  return RHLLC_OK;
}

static int add_rad_flux(rhllc_solver *h, double tau, const float *Ur, const float *Fr3, const float *Tr9) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  if (!Ur || !Fr3) return fail(RHLLC_ERR_ARG, "null radiation array");
  if (!(tau >= 0.0) || !std::isfinite(tau)) return fail(RHLLC_ERR_ARG, "tau must be finite and >= 0");
  if (!h->tab) return fail(RHLLC_ERR_STATE, "no opacity table set (rhllc_set_opacity_table)");
  if (!h->ready()) return fail(RHLLC_ERR_STATE, "no state uploaded");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int32_t N = h->n_cells;
  std::vector<float4> fu(N);
  for (int32_t i = 0; i < N; ++i) {
    const int32_t c = h->cell_new_to_old[i];
    fu[i] = make_float4(Fr3[3 * (size_t)c], Fr3[3 * (size_t)c + 1], Fr3[3 * (size_t)c + 2], Ur[c]);
  }
  // Its own buffers: the moments of rhllc_set_radiation (the step's fused
  // source, also held by a captured graph) stay as they were.
  if (!h->rad_FU_add) HIP_TRY(hipMalloc((void **)&h->rad_FU_add, sizeof(float4) * N));
  HIP_TRY(hipMemcpy(h->rad_FU_add, fu.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
  Dev x = h->now();
  x.rad_FU = h->rad_FU_add;
  x.rad_T = nullptr;
  if (Tr9) {
    if (int e = upload_tensor(h, Tr9, &h->rad_T_add)) return e;
    x.rad_T = h->rad_T_add;
  }
  HIP_TRY(hipMemcpy(h->host_sc, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
  const unsigned long long fix_fail_before = h->host_sc->fix_fail;
  HIP_TRY(launch_add_rad(x, h->mixed, tau, h->stream));
  HIP_TRY(hipMemcpyAsync(h->host_sc, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const unsigned long long now_failed = h->host_sc->fix_fail - fix_fail_before;
  h->fix_fail_seen += now_failed;  // reported here; earlier steps' failures stay for rhllc_sync
  if (now_failed)
    return fail(RHLLC_ERR_C2P, std::to_string(now_failed) + " cells failed PhysPressureFix after the radiation source");
  return RHLLC_OK;
}

extern "C" int rhllc_add_rad_flux(rhllc_solver *h, double tau, const float *Ur, const float *Fr3) {
  return add_rad_flux(h, tau, Ur, Fr3, nullptr);
}

extern "C" int rhllc_add_rad_flux_tensor(rhllc_solver *h, double tau, const float *Ur, const float *Fr3,
                                         const float *Tr9) {
  if (!Tr9) return fail(RHLLC_ERR_ARG, "null pressure tensor");
  return add_rad_flux(h, tau, Ur, Fr3, Tr9);
}

extern "C" int rhllc_set_radiation_emission(rhllc_solver *h, int32_t on) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->drop_graph();  // the graph holds the kernel arguments by value
  h->d.rad_emission = on != 0;
  return RHLLC_OK;
}

// Table axes as the lookup uses them: finite, min <= max, positive steps.
static int check_table(int32_t nx, int32_t ny, const double *hdr) {
  if (nx < 1 || ny < 2 || (int64_t)nx * ny > (int64_t)1 << 26) return fail(RHLLC_ERR_ARG, "bad table size");
  for (int k = 0; k < 6; ++k)
    if (!std::isfinite(hdr[k])) return fail(RHLLC_ERR_ARG, "non-finite table header");
  if (!(hdr[0] <= hdr[1]) || !(hdr[2] > 0) || !(hdr[3] <= hdr[4]) || !(hdr[5] > 0))
    return fail(RHLLC_ERR_ARG, "table header needs min <= max and step > 0 on both axes");
  return RHLLC_OK;
}

extern "C" int rhllc_set_opacity_table(rhllc_solver *h, int32_t nx, int32_t ny, const double header[6],
                                       const double *data) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->drop_graph();
  if (!data) {
    if (h->rad && !h->rad_kappa)
      return fail(RHLLC_ERR_STATE, "the radiation source has no other opacities: remove it (rhllc_set_radiation "
                                   "with Ur = NULL) or give kappa_a/kappa_s first");
    if (h->tab) HIP_TRY(hipFree(h->tab));
    h->tab = nullptr;
    h->d.tab = nullptr;
    return RHLLC_OK;
  }
  if (!header) return fail(RHLLC_ERR_ARG, "null table header");
  if (int e = check_table(nx, ny, header)) return e;
  double *t = nullptr;
  HIP_TRY(hipMalloc((void **)&t, sizeof(double) * nx * ny));
  const hipError_t e = hipMemcpy(t, data, sizeof(double) * nx * ny, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(t);
    HIP_TRY(e);
  }
  if (h->tab) HIP_TRY(hipFree(h->tab));
  h->tab = t;
  Dev &d = h->d;
  d.tab = t;
  d.tab_nx = nx, d.tab_ny = ny;
  d.tab_min_x = header[0], d.tab_max_x = header[1], d.tab_step_x = header[2];
  d.tab_min_y = header[3], d.tab_max_y = header[4], d.tab_step_y = header[5];
  return RHLLC_OK;
}

extern "C" int rhllc_set_mlp(rhllc_solver *h, const float *model) {
  if (!h || !model) return fail(RHLLC_ERR_ARG, "null argument");
  const double mg = model[RHLLC_MLP_MODEL - 1];
  if (!(std::fabs(mg - h->prm.gamma) <= 1e-6 * h->prm.gamma))
    return fail(RHLLC_ERR_ARG, "model was trained for gamma " + std::to_string(mg) + ", the handle uses " +
                                   std::to_string(h->prm.gamma));
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->drop_graph();
  if (!h->mlp_P) HIP_TRY(hipMalloc((void **)&h->mlp_P, sizeof(float) * RHLLC_MLP_MODEL));
  if (!h->mlp_W2b) HIP_TRY(hipMalloc((void **)&h->mlp_W2b, sizeof(short) * 64 * 64 * 2));
  HIP_TRY(hipMemcpy(h->mlp_P, model, sizeof(float) * RHLLC_MLP_MODEL, hipMemcpyHostToDevice));
  HIP_TRY(rhllc_mlp::launch_pack(h->mlp_P, h->mlp_W2b, h->mlp_W2b + 64 * 64, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->d.mlpP = h->mlp_P;
  h->d.mlpW2b = h->mlp_W2b;
  return RHLLC_OK;
}

extern "C" int rhllc_time_kernels(rhllc_solver *h, int32_t nsteps, double out[4]) {
  if (!h || !out) return fail(RHLLC_ERR_ARG, "null argument");
  if (!h->ready()) return fail(RHLLC_ERR_STATE, "no state uploaded (or no MLP model set)");
  if (nsteps < 1 || nsteps > 4096) return fail(RHLLC_ERR_ARG, "nsteps must be in 1..4096");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  // Twenty untimed lead-in steps, then nsteps steps with four events each, all
  // enqueued back to back: the device runs them without the idle gaps that a
  // host wait per step would leave, as it runs the graph-replayed steps.  After
  // an idle period the device clock takes some fifteen steps to settle (the
  // first kernel runs fast, the next few up to 30% slower than steady state).
  constexpr int kLead = 20;
  std::vector<hipEvent_t> ev(4 * (size_t)nsteps, nullptr);
  int rc = RHLLC_OK;
  for (auto &e : ev)
    if (hipError_t err = hipEventCreate(&e); err != hipSuccess) {
      rc = hip_fail(err, "kernel timing");
      break;
    }
  const bool stats = h->prm.collect_stats != 0;
  for (int32_t k = 0; k < kLead + nsteps && rc == RHLLC_OK; ++k) {
    hipEvent_t *e4 = k >= kLead ? &ev[4 * (size_t)(k - kLead)] : nullptr;
    const Dev x = h->tiled ? h->dev(h->cur, h->cur ^ 1) : h->now();
    hipError_t e = e4 ? hipEventRecord(e4[0], h->stream) : hipSuccess;
    if (e == hipSuccess)
      e = h->tiled ? launch_tile(x, h->mixed, h->variant, h->rad, stats, h->n_tiles, h->tile_slots, h->stream)
                   : launch_face(x, h->mixed, h->stream);
    if (e == hipSuccess && e4) e = hipEventRecord(e4[1], h->stream);
    if (e == hipSuccess) {
      if (!h->tiled)
        e = h->enqueue_cells(x);
      else if (h->variant == kMixedMlp)
        e = launch_cell_mlp(x, h->rad, stats, h->stream);
    }
    if (e == hipSuccess && e4) e = hipEventRecord(e4[2], h->stream);
    if (e == hipSuccess) e = launch_dt(x, h->stream);
    if (e == hipSuccess && e4) e = hipEventRecord(e4[3], h->stream);
    if (e != hipSuccess) {
      rc = hip_fail(e, "kernel timing");
      break;
    }
    h->last_in = h->cur;
    if (h->tiled) h->cur ^= 1;
  }
  double acc[4] = {0, 0, 0, 0};
  if (rc == RHLLC_OK) {
    if (hipError_t e = hipStreamSynchronize(h->stream); e != hipSuccess) rc = hip_fail(e, "kernel timing");
  }
  for (int32_t k = 0; k < nsteps && rc == RHLLC_OK; ++k) {
    const hipEvent_t *e4 = &ev[4 * (size_t)k];
    float ms[4] = {0, 0, 0, 0};
    hipError_t e = hipEventElapsedTime(&ms[0], e4[0], e4[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[1], e4[1], e4[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[2], e4[2], e4[3]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[3], e4[0], e4[3]);
    if (e != hipSuccess) rc = hip_fail(e, "kernel timing");
    for (int i = 0; i < 4; ++i) acc[i] += ms[i];
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc != RHLLC_OK) return rc;
  for (int i = 0; i < 4; ++i) out[i] = acc[i] / nsteps;
  return RHLLC_OK;
}

extern "C" int rhllc_info(rhllc_solver *h, int64_t *n_cells, int64_t *n_faces, int64_t *device_bytes) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  if (n_cells) *n_cells = h->n_cells;
  if (n_faces) *n_faces = h->n_faces;
  if (device_bytes) {
    const int64_t N = h->n_cells, F = h->n_faces;
    // The two-kernel structure's compulsory bytes at the variant's storage
    // precision: FP64 face 80 N + 80 F, cell 153 N + 40 F; mixed face
    // 20 N + 44 F, cell 141 N + 20 F; the radiation fields 24 N.
    *device_bytes = h->mixed ? 161 * N + 64 * F : 233 * N + 120 * F;
    if (h->rad) *device_bytes += 24 * N;
  }
  return RHLLC_OK;
}

extern "C" int rhllc_tile_layout(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
                                 int32_t reorder, int32_t *n_tiles, int64_t *n_extra, int32_t *max_slots,
                                 int32_t *cell_new_to_old, int32_t *face_new_to_old, int32_t *tile_meta,
                                 int32_t *tile_extra, uint32_t *cell_slots) {
  if (!geo_cells || !geo_faces || !n_tiles || !n_extra || !max_slots) return fail(RHLLC_ERR_ARG, "null argument");
  if (n_cells < 1 || n_cells >= ((int64_t)1 << 25) || n_faces < 1 || n_faces > 0x7FFFFFFF)
    return fail(RHLLC_ERR_ARG, "mesh size out of range");
  rhllc_layout::Options opt;
  opt.reorder = reorder != 0;
  opt.tiled = true;
  opt.has_out_source = true;
  opt.tile_cells = kTileCells;
  rhllc_layout::Layout L;
  std::string err;
  if (!rhllc_layout::build(opt, n_cells, geo_cells, n_faces, geo_faces, L, err)) return fail(RHLLC_ERR_ARG, err);
  *n_tiles = L.n_tiles;
  *n_extra = (int64_t)L.tile_extra.size();
  *max_slots = L.max_slots;
  if (cell_new_to_old) std::memcpy(cell_new_to_old, L.cell_new_to_old.data(), sizeof(int32_t) * n_cells);
  if (face_new_to_old) std::memcpy(face_new_to_old, L.face_new_to_old.data(), sizeof(int32_t) * n_faces);
  if (tile_meta) std::memcpy(tile_meta, L.tile_meta.data(), sizeof(int32_t) * L.tile_meta.size());
  if (tile_extra && !L.tile_extra.empty())
    std::memcpy(tile_extra, L.tile_extra.data(), sizeof(int32_t) * L.tile_extra.size());
  if (cell_slots) std::memcpy(cell_slots, L.cell_slots.data(), sizeof(uint32_t) * L.cell_slots.size());
  return RHLLC_OK;
}

extern "C" int rhllc_pack_face_vectors(int64_t n, const double *Sn, float *out) {
  if (n < 0) return fail(RHLLC_ERR_ARG, "n < 0");
  if (n > 0 && (!Sn || !out)) return fail(RHLLC_ERR_ARG, "null argument");
  for (int64_t i = 0; i < n; ++i) rhllc_layout::pack_sn(Sn + 3 * i, out + 4 * i);
  return RHLLC_OK;
}

// ------------------------------------------------------------ pointwise evals
namespace {
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};
hipError_t to_dev(DevBuf &b, const void *src, size_t bytes) {
  hipError_t e = hipMalloc(&b.p, bytes + 64);
  if (e == hipSuccess && src) e = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
  return e;
}
}  // namespace

namespace {
bool bad_eos(int32_t eos, double gamma) {
  return !(eos == RHLLC_EOS_TAUB || (eos == RHLLC_EOS_IDEAL && gamma > 1.0));
}
}  // namespace

extern "C" int rhllc_eval_flux(int32_t precision, int32_t eos, double gamma, int64_t n, const double *UL,
                               const double *UR, const double *WL, const double *WR, const double *nS4, double *F,
                               double *speed) {
  if (n < 0 || !UL || !UR || !WL || !WR || !nS4 || !F || !speed) return fail(RHLLC_ERR_ARG, "bad argument");
  if (bad_eos(eos, gamma)) return fail(RHLLC_ERR_ARG, "unknown eos, or gamma <= 1 for the ideal gas");
  DevBuf a, b, c, dd, e, f, g;
  const size_t s5 = sizeof(double) * 5 * n;
  HIP_TRY(to_dev(a, UL, s5));
  HIP_TRY(to_dev(b, UR, s5));
  HIP_TRY(to_dev(c, WL, s5));
  HIP_TRY(to_dev(dd, WR, s5));
  HIP_TRY(to_dev(e, nS4, sizeof(double) * 4 * n));
  HIP_TRY(to_dev(f, nullptr, s5));
  HIP_TRY(to_dev(g, nullptr, sizeof(double) * n));
  HIP_TRY(launch_eval_flux(precision == RHLLC_PREC_MIXED, eos, gamma, n, (double *)a.p, (double *)b.p, (double *)c.p,
                           (double *)dd.p, (double *)e.p, (double *)f.p, (double *)g.p, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(F, f.p, s5, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(speed, g.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_eval_c2p(int32_t precision, int32_t eos, double gamma, int64_t n, const double *U,
                              const double *guess, double *W, int32_t *codes, int32_t *iters) {
  if (n < 0 || !U || !W || !codes || !iters) return fail(RHLLC_ERR_ARG, "bad argument");
  if (bad_eos(eos, gamma)) return fail(RHLLC_ERR_ARG, "unknown eos, or gamma <= 1 for the ideal gas");
  if (precision != RHLLC_PREC_FP64 && precision != RHLLC_PREC_MIXED) return fail(RHLLC_ERR_ARG, "bad precision");
  DevBuf a, b, c, dd, e;
  HIP_TRY(to_dev(a, U, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(c, W, sizeof(double) * 5 * n));
  if (guess) HIP_TRY(to_dev(b, guess, sizeof(double) * n));
  HIP_TRY(to_dev(dd, nullptr, sizeof(int32_t) * n));
  HIP_TRY(to_dev(e, nullptr, sizeof(int32_t) * n));
  HIP_TRY(launch_eval_c2p(precision == RHLLC_PREC_MIXED, eos, gamma, n, (double *)a.p, (double *)b.p, (double *)c.p, (int32_t *)dd.p, (int32_t *)e.p,
                          nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(W, c.p, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(codes, dd.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(iters, e.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_eval_p2c(int32_t eos, double gamma, int64_t n, double *W, double *U) {
  if (n < 0 || !W || !U) return fail(RHLLC_ERR_ARG, "bad argument");
  if (bad_eos(eos, gamma)) return fail(RHLLC_ERR_ARG, "unknown eos, or gamma <= 1 for the ideal gas");
  DevBuf a, b;
  HIP_TRY(to_dev(a, W, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(b, nullptr, sizeof(double) * 5 * n));
  HIP_TRY(launch_eval_p2c(eos, gamma, n, (double *)a.p, (double *)b.p, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(W, a.p, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(U, b.p, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_eval_pressure_fix(int32_t eos, double gamma, int64_t n, double *U, double *W, int32_t *codes) {
  if (n < 0 || !U || !W || !codes) return fail(RHLLC_ERR_ARG, "bad argument");
  if (bad_eos(eos, gamma)) return fail(RHLLC_ERR_ARG, "unknown eos, or gamma <= 1 for the ideal gas");
  DevBuf a, b, c;
  HIP_TRY(to_dev(a, U, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(b, W, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(c, nullptr, sizeof(int32_t) * n));
  HIP_TRY(launch_eval_fix(eos, gamma, n, (double *)a.p, (double *)b.p, (int32_t *)c.p, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(U, a.p, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(W, b.p, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(codes, c.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_eval_rad_source(int32_t precision, int64_t n, const double *W, const double *Ur,
                                     const double *Fr3, const double *ka, const double *ks, int32_t emission,
                                     const float *Tr9, double *G) {
  if (n < 0 || !W || !Ur || !Fr3 || !ka || !ks || !G) return fail(RHLLC_ERR_ARG, "bad argument");
  if (precision != RHLLC_PREC_FP64 && precision != RHLLC_PREC_MIXED) return fail(RHLLC_ERR_ARG, "bad precision");
  DevBuf a, b, c, dd, e, f, t;
  if (Tr9) HIP_TRY(to_dev(t, Tr9, sizeof(float) * 9 * n));
  HIP_TRY(to_dev(a, W, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(b, Ur, sizeof(double) * n));
  HIP_TRY(to_dev(c, Fr3, sizeof(double) * 3 * n));
  HIP_TRY(to_dev(dd, ka, sizeof(double) * n));
  HIP_TRY(to_dev(e, ks, sizeof(double) * n));
  HIP_TRY(to_dev(f, nullptr, sizeof(double) * 4 * n));
  HIP_TRY(launch_eval_rad(precision == RHLLC_PREC_MIXED, n, (double *)a.p, (double *)b.p, (double *)c.p,
                          (double *)dd.p, (double *)e.p, emission, (const float *)t.p, (double *)f.p, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(G, f.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_eval_rad_source_opt(int32_t nx, int32_t ny, const double header[6], const double *table,
                                         int64_t n, const double *W, const double *Ur, const double *Fr3,
                                         const float *Tr9, double *G, double *cell) {
  if (n < 0 || !header || !table || !W || !Ur || !Fr3 || !G || !cell) return fail(RHLLC_ERR_ARG, "bad argument");
  if (int e = check_table(nx, ny, header)) return e;
  DevBuf t, a, b, c, g, o, tt;
  if (Tr9) HIP_TRY(to_dev(tt, Tr9, sizeof(float) * 9 * n));
  HIP_TRY(to_dev(t, table, sizeof(double) * nx * ny));
  HIP_TRY(to_dev(a, W, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(b, Ur, sizeof(double) * n));
  HIP_TRY(to_dev(c, Fr3, sizeof(double) * 3 * n));
  HIP_TRY(to_dev(g, nullptr, sizeof(double) * 4 * n));
  HIP_TRY(to_dev(o, nullptr, sizeof(double) * 5 * n));
  HIP_TRY(launch_eval_rad_opt(nx, ny, header, (double *)t.p, n, (double *)a.p, (double *)b.p, (double *)c.p,
                              (const float *)tt.p, (double *)g.p, (double *)o.p, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(G, g.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(cell, o.p, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_copy_probe(int32_t device, int64_t nbytes, int32_t reps, double *gbps) {
  if (nbytes < 16 || reps < 1 || !gbps) return fail(RHLLC_ERR_ARG, "bad argument");
  HIP_TRY(hipSetDevice(device));
  const int64_t n16 = nbytes / 16;
  DevBuf a, b;
  HIP_TRY(to_dev(a, nullptr, (size_t)n16 * 16));
  HIP_TRY(to_dev(b, nullptr, (size_t)n16 * 16));
  HIP_TRY(hipMemset(a.p, 0, (size_t)n16 * 16));
  HIP_TRY(launch_copy(a.p, b.p, n16, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, nullptr));
  for (int32_t r = 0; r < reps; ++r) HIP_TRY(launch_copy(a.p, b.p, n16, nullptr));
  HIP_TRY(hipEventRecord(e1, nullptr));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = 2.0 * (double)n16 * 16.0 * reps / (ms * 1e-3) / 1e9;
  return RHLLC_OK;
}

extern "C" int rhllc_device_query(int32_t device, char *name, int32_t name_len, int32_t *cus, double *hbm_gb) {
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (name && name_len > 0) {
    std::strncpy(name, prop.gcnArchName, (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cus) *cus = prop.multiProcessorCount;
  if (hbm_gb) *hbm_gb = (double)prop.totalGlobalMem / 1e9;
  return RHLLC_OK;
}
