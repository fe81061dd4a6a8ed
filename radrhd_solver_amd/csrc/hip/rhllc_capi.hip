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
  int variant = kFp64Ref;
  bool has_state = false;
  bool rad = false;
  bool rad_kappa = false;  // rhllc_set_radiation was given kappa_a / kappa_s
  hipStream_t stream = nullptr;
  Dev d{};
  std::vector<void *> allocs;
  std::vector<int32_t> cell_new_to_old, face_new_to_old;
  Scalars *host_sc = nullptr;  // pinned staging of the device scalars
  float4 *rad_FU = nullptr;
  float2 *rad_k = nullptr;
  double *tab = nullptr;  // opacity table (nx*ny), owned
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_src = nullptr;
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
    if (graph) (void)hipGraphExecDestroy(graph);
    if (graph_src) (void)hipGraphDestroy(graph_src);
    graph = nullptr;
    graph_src = nullptr;
  }

  ~rhllc_solver() {
    drop_graph();
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : allocs) (void)hipFree(p);
    if (rad_FU) (void)hipFree(rad_FU);
    if (rad_k) (void)hipFree(rad_k);
    if (tab) (void)hipFree(tab);
    if (mlp_P) (void)hipFree(mlp_P);
    if (mlp_W2b) (void)hipFree(mlp_W2b);
    if (host_sc) (void)hipHostFree(host_sc);
    if (stream) (void)hipStreamDestroy(stream);
  }

  bool ready() const { return has_state && (variant != kMixedMlp || mlp_P != nullptr); }

  // Cell kernel, plus the list kernel of the surrogate-start variant.
  hipError_t enqueue_cells() {
    hipError_t e = launch_cell(d, variant, rad, prm.collect_stats != 0, stream);
    if (e == hipSuccess && variant == kMixedMlp) e = launch_cell_mlp(d, rad, prm.collect_stats != 0, stream);
    return e;
  }
  hipError_t enqueue_step() {
    hipError_t e = launch_face(d, mixed, stream);
    if (e != hipSuccess) return e;
    e = enqueue_cells();
    if (e != hipSuccess) return e;
    return launch_dt(d, stream);
  }
};

extern "C" const char *rhllc_last_error(void) { return g_err.c_str(); }

extern "C" int rhllc_create(const rhllc_params *params, int64_t n_cells, const void *geo_cells, int64_t n_faces,
                            const void *geo_faces, rhllc_solver **out) {
  if (!params || !out || !geo_cells || !geo_faces) return fail(RHLLC_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_cells < 1 || n_cells >= ((int64_t)1 << 25))
    return fail(RHLLC_ERR_ARG, "n_cells must be in [1, 2^25) (geo_face_t::id_l is 25 bits)");
  if (n_faces < 1 || n_faces > 0x7FFFFFFF) return fail(RHLLC_ERR_ARG, "n_faces out of range");
  const rhllc_params &p = *params;
  if (p.eos != RHLLC_EOS_IDEAL && p.eos != RHLLC_EOS_TAUB) return fail(RHLLC_ERR_ARG, "unknown eos");
  if (p.eos == RHLLC_EOS_IDEAL && !(p.gamma > 1.0)) return fail(RHLLC_ERR_ARG, "gamma must be > 1");
  if (p.eos == RHLLC_EOS_TAUB && p.c2p_guess == RHLLC_GUESS_MLP)
    return fail(RHLLC_ERR_ARG, "the surrogate start is trained for the ideal gas; use RHLLC_EOS_IDEAL");
  if (!(p.cfl > 0.0) || !(p.h_min > 0.0)) return fail(RHLLC_ERR_ARG, "CFL and h_min must be > 0");
  if (p.precision != RHLLC_PREC_FP64 && p.precision != RHLLC_PREC_MIXED)
    return fail(RHLLC_ERR_ARG, "unknown precision");
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

  const int32_t N = (int32_t)n_cells, F = (int32_t)n_faces;
  const uint8_t *gc = static_cast<const uint8_t *>(geo_cells);
  const uint8_t *gf = static_cast<const uint8_t *>(geo_faces);

  // Decode and validate so every device gather stays in bounds.
  std::vector<int32_t> fl(F), fr(F);
  for (int32_t f = 0; f < F; ++f) {
    const uint8_t *r = gf + (size_t)f * RHLLC_GEO_FACE_BYTES;
    const int32_t l = (int32_t)rhllc_mesh::face_id_l(r);
    const int32_t rr = rhllc_mesh::face_id_r(r);
    if (l >= N) return fail(RHLLC_ERR_ARG, "face " + std::to_string(f) + ": id_l out of range");
    if (rr >= N) return fail(RHLLC_ERR_ARG, "face " + std::to_string(f) + ": id_r out of range");
    if (rr < RHLLC_BOUND_OUTER_SURFACE)
      return fail(RHLLC_ERR_ARG, "face " + std::to_string(f) + ": unknown boundary code " + std::to_string(rr));
    if (rr == RHLLC_BOUND_OUT_SOURCE && !p.has_out_source)
      return fail(RHLLC_ERR_ARG, "mesh has e_bound_out_source (-3) faces but no out_source state was given");
    fl[f] = l;
    fr[f] = rr;
  }
  for (int32_t c = 0; c < N; ++c) {
    const uint8_t *r = gc + (size_t)c * RHLLC_GEO_CELL_BYTES;
    int32_t ids[4];
    rhllc_mesh::cell_ids(r, ids);
    const uint8_t sign = r[48];
    for (int j = 0; j < 4; ++j) {
      if (ids[j] < 0 || ids[j] >= F) return fail(RHLLC_ERR_ARG, "cell " + std::to_string(c) + ": face id out of range");
      const bool is_left = (sign >> j) & 1;
      if (is_left ? fl[ids[j]] != c : fr[ids[j]] != c)
        return fail(RHLLC_ERR_ARG, "cell " + std::to_string(c) + ": slot " + std::to_string(j) +
                                       " disagrees with its face's id_l/id_r and sign bit");
    }
    if (!(rhllc_mesh::cell_V(r) > 0.0)) return fail(RHLLC_ERR_ARG, "cell " + std::to_string(c) + ": V <= 0");
  }

  std::unique_ptr<rhllc_solver> h(new rhllc_solver());
  h->prm = p;
  h->n_cells = N;
  h->n_faces = F;
  h->mixed = p.precision == RHLLC_PREC_MIXED;
  h->variant = variant;
  HIP_TRY(hipSetDevice(p.device));
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipHostMalloc((void **)&h->host_sc, sizeof(Scalars), hipHostMallocDefault));

  // Relabel (or keep) the numbering.
  h->cell_new_to_old.resize(N);
  h->face_new_to_old.resize(F);
  if (p.reorder) {
    rhllc_mesh::curve_order_impl(N, geo_cells, F, geo_faces, h->cell_new_to_old.data(), h->face_new_to_old.data());
  } else {
    for (int32_t i = 0; i < N; ++i) h->cell_new_to_old[i] = i;
    for (int32_t i = 0; i < F; ++i) h->face_new_to_old[i] = i;
  }
  std::vector<int32_t> cnew(N), fnew(F);
  for (int32_t i = 0; i < N; ++i) cnew[h->cell_new_to_old[i]] = i;
  for (int32_t i = 0; i < F; ++i) fnew[h->face_new_to_old[i]] = i;

  std::vector<int2> lr(F);
  std::vector<double4> nS(h->mixed ? 0 : F);
  std::vector<float3> Sn32(h->mixed ? F : 0);
  for (int32_t i = 0; i < F; ++i) {
    const int32_t f = h->face_new_to_old[i];
    const uint8_t *r = gf + (size_t)f * RHLLC_GEO_FACE_BYTES;
    double n[3], S;
    std::memcpy(n, r + 8, 24);
    std::memcpy(&S, r + 32, 8);
    const int32_t l = cnew[fl[f]];
    // Copy-type boundaries (Cone geometry: every code but -3) read the cell itself.
    const int32_t rr = fr[f] >= 0 ? cnew[fr[f]] : (fr[f] == RHLLC_BOUND_OUT_SOURCE ? -1 : l);
    lr[i] = make_int2(l, rr);
    if (h->mixed)
      Sn32[i] = make_float3((float)(n[0] * S), (float)(n[1] * S), (float)(n[2] * S));
    else
      nS[i] = make_double4(n[0], n[1], n[2], S);
  }
  std::vector<int4> cf(N);
  std::vector<double> V(N);
  for (int32_t i = 0; i < N; ++i) {
    const int32_t c = h->cell_new_to_old[i];
    const uint8_t *r = gc + (size_t)c * RHLLC_GEO_CELL_BYTES;
    int32_t ids[4], e[4];
    rhllc_mesh::cell_ids(r, ids);
    for (int j = 0; j < 4; ++j) e[j] = fnew[ids[j]] | (((r[48] >> j) & 1) ? 0 : (int32_t)0x80000000u);
    cf[i] = make_int4(e[0], e[1], e[2], e[3]);
    V[i] = rhllc_mesh::cell_V(r);
  }

  Dev &d = h->d;
  d.n_cells = N;
  d.n_faces = F;
  d.eos = p.eos;
  d.gamma = p.eos == RHLLC_EOS_IDEAL ? p.gamma : 0.0;
  d.Gg = p.eos == RHLLC_EOS_IDEAL ? p.gamma / (p.gamma - 1.0) : 0.0;
  d.cfl_hmin = p.cfl * p.h_min;
  for (int k = 0; k < 5; ++k) d.ghostW[k] = p.out_source_W[k], d.ghostU[k] = p.out_source_U[k];

  int2 *lr_d;
  int4 *cf_d;
  double *V_d;
  HIP_TRY(h->alloc(&lr_d, F));
  HIP_TRY(h->alloc(&cf_d, N));
  HIP_TRY(hipMemcpy(lr_d, lr.data(), sizeof(int2) * F, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(cf_d, cf.data(), sizeof(int4) * N, hipMemcpyHostToDevice));
  d.face_lr = lr_d;
  d.cell_faces = cf_d;
  if (!h->mixed) {
    HIP_TRY(h->alloc(&V_d, N));
    HIP_TRY(hipMemcpy(V_d, V.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    d.cell_V = V_d;
  } else {
    // (V, delta): delta = sum over the cell's faces of +-S n as the face
    // kernel sees it (FP32), summed in FP64; the cell kernel subtracts the
    // cell's own flux through it.
    std::vector<float4> Vd(N);
    for (int32_t i = 0; i < N; ++i) {
      const int ids[4] = {cf[i].x, cf[i].y, cf[i].z, cf[i].w};
      double dl[3] = {0.0, 0.0, 0.0};
      for (int j = 0; j < 4; ++j) {
        const float3 q = Sn32[ids[j] & 0x7FFFFFFF];
        const double sg = ids[j] >= 0 ? 1.0 : -1.0;
        dl[0] += sg * (double)q.x, dl[1] += sg * (double)q.y, dl[2] += sg * (double)q.z;
      }
      Vd[i] = make_float4((float)V[i], (float)dl[0], (float)dl[1], (float)dl[2]);
    }
    float4 *vd_d;
    HIP_TRY(h->alloc(&vd_d, N));
    HIP_TRY(hipMemcpy(vd_d, Vd.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
    d.cell_Vd = vd_d;
  }
  if (h->mixed) {
    float3 *q;
    HIP_TRY(h->alloc(&q, F));
    HIP_TRY(hipMemcpy(q, Sn32.data(), sizeof(float3) * F, hipMemcpyHostToDevice));
    d.face_Sn32 = q;
    HIP_TRY(h->alloc(&d.F4f, F));
    HIP_TRY(h->alloc(&d.FXf, F));
  } else {
    double4 *q;
    HIP_TRY(h->alloc(&q, F));
    HIP_TRY(hipMemcpy(q, nS.data(), sizeof(double4) * F, hipMemcpyHostToDevice));
    d.face_nS = q;
    HIP_TRY(h->alloc(&d.F4, F));
    HIP_TRY(h->alloc(&d.FE, F));
  }
  HIP_TRY(h->alloc(&d.U4, N));
  HIP_TRY(h->alloc(&d.UE, N));
  HIP_TRY(h->alloc(&d.W4, N));
  HIP_TRY(h->alloc(&d.Wp, N));
  if (h->mixed) {
    HIP_TRY(h->alloc(&d.Wva, N));
    HIP_TRY(h->alloc(&d.Wsd, N));
  }
  if (variant == kMixedMlp) HIP_TRY(h->alloc(&d.mlp_list, N));
  HIP_TRY(h->alloc(&d.sc, 1));
  HIP_TRY(h->alloc(&d.speed_slots, (size_t)kSpeedSlots * kSlotStride));
  std::memset(h->host_sc, 0, sizeof(Scalars));
  h->host_sc->tau = d.cfl_hmin / 1.0;
  h->host_sc->last_speed = 1.0;
  HIP_TRY(hipMemcpy(d.sc, h->host_sc, sizeof(Scalars), hipMemcpyHostToDevice));
  HIP_TRY(launch_clear_scalars(d, h->stream));
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
  const int32_t N = h->n_cells;
  const uint8_t *base = static_cast<const uint8_t *>(recs);
  std::vector<double4> w4(N), u4(off_conv >= 0 ? N : 0);
  std::vector<double> wp(N), ue(off_conv >= 0 ? N : 0);
  for (int32_t i = 0; i < N; ++i) {
    const uint8_t *r = base + (size_t)h->cell_new_to_old[i] * (size_t)stride;
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
  Dev &d = h->d;
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
  const Dev &d = h->d;
  uint8_t *base = static_cast<uint8_t *>(recs);
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
      std::memcpy(base + (size_t)h->cell_new_to_old[i] * stride + off_phys, w, 40);
    }
  }
  if (off_conv >= 0) {
    std::vector<double4> u4(N);
    std::vector<double> ue(N);
    HIP_TRY(hipMemcpy(u4.data(), d.U4, sizeof(double4) * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ue.data(), d.UE, sizeof(double) * N, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < N; ++i) {
      const double u[5] = {u4[i].x, u4[i].y, u4[i].z, u4[i].w, ue[i]};
      std::memcpy(base + (size_t)h->cell_new_to_old[i] * stride + off_conv, u, 40);
    }
  }
  return RHLLC_OK;
}

extern "C" int rhllc_step(rhllc_solver *h, double tau, double *max_speed) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  if (!h->ready()) return fail(RHLLC_ERR_STATE, "no state uploaded (or no MLP model set for RHLLC_GUESS_MLP)");
  HIP_TRY(hipSetDevice(h->prm.device));
  const unsigned long long fix_fail_before = h->host_sc->fix_fail;
  h->host_sc->tau = tau;
  HIP_TRY(hipMemcpyAsync(&h->d.sc->tau, &h->host_sc->tau, sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h->enqueue_step());
  HIP_TRY(hipMemcpyAsync(h->host_sc, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const double s = h->host_sc->last_speed;
  if (max_speed) *max_speed = s;
  if (!std::isfinite(s)) return fail(RHLLC_ERR_NUMERIC, "non-finite max signal speed");
  if (h->host_sc->fix_fail > fix_fail_before)
    return fail(RHLLC_ERR_C2P, std::to_string(h->host_sc->fix_fail - fix_fail_before) +
                                   " cells failed PhysPressureFix in this step");
  return RHLLC_OK;
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
  if (!h->graph) {
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = h->enqueue_step();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(e, "step capture");
    }
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    h->graph_src = g;
    HIP_TRY(hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
  }
  for (int32_t k = 0; k < nsteps; ++k) HIP_TRY(hipGraphLaunch(h->graph, h->stream));
  return RHLLC_OK;
}

extern "C" int rhllc_sync(rhllc_solver *h) {
  if (!h) return fail(RHLLC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
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
  HIP_TRY(launch_clear_scalars(h->d, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return RHLLC_OK;
}

extern "C" int rhllc_download_face_flux(rhllc_solver *h, double *F) {
  if (!h || !F) return fail(RHLLC_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(h->prm.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int32_t nf = h->n_faces;
  if (!h->mixed) {
    std::vector<double4> a(nf);
    std::vector<double> e(nf);
    HIP_TRY(hipMemcpy(a.data(), h->d.F4, sizeof(double4) * nf, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(e.data(), h->d.FE, sizeof(double) * nf, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < nf; ++i) {
      double *o = F + 5 * (size_t)h->face_new_to_old[i];
      o[0] = a[i].x, o[1] = a[i].y, o[2] = a[i].z, o[3] = a[i].w, o[4] = e[i];
    }
  } else {
    std::vector<float4> a(nf);
    std::vector<uint2> e(nf);
    HIP_TRY(hipMemcpy(a.data(), h->d.F4f, sizeof(float4) * nf, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(e.data(), h->d.FXf, sizeof(uint2) * nf, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < nf; ++i) rhllc_ext::unpack(a[i], e[i], F + 5 * (size_t)h->face_new_to_old[i]);
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

extern "C" int rhllc_add_rad_flux(rhllc_solver *h, double tau, const float *Ur, const float *Fr3) {
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
  if (!h->rad_FU) HIP_TRY(hipMalloc((void **)&h->rad_FU, sizeof(float4) * N));
  HIP_TRY(hipMemcpy(h->rad_FU, fu.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
  h->d.rad_FU = h->rad_FU;
  HIP_TRY(hipMemcpy(h->host_sc, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
  const unsigned long long fix_fail_before = h->host_sc->fix_fail;
  HIP_TRY(launch_add_rad(h->d, h->mixed, tau, h->stream));
  HIP_TRY(hipMemcpyAsync(h->host_sc, h->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->host_sc->fix_fail > fix_fail_before)
    return fail(RHLLC_ERR_C2P, std::to_string(h->host_sc->fix_fail - fix_fail_before) +
                                   " cells failed PhysPressureFix after the radiation source");
  return RHLLC_OK;
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
  if (nsteps < 1) return fail(RHLLC_ERR_ARG, "nsteps < 1");
  HIP_TRY(hipSetDevice(h->prm.device));
  hipEvent_t ev[4];
  for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  int rc = RHLLC_OK;
  for (int32_t k = 0; k < nsteps && rc == RHLLC_OK; ++k) {
    hipError_t e = hipEventRecord(ev[0], h->stream);
    if (e == hipSuccess) e = launch_face(h->d, h->mixed, h->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[1], h->stream);
    if (e == hipSuccess) e = h->enqueue_cells();
    if (e == hipSuccess) e = hipEventRecord(ev[2], h->stream);
    if (e == hipSuccess) e = launch_dt(h->d, h->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[3], h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[3]);
    float ms[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[0], ev[0], ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[1], ev[1], ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[2], ev[2], ev[3]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[3], ev[0], ev[3]);
    if (e != hipSuccess) {
      rc = hip_fail(e, "kernel timing");
      break;
    }
    for (int i = 0; i < 4; ++i) acc[i] += ms[i];
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
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
    // Algorithmic bytes of one step (both kernels), each array element counted
    // once per kernel that touches it.
    *device_bytes = h->mixed ? 180 * N + 68 * F : 233 * N + 120 * F;
    if (h->rad) *device_bytes += 24 * N;
  }
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

extern "C" int rhllc_eval_rad_source(int64_t n, const double *W, const double *Ur, const double *Fr3,
                                     const double *ka, const double *ks, int32_t emission, double *G) {
  if (n < 0 || !W || !Ur || !Fr3 || !ka || !ks || !G) return fail(RHLLC_ERR_ARG, "bad argument");
  DevBuf a, b, c, dd, e, f;
  HIP_TRY(to_dev(a, W, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(b, Ur, sizeof(double) * n));
  HIP_TRY(to_dev(c, Fr3, sizeof(double) * 3 * n));
  HIP_TRY(to_dev(dd, ka, sizeof(double) * n));
  HIP_TRY(to_dev(e, ks, sizeof(double) * n));
  HIP_TRY(to_dev(f, nullptr, sizeof(double) * 4 * n));
  HIP_TRY(launch_eval_rad(n, (double *)a.p, (double *)b.p, (double *)c.p, (double *)dd.p, (double *)e.p,
                          emission, (double *)f.p, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(G, f.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_eval_rad_source_opt(int32_t nx, int32_t ny, const double header[6], const double *table,
                                         int64_t n, const double *W, const double *Ur, const double *Fr3, double *G,
                                         double *cell) {
  if (n < 0 || !header || !table || !W || !Ur || !Fr3 || !G || !cell) return fail(RHLLC_ERR_ARG, "bad argument");
  if (int e = check_table(nx, ny, header)) return e;
  DevBuf t, a, b, c, g, o;
  HIP_TRY(to_dev(t, table, sizeof(double) * nx * ny));
  HIP_TRY(to_dev(a, W, sizeof(double) * 5 * n));
  HIP_TRY(to_dev(b, Ur, sizeof(double) * n));
  HIP_TRY(to_dev(c, Fr3, sizeof(double) * 3 * n));
  HIP_TRY(to_dev(g, nullptr, sizeof(double) * 4 * n));
  HIP_TRY(to_dev(o, nullptr, sizeof(double) * 5 * n));
  HIP_TRY(launch_eval_rad_opt(nx, ny, header, (double *)t.p, n, (double *)a.p, (double *)b.p, (double *)c.p,
                              (double *)g.p, (double *)o.p, nullptr));
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
