// C ABI of the c2p surrogate (include/rhllc_hip.h): training on the device,
// evaluation, one-batch gradients (for checks) and the MFMA layout self test.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/rhllc_hip.h"
#include "../mesh/mesh_common.h"
#include "mlp.h"
#include "mlp_internal.h"

using namespace rhllc_mlp;

extern "C" const char *rhllc_last_error(void);

namespace rhllc_capi_detail {
int set_error(int code, const std::string &msg);
}

namespace {

int fail(int code, const std::string &msg) { return rhllc_capi_detail::set_error(code, msg); }

#define HIP_TRY(expr)                                                                                          \
  do {                                                                                                         \
    hipError_t e_ = (expr);                                                                                    \
    if (e_ != hipSuccess) return fail(RHLLC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

struct Buf {
  void *p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

hipError_t alloc(Buf &b, size_t bytes) { return hipMalloc(&b.p, bytes + 64); }

}  // namespace

extern "C" int rhllc_mlp_selftest(int32_t device, float *out) {
  if (!out) return fail(RHLLC_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(device));
  Buf o;
  HIP_TRY(alloc(o, 256 * sizeof(float)));
  HIP_TRY(launch_selftest(o.as<float>(), nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(out, o.p, 256 * sizeof(float), hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_mlp_eval(int32_t device, const float *model, int64_t n, const double *U, double *p_out) {
  if (!model || n < 0 || (n > 0 && (!U || !p_out))) return fail(RHLLC_ERR_ARG, "bad argument");
  if (n == 0) return RHLLC_OK;
  HIP_TRY(hipSetDevice(device));
  Buf P, W2, Ud, pd;
  HIP_TRY(alloc(P, sizeof(float) * kModel));
  HIP_TRY(alloc(W2, sizeof(short) * kH * kH * 2));
  HIP_TRY(alloc(Ud, sizeof(double) * 5 * n));
  HIP_TRY(alloc(pd, sizeof(double) * n));
  HIP_TRY(hipMemcpy(P.p, model, sizeof(float) * kModel, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(Ud.p, U, sizeof(double) * 5 * n, hipMemcpyHostToDevice));
  HIP_TRY(launch_pack(P.as<float>(), W2.as<short>(), W2.as<short>() + kH * kH, nullptr));
  HIP_TRY(launch_eval(P.as<float>(), W2.as<short>(), n, Ud.as<double>(), pd.as<double>(), nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(p_out, pd.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RHLLC_OK;
}

extern "C" int rhllc_mlp_grad(int32_t device, const float *model, int32_t n, const float *X, const float *y,
                              const float *ok, float *grad, float *loss) {
  if (!model || !X || !y || !ok || !grad || !loss || n <= 0 || n % 256)
    return fail(RHLLC_ERR_ARG, "bad argument (n must be a positive multiple of 256)");
  HIP_TRY(hipSetDevice(device));
  const int nw = n / 64;
  Buf P, W2, Xd, Yd, Od, Gd, Ld;
  HIP_TRY(alloc(P, sizeof(float) * kModel));
  HIP_TRY(alloc(W2, sizeof(short) * kH * kH * 2));
  HIP_TRY(alloc(Xd, sizeof(float) * 3 * n));
  HIP_TRY(alloc(Yd, sizeof(float) * n));
  HIP_TRY(alloc(Od, sizeof(float) * n));
  HIP_TRY(alloc(Gd, sizeof(float) * kParams * nw));
  HIP_TRY(alloc(Ld, sizeof(float) * nw));
  HIP_TRY(hipMemcpy(P.p, model, sizeof(float) * kModel, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(Xd.p, X, sizeof(float) * 3 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(Yd.p, y, sizeof(float) * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(Od.p, ok, sizeof(float) * n, hipMemcpyHostToDevice));
  HIP_TRY(launch_pack(P.as<float>(), W2.as<short>(), W2.as<short>() + kH * kH, nullptr));
  HIP_TRY(launch_grad(P.as<float>(), W2.as<short>(), W2.as<short>() + kH * kH, Xd.as<float>(), Yd.as<float>(),
                      Od.as<float>(), 1.0f / (float)n, n, Gd.as<float>(), Ld.as<float>(), nullptr));
  HIP_TRY(hipDeviceSynchronize());
  std::vector<float> g((size_t)kParams * nw), l(nw);
  HIP_TRY(hipMemcpy(g.data(), Gd.p, sizeof(float) * g.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(l.data(), Ld.p, sizeof(float) * nw, hipMemcpyDeviceToHost));
  for (int p = 0; p < kParams; ++p) {
    float s = 0.f;
    for (int w = 0; w < nw; ++w) s += g[(size_t)w * kParams + p];
    grad[p] = s;
  }
  float s = 0.f;
  for (int w = 0; w < nw; ++w) s += l[w];
  *loss = s;
  return RHLLC_OK;
}

extern "C" int rhllc_mlp_train(const rhllc_mlp_config *cfg, float *model_out, float *loss_hist) {
  if (!cfg || !model_out) return fail(RHLLC_ERR_ARG, "null argument");
  if (cfg->batch <= 0 || cfg->batch % 256 || cfg->steps <= 0 || cfg->norm_samples < 1024 || !(cfg->gamma > 1.0) ||
      !(cfg->lr > 0.0))
    return fail(RHLLC_ERR_ARG, "bad training configuration (batch a multiple of 256, steps > 0, ...)");
  HIP_TRY(hipSetDevice(cfg->device));
  hipStream_t st;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};

  // Normalisation from a separate sample stream.
  const int nn = cfg->norm_samples;
  Buf Xn, Yn, On;
  HIP_TRY(alloc(Xn, sizeof(float) * 3 * nn));
  HIP_TRY(alloc(Yn, sizeof(float) * nn));
  HIP_TRY(alloc(On, sizeof(float) * nn));
  HIP_TRY(launch_gen(cfg->gamma, cfg->seed + 17, nullptr, nn, nullptr, Xn.as<float>(), Yn.as<float>(), On.as<float>(), st));
  HIP_TRY(hipStreamSynchronize(st));
  std::vector<float> xh(3 * (size_t)nn), yh(nn), oh(nn);
  HIP_TRY(hipMemcpy(xh.data(), Xn.p, sizeof(float) * xh.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(yh.data(), Yn.p, sizeof(float) * nn, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(oh.data(), On.p, sizeof(float) * nn, hipMemcpyDeviceToHost));
  double sum[4] = {0, 0, 0, 0}, sq[4] = {0, 0, 0, 0}, cnt = 0;
  for (int i = 0; i < nn; ++i) {
    if (oh[i] == 0.f) continue;
    const double v[4] = {xh[3 * i], xh[3 * i + 1], xh[3 * i + 2], yh[i]};
    for (int k = 0; k < 4; ++k) sum[k] += v[k], sq[k] += v[k] * v[k];
    cnt += 1;
  }
  if (cnt < 16) return fail(RHLLC_ERR_NUMERIC, "too few valid normalisation samples");
  std::vector<float> model(kModel, 0.f);
  for (int k = 0; k < 4; ++k) {
    const double mu = sum[k] / cnt;
    const double sd = std::sqrt(std::max(sq[k] / cnt - mu * mu, 1e-12));
    const int im = k < 3 ? kOffNorm + k : kOffNorm + 6, is = k < 3 ? kOffNorm + 3 + k : kOffNorm + 7;
    model[im] = (float)mu, model[is] = (float)sd;
  }
  model[kOffNorm + 8] = (float)cfg->gamma;
  // Glorot-uniform weights, zero biases (seeded).
  uint64_t rs = cfg->seed * 0x2545F4914F6CDD1Dull + 1;
  auto uni = [&](double a) {
    rs += 0x9E3779B97F4A7C15ull;
    return (float)((2.0 * rhllc_mesh::u01(rhllc_mesh::mix64(rs)) - 1.0) * a);
  };
  for (int i = 0; i < kH * kIn; ++i) model[kOffW1 + i] = uni(std::sqrt(6.0 / (kIn + kH)));
  for (int i = 0; i < kH * kH; ++i) model[kOffW2 + i] = uni(std::sqrt(6.0 / (kH + kH)));
  for (int i = 0; i < kH; ++i) model[kOffW3 + i] = uni(std::sqrt(6.0 / (kH + 1)));

  const int B = cfg->batch, nw = B / 64;
  Buf P, M, V, W2, X, Y, O, G, L, H, S;
  HIP_TRY(alloc(P, sizeof(float) * kModel));
  HIP_TRY(alloc(M, sizeof(float) * kParams));
  HIP_TRY(alloc(V, sizeof(float) * kParams));
  HIP_TRY(alloc(W2, sizeof(short) * kH * kH * 2));
  HIP_TRY(alloc(X, sizeof(float) * 3 * B));
  HIP_TRY(alloc(Y, sizeof(float) * B));
  HIP_TRY(alloc(O, sizeof(float) * B));
  HIP_TRY(alloc(G, sizeof(float) * kParams * nw));
  HIP_TRY(alloc(L, sizeof(float) * nw));
  HIP_TRY(alloc(H, sizeof(float) * cfg->steps));
  HIP_TRY(alloc(S, sizeof(uint32_t)));
  HIP_TRY(hipMemcpy(P.p, model.data(), sizeof(float) * kModel, hipMemcpyHostToDevice));
  std::vector<float> zeros(kParams, 0.f);
  HIP_TRY(hipMemcpy(M.p, zeros.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(V.p, zeros.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  const uint32_t zero = 0;
  HIP_TRY(hipMemcpy(S.p, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(launch_pack(P.as<float>(), W2.as<short>(), W2.as<short>() + kH * kH, st));
  const float *norm = P.as<float>() + kOffNorm;

  // One training step, captured once and replayed.
  HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_gen(cfg->gamma, cfg->seed, S.as<uint32_t>(), B, norm, X.as<float>(), Y.as<float>(),
                            O.as<float>(), st);
  if (e == hipSuccess)
    e = launch_grad(P.as<float>(), W2.as<short>(), W2.as<short>() + kH * kH, X.as<float>(), Y.as<float>(),
                    O.as<float>(), 1.0f / (float)B, B, G.as<float>(), L.as<float>(), st);
  if (e == hipSuccess)
    e = launch_adam(P.as<float>(), M.as<float>(), V.as<float>(), W2.as<short>(), W2.as<short>() + kH * kH,
                    G.as<float>(), nw, S.as<uint32_t>(), cfg->steps, (float)cfg->lr, L.as<float>(), H.as<float>(), st);
  if (e == hipSuccess) e = launch_tick(S.as<uint32_t>(), st);
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(st, &graph);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(RHLLC_ERR_HIP, std::string("training step capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(RHLLC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  for (int k = 0; k < cfg->steps && e == hipSuccess; ++k) e = hipGraphLaunch(exec, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipGraphExecDestroy(exec);
  if (e != hipSuccess) return fail(RHLLC_ERR_HIP, std::string("training: ") + hipGetErrorString(e));
  HIP_TRY(hipMemcpy(model_out, P.p, sizeof(float) * kModel, hipMemcpyDeviceToHost));
  if (loss_hist) {
    HIP_TRY(hipMemcpy(loss_hist, H.p, sizeof(float) * cfg->steps, hipMemcpyDeviceToHost));
    for (int k = 0; k < cfg->steps; ++k) loss_hist[k] /= (float)B;
  }
  return RHLLC_OK;
}

// ---------------------------------------------------------------- no-alloc
// Scratch of rhllc_mlp_grad_noalloc: the packed BF16 weights (2 x 64 x 64
// shorts = 4096 floats), the per-wave gradient partials (kParams floats a
// wave) and the per-wave losses, for batches of up to kNoAllocMaxN samples.
// Each GPU holds its own copy until the process ends.  Host code uses the
// device address hipGetSymbolAddress returns, never the array's name.
namespace {
constexpr int kNoAllocMaxN = 65536;
constexpr int kNoAllocMaxWaves = kNoAllocMaxN / 64;
constexpr size_t kScratchW2 = 0;                                              // floats
constexpr size_t kScratchGrad = kScratchW2 + (size_t)kH * kH;                 // 2 x 4096 shorts
constexpr size_t kScratchLoss = kScratchGrad + (size_t)kParams * kNoAllocMaxWaves;
constexpr size_t kScratchFloats = kScratchLoss + kNoAllocMaxWaves;
}  // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {
std::mutex g_scratch_mu;
std::vector<float *> g_scratch_ptr;  // per device; null until its first call

// The device address of `scratch` on the current device (looked up once per device).
int scratch_on(int device, float **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(RHLLC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<float *>(p);
  }
  *out = g_scratch_ptr[device];
  return RHLLC_OK;
}
}  // namespace

extern "C" int rhllc_mlp_grad_noalloc(int32_t device, const float *model, int32_t n, const float *X, const float *y,
                                      const float *ok, float *grad, float *loss, void *stream) {
  if (!model || !X || !y || !ok || !grad || !loss || n <= 0 || n % 256)
    return fail(RHLLC_ERR_ARG, "bad argument (n must be a positive multiple of 256)");
  if (n > kNoAllocMaxN)
    return fail(RHLLC_ERR_ARG, "n = " + std::to_string(n) + " exceeds the no-alloc scratch (at most " +
                                   std::to_string(kNoAllocMaxN) + " samples a call)");
  if (device < 0 || device > 4096) return fail(RHLLC_ERR_ARG, "bad device ordinal");
  HIP_TRY(hipSetDevice(device));
  float *sc = nullptr;
  if (int e = scratch_on(device, &sc)) return e;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const int nw = n / 64;
  short *W2b = reinterpret_cast<short *>(sc + kScratchW2);
  float *part = sc + kScratchGrad, *lpart = sc + kScratchLoss;
  // Every scratch element these kernels read is written first in the same call
  // (pack: all of W2; grad: nw partial rows and nw losses), so nothing is cleared.
  HIP_TRY(launch_pack(model, W2b, W2b + kH * kH, st));
  HIP_TRY(launch_grad(model, W2b, W2b + kH * kH, X, y, ok, 1.0f / (float)n, n, part, lpart, st));
  HIP_TRY(launch_sum_partials(part, nw, kParams, grad, st));
  HIP_TRY(launch_sum_partials(lpart, nw, 1, loss, st));
  return RHLLC_OK;
}
