// Training and evaluation kernels of the c2p surrogate MLP (see mlp.h).
//
// Training step (captured in a hipGraph and replayed):
//   k_mlp_gen    seeded primitives -> FP64 p2c -> FP64 c2p (reference start)
//                -> normalised features/targets (failed states are masked);
//   k_mlp_grad   one wave per 64 samples: forward (BF16 MFMA with K = 32,
//                FP32 accumulate),
//                backward, per-wave gradient partials written to global memory;
//   k_mlp_adam   fixed-order sum of the partials (deterministic), Adam on the
//                FP32 master weights, BF16 copies of W2 and W2^T;
//   k_mlp_tick   advances the device step counter.
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "mlp_internal.h"
#include "rhllc_physics.h"

namespace rhllc_mlp {

using rhllc_dev::Ieee;
using rhllc_dev::State;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Primitive sample k of the training distribution (log-uniform rho in
// [1e-4, 1e2], p/rho in [1e-6, 1e2], Lorentz factor in [1, 50], isotropic v).
__device__ void sample_state(uint64_t seed, uint64_t k, State<double> &W) {
  const uint64_t b = mix64(seed) ^ (k * 0x9E3779B97F4A7C15ull);
  const double rho = exp(2.302585092994046 * (-4.0 + 6.0 * u01(mix64(b + 1))));
  const double pr = exp(2.302585092994046 * (-6.0 + 8.0 * u01(mix64(b + 2))));
  const double lor = exp(3.912023005428146 * u01(mix64(b + 3)));  // ln 50
  const double v = sqrt(1.0 - 1.0 / (lor * lor));
  const double z = 2.0 * u01(mix64(b + 4)) - 1.0;
  const double ph = 6.283185307179586 * u01(mix64(b + 5));
  const double s = sqrt(fmax(0.0, 1.0 - z * z));
  W.d = rho;
  W.v[0] = v * s * cos(ph), W.v[1] = v * s * sin(ph), W.v[2] = v * z;
  W.p = rho * pr;
}

// Raw (un-normalised) features and target of sample k; ok = 0 when c2p fails.
__device__ void make_sample(double gamma, uint64_t seed, uint64_t k, float x[3], float &y, int &ok) {
  State<double> W, U, Wc;
  sample_state(seed, k, W);
  const double Gg = gamma / (gamma - 1.0);
  rhllc_dev::p2c<double, Ieee>(Gg, W, U);
  int it = 0;
  const int code = rhllc_dev::c2p(Gg, U, Wc, -1.0, it);
  const double m = sqrt(rhllc_dev::dot3(U.v, U.v));
  double p0;
  raw_features(Gg, U.d, m, U.p, x, p0);
  ok = code == 0 && Wc.p > 0.0 && p0 > 0.0 && isfinite(x[0] + x[1] + x[2]);
  y = ok ? (float)log(Wc.p / p0) : 0.0f;
  if (!ok) x[0] = x[1] = x[2] = 0.0f;
}

__global__ void k_mlp_gen(double gamma, uint64_t seed, const uint32_t *step, int n, const float *norm, float *X,
                          float *Y, float *OK) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = (uint64_t)(step ? *step : 0u) * (uint64_t)n + (uint64_t)i;
  float x[3], y;
  int ok;
  make_sample(gamma, seed, k, x, y, ok);
  if (norm) {
    for (int f = 0; f < 3; ++f) x[f] = (x[f] - norm[f]) / norm[3 + f];
    y = (y - norm[6]) / norm[7];
  }
  X[3 * i] = x[0], X[3 * i + 1] = x[1], X[3 * i + 2] = x[2];
  Y[i] = ok ? y : 0.0f;
  OK[i] = ok ? 1.0f : 0.0f;
}

// One wave = 64 consecutive samples.  grad[wave][kParams] receives the
// wave's gradient of sum_s ok_s (out_s - y_s)^2 * inv_n; loss[wave] the sum
// of squared errors.
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ P, const short *__restrict__ W2b,
                                                  const short *__restrict__ W2Tb, const float *__restrict__ X,
                                                  const float *__restrict__ Y, const float *__restrict__ OK,
                                                  float inv_n, int n, float *__restrict__ grad,
                                                  float *__restrict__ loss) {
  __shared__ short lds[4][2][kH * kH];  // per wave: dZ2^T [h2][s] and H1^T [h1][s], bf16
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = l >> 4, c = l & 15;
  const int wave = blockIdx.x * 4 + w;
  const int base = wave * 64;
  // own sample (lane l) and the 4 samples s_t = 16 t + c this lane works on
  float xl[3] = {0.f, 0.f, 0.f}, yl = 0.f, okl = 0.f;
  if (base + l < n) {
    xl[0] = X[3 * (base + l)], xl[1] = X[3 * (base + l) + 1], xl[2] = X[3 * (base + l) + 2];
    yl = Y[base + l], okl = OK[base + l];
  }
  float x[4][3], y[4], ok[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int f = 0; f < 3; ++f) x[t][f] = __shfl(xl[f], 16 * t + c, 64);
    y[t] = __shfl(yl, 16 * t + c, 64);
    ok[t] = __shfl(okl, 16 * t + c, 64);
  }
  // ---- forward
  float z1[4][4][4];  // [q][t][e]
  bf16x4 h1[4][4];    // [q][t]
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int h = 16 * q + 4 * g + e;
      const float w0 = P[kOffW1 + 3 * h], w1 = P[kOffW1 + 3 * h + 1], w2 = P[kOffW1 + 3 * h + 2], bb = P[kOffB1 + h];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float z = ((w0 * x[t][0] + w1 * x[t][1]) + w2 * x[t][2]) + bb;
        z1[q][t][e] = z;
        h1[q][t][e] = f2bf(silu(z));
      }
    }
  float z2[4][4][4];  // [u][t][r] pre-activation of h2 = 16 u + 4 g + r, sample s_t
  float out[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    bf16x4 a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] = *reinterpret_cast<const bf16x4 *>(W2b + (16 * u + c) * kH + 16 * q + 4 * g);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 4; q += 2) acc = mfma32(a[q], a[q + 1], h1[q][t], h1[q + 1][t], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int h2 = 16 * u + 4 * g + r;
        z2[u][t][r] = acc[r] + P[kOffB2 + h2];
        out[t] += P[kOffW3 + h2] * silu(z2[u][t][r]);
      }
    }
  }
  float dout[4];
  float sq = 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    float v = out[t];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    const float d = (v + P[kOffB3] - y[t]) * ok[t];
    dout[t] = 2.0f * d * inv_n;
    sq += d * d;
  }
  // ---- layer 3 and b2, W3 gradients; dZ2
  bf16x4 dz2b[4][4];  // [u][t] as the B operand of dH1^T = W2^T dZ2^T
  float gW3[4][4], gB2[4][4];  // [u][r]
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int h2 = 16 * u + 4 * g + r;
      const float w3 = P[kOffW3 + h2];
      float aw = 0.f, ab = 0.f;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float z = z2[u][t][r];
        aw += dout[t] * silu(z);
        const float dz = dout[t] * w3 * dsilu(z);
        ab += dz;
        dz2b[u][t][r] = f2bf(dz);
      }
      gW3[u][r] = aw, gB2[u][r] = ab;
    }
  // ---- dH1^T = W2^T dZ2^T  (tiles q over h1, t over samples), then dZ1
  float gW1[4][4][3], gB1[4][4];  // [q][e]
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    bf16x4 a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = *reinterpret_cast<const bf16x4 *>(W2Tb + (16 * q + c) * kH + 16 * u + 4 * g);
#pragma unroll
    for (int e = 0; e < 4; ++e) gW1[q][e][0] = gW1[q][e][1] = gW1[q][e][2] = gB1[q][e] = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int u = 0; u < 4; u += 2) acc = mfma32(a[u], a[u + 1], dz2b[u][t], dz2b[u + 1][t], acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float dz = acc[e] * dsilu(z1[q][t][e]);
        gB1[q][e] += dz;
        gW1[q][e][0] += dz * x[t][0];
        gW1[q][e][1] += dz * x[t][1];
        gW1[q][e][2] += dz * x[t][2];
      }
    }
  }
  // ---- dW2 = dZ2^T H1 via an LDS transpose of both operands
  short *dzT = lds[w][0];
  short *h1T = lds[w][1];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) dzT[(16 * u + 4 * g + r) * kH + 16 * t + c] = dz2b[u][t][r];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) h1T[(16 * q + 4 * g + e) * kH + 16 * t + c] = h1[q][t][e];
  __syncthreads();
  float *G = grad + (size_t)wave * kParams;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t = 0; t < 4; t += 2) {
        const short *ar = dzT + (16 * u + c) * kH + 4 * g, *br = h1T + (16 * q + c) * kH + 4 * g;
        acc = mfma32(*reinterpret_cast<const bf16x4 *>(ar + 16 * t), *reinterpret_cast<const bf16x4 *>(ar + 16 * t + 16),
                     *reinterpret_cast<const bf16x4 *>(br + 16 * t), *reinterpret_cast<const bf16x4 *>(br + 16 * t + 16),
                     acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) G[kOffW2 + (16 * u + 4 * g + r) * kH + 16 * q + c] = acc[r];
    }
  // ---- reduce the per-lane partials over the 16 sample lanes (c) of each group g
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float a = gW3[u][r], b = gB2[u][r];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) a += __shfl_xor(a, o, 64), b += __shfl_xor(b, o, 64);
      if (c == 0) G[kOffW3 + 16 * u + 4 * g + r] = a, G[kOffB2 + 16 * u + 4 * g + r] = b;
    }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float v[4] = {gW1[q][e][0], gW1[q][e][1], gW1[q][e][2], gB1[q][e]};
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) v[k] += __shfl_xor(v[k], o, 64);
      const int h = 16 * q + 4 * g + e;
      if (c == 0) G[kOffW1 + 3 * h] = v[0], G[kOffW1 + 3 * h + 1] = v[1], G[kOffW1 + 3 * h + 2] = v[2], G[kOffB1 + h] = v[3];
    }
  // b3: sum of dout over the wave's samples (group g = 0 holds each sample once)
  float db = (g == 0) ? (dout[0] + dout[1] + dout[2] + dout[3]) : 0.f;
  float ls = (g == 0) ? sq : 0.f;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) db += __shfl_xor(db, o, 64), ls += __shfl_xor(ls, o, 64);
  if (l == 0) G[kOffB3] = db, loss[wave] = ls;
}

// Sum of the wave partials in wave order, then Adam (bias-corrected) with a
// cosine learning-rate schedule.
__global__ void k_mlp_adam(float *P, float *Mo, float *Vo, short *W2b, short *W2Tb, const float *grad, int nwaves,
                           const uint32_t *step, int total_steps, float lr0, const float *loss, float *loss_hist) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = *step;
  if (p < kParams) {
    float gsum = 0.f;
    for (int wv = 0; wv < nwaves; ++wv) gsum += grad[(size_t)wv * kParams + p];
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * Mo[p] + (1.f - b1) * gsum;
    const float v = b2 * Vo[p] + (1.f - b2) * gsum * gsum;
    Mo[p] = m, Vo[p] = v;
    const float tt = (float)(s + 1);
    const float mh = m / (1.f - powf(b1, tt)), vh = v / (1.f - powf(b2, tt));
    const float frac = (float)s / (float)(total_steps > 1 ? total_steps - 1 : 1);
    const float lr = lr0 * (0.02f + 0.98f * 0.5f * (1.f + cosf(3.14159265f * frac)));
    const float np = P[p] - lr * mh / (sqrtf(vh) + eps);
    P[p] = np;
    if (p >= kOffW2 && p < kOffB2) {
      const int i = (p - kOffW2) / kH, j = (p - kOffW2) % kH;
      W2b[i * kH + j] = f2bf(np);
      W2Tb[j * kH + i] = f2bf(np);
    }
  }
  if (p == 0 && loss_hist) {
    float t = 0.f;
    for (int wv = 0; wv < nwaves; ++wv) t += loss[wv];
    loss_hist[s] = t;
  }
}

__global__ void k_mlp_tick(uint32_t *step) { *step += 1u; }

// out[p] = sum over w < nw of part[w * np + p], added in w order in FP32 (the
// order of the host reduction of rhllc_mlp_grad).
__global__ void k_mlp_sum_partials(const float *part, int nw, int np, float *out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  float s = 0.f;
  for (int w = 0; w < nw; ++w) s += part[(size_t)w * np + p];
  out[p] = s;
}

__global__ void k_mlp_pack(const float *P, short *W2b, short *W2Tb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kH * kH) return;
  const int i = k / kH, j = k % kH;
  const short b = f2bf(P[kOffW2 + k]);
  W2b[i * kH + j] = b;
  W2Tb[j * kH + i] = b;
}

// Inference: p = p0 exp(out * sd_y + mu_y) for n conserved states (U: n x 5).
__global__ __launch_bounds__(256) void k_mlp_eval(const float *__restrict__ P, const short *__restrict__ W2b, int64_t n,
                                                  const double *__restrict__ U, double *__restrict__ p_out) {
  const int l = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float xl[3] = {0.f, 0.f, 0.f};
  const float *norm = P + kOffNorm;
  const double gamma = norm[8];
  double p0 = 0.0;
  if (i < n) {
    const double *u = U + 5 * i;
    const double m = sqrt((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3]);
    features(norm, gamma / (gamma - 1.0), u[0], m, u[4], xl, p0);
  }
  float x[4][3];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int f = 0; f < 3; ++f) x[t][f] = __shfl(xl[f], 16 * t + (l & 15), 64);
  float out[4];
  forward_wave(P, W2b, x, out);
  const float o = out[l >> 4];  // this lane's own sample 16 (l/16) + l%16 = l
  if (i < n) p_out[i] = predicted_pressure(norm, o, p0);
}

// MFMA lane-layout self test of mfma32 as the kernels use it: units k of two
// 16-unit blocks, lane group g holding k = 4 g + e (lo) and 16 + 4 g + e (hi).
// A[i][3] = i + 1 and A[i][23] = 0.5, B[3][j] = 16 (j + 1) and B[23][j] = 2, all
// other entries 0, so C[i][j] = 16 (i + 1)(j + 1) + 1 exactly; out[64*4] = C
// per lane (C[i][j] expected in lane j + 16 (i/4), element i%4).
__global__ void k_mlp_selftest(float *out) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4, c = l & 15;
  bf16x4 a[2], b[2];
  for (int h = 0; h < 2; ++h)
    for (int e = 0; e < 4; ++e) {
      const int k = 16 * h + 4 * g + e;
      a[h][e] = f2bf((k == 3 ? (float)(c + 1) : 0.0f) + (k == 23 ? 0.5f : 0.0f));
      b[h][e] = f2bf((k == 3 ? 16.0f * (float)(c + 1) : 0.0f) + (k == 23 ? 2.0f : 0.0f));
    }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma32(a[0], a[1], b[0], b[1], acc);
  for (int r = 0; r < 4; ++r) out[4 * l + r] = acc[r];
}

hipError_t launch_gen(double gamma, uint64_t seed, const uint32_t *step, int n, const float *norm, float *X, float *Y,
                      float *OK, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_gen, dim3((n + 255) / 256), dim3(256), 0, s, gamma, seed, step, n, norm, X, Y, OK);
  return hipGetLastError();
}
hipError_t launch_grad(const float *P, const short *W2b, const short *W2Tb, const float *X, const float *Y,
                       const float *OK, float inv_n, int n, float *grad, float *loss, hipStream_t s) {
  // n is a multiple of 256 (4 waves of 64 samples per workgroup)
  hipLaunchKernelGGL(k_mlp_grad, dim3(n / 256), dim3(256), 0, s, P, W2b, W2Tb, X, Y, OK, inv_n, n, grad, loss);
  return hipGetLastError();
}
hipError_t launch_adam(float *P, float *M, float *V, short *W2b, short *W2Tb, const float *grad, int nwaves,
                       const uint32_t *step, int total_steps, float lr0, const float *loss, float *loss_hist,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, P, M, V, W2b, W2Tb, grad, nwaves, step,
                     total_steps, lr0, loss, loss_hist);
  return hipGetLastError();
}
hipError_t launch_sum_partials(const float *part, int nw, int np, float *out, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_sum_partials, dim3((np + 255) / 256), dim3(256), 0, s, part, nw, np, out);
  return hipGetLastError();
}
hipError_t launch_tick(uint32_t *step, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(1), 0, s, step);
  return hipGetLastError();
}
hipError_t launch_pack(const float *P, short *W2b, short *W2Tb, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(kH * kH / 256), dim3(256), 0, s, P, W2b, W2Tb);
  return hipGetLastError();
}
hipError_t launch_eval(const float *P, const short *W2b, int64_t n, const double *U, double *p_out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, W2b, n, U, p_out);
  return hipGetLastError();
}
hipError_t launch_selftest(float *out, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_selftest, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

}  // namespace rhllc_mlp
