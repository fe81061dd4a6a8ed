// MLP surrogate of the c2p pressure, 3 -> 64 -> 64 -> 1 with SiLU, no
// counterpart in the reference.
//
// Inputs are log-scaled functions of (D, |m|, tau = E - D):
//   x = (log D, log1p(|m|/D), log(p0/D)),   output log(p/p0),
// with p0 the pressure of the c2p equation E + p = D W + Gg p W^2 solved with
// the Lorentz factor frozen at its p = 0 value W0 = E / sqrt(E^2 - m^2).
// p >= p0 always (W falls as p grows) and p/p0 stays in [1, 5] over the
// training range.  log(tau/D) itself is not usable as an input: at W ~ 50 it
// carries the thermal energy as a 1e-4 relative difference against |m|/D,
// below BF16 resolution, while p0 is formed in FP64 before rounding.
//
// Parameter vector (FP32, kParams floats):
//   W1[64][3] b1[64] W2[64][64] b2[64] W3[64] b3[1]     (weights stored [out][in])
// followed by mu_x[3] sd_x[3] mu_y sd_y (normalisation) and the adiabatic
// index gamma the model was trained for.
//
// Wave-level layout.  A wave evaluates 64 samples with the 64x64 layer on
// gfx950's v_mfma_f32_16x16x32_bf16 (FP32 accumulate), two instructions per
// 16x16 output tile.  Its operand/result lane maps are
//   A[i][k]: lane i + 16*(k/8), element k%8
//   B[k][j]: lane j + 16*(k/8), element k%8
//   C[i][j]: lane j + 16*(i/4), element i%4
// The hidden units are held in 16-unit blocks q, four per lane (units
// 16 q + 4 (l/16) + e, e = 0..3, in lane l); one instruction contracts the
// blocks 2 q' and 2 q' + 1 by giving it, in both operands, the lane's four
// units of the first block as elements 0..3 and of the second as 4..7 (the
// sum over k does not depend on which unit sits at which k, as long as A and
// B agree; rhllc_mlp_selftest checks this layout on the device).  The forward
// product is Z2^T = W2 * H1^T, so lane l holds, for samples s_t = 16 t + l%16
// (t = 0..3), the hidden units 16 q + 4 (l/16) + e of both layers -- layer 1
// is computed directly in that layout and no transpose is needed until dW2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhllc_mlp {

constexpr int kH = 64;
constexpr int kIn = 3;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;  // 192
constexpr int kOffW2 = kOffB1 + kH;        // 256
constexpr int kOffB2 = kOffW2 + kH * kH;   // 4352
constexpr int kOffW3 = kOffB2 + kH;        // 4416
constexpr int kOffB3 = kOffW3 + kH;        // 4480
constexpr int kParams = kOffB3 + 1;        // 4481
constexpr int kOffNorm = kParams;          // mu_x[3], sd_x[3], mu_y, sd_y, gamma
constexpr int kModel = kParams + 9;        // 4490

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// acc += A B over 32 k: lo holds the lane's k elements 0..3, hi 4..7 (of A
// and of B alike).
__device__ __forceinline__ f32x4 mfma32(bf16x4 a_lo, bf16x4 a_hi, bf16x4 b_lo, bf16x4 b_hi, f32x4 acc) {
  const s16x8 a = {a_lo[0], a_lo[1], a_lo[2], a_lo[3], a_hi[0], a_hi[1], a_hi[2], a_hi[3]};
  const s16x8 b = {b_lo[0], b_lo[1], b_lo[2], b_lo[3], b_hi[0], b_hi[1], b_hi[2], b_hi[3]};
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                 0, 0, 0);
}

__device__ __forceinline__ short f2bf(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
// d silu / dz = s (1 + z (1 - s)), s = sigmoid(z)
__device__ __forceinline__ float dsilu(float z) {
  const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Frozen-Lorentz-factor pressure p0 (see the top of this file), floored at
// 1e-12 D; NaN for E <= |m| (no physical state).
__device__ __forceinline__ double cold_pressure(double Gg, double D, double m, double E) {
  const double q = (E - m) * (E + m);
  const double W0 = E / sqrt(q);
  const double p0 = (E - D * W0) / (Gg * (E * E / q) - 1.0);
  return q > 0.0 ? fmax(p0, 1e-12 * D) : __builtin_nan("");
}

// Raw (un-normalised) features of a conserved state; *p0 receives the cold pressure.
__device__ __forceinline__ void raw_features(double Gg, double D, double m, double E, float r[3], double &p0) {
  p0 = cold_pressure(Gg, D, m, E);
  r[0] = logf((float)D);
  r[1] = log1pf((float)(m / D));
  r[2] = logf((float)(p0 / D));
}

// Normalised features; *p0 as raw_features.
__device__ __forceinline__ void features(const float *norm, double Gg, double D, double m, double E, float x[3],
                                         double &p0) {
  float r[3];
  raw_features(Gg, D, m, E, r, p0);
  x[0] = (r[0] - norm[0]) / norm[3];
  x[1] = (r[1] - norm[1]) / norm[4];
  x[2] = (r[2] - norm[2]) / norm[5];
}

// Pressure from the network output o (normalised log(p/p0)).
__device__ __forceinline__ double predicted_pressure(const float *norm, float o, double p0) {
  return p0 * (double)expf(o * norm[7] + norm[6]);
}

// Forward pass of one wave over 64 samples.  x[t][f] holds the normalised
// features of sample 16 t + lane%16 (the caller fills them, e.g. by shuffles).
// Returns, in every lane, the normalised outputs of its 4 samples s_t.
// W2b: W2 as bf16 row-major [out][in]; P: FP32 parameters.
__device__ __forceinline__ void forward_wave(const float *__restrict__ P, const short *__restrict__ W2b,
                                             const float x[4][3], float out[4]) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4, c = l & 15;
  // Layer 1 in the B-operand layout of H1^T: hidden 16 q + 4 g + e of sample t.
  bf16x4 h1[4][4];  // [q][t]
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int h = 16 * q + 4 * g + e;
      const float w0 = P[kOffW1 + 3 * h], w1 = P[kOffW1 + 3 * h + 1], w2 = P[kOffW1 + 3 * h + 2], bb = P[kOffB1 + h];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float z = ((w0 * x[t][0] + w1 * x[t][1]) + w2 * x[t][2]) + bb;
        h1[q][t][e] = f2bf(silu(z));
      }
    }
  }
  // Z2^T[h2][s] = sum_h1 W2[h2][h1] H1^T[h1][s]: tiles (u over h2, t over samples).
  float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    bf16x4 a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      a[q] = *reinterpret_cast<const bf16x4 *>(W2b + (16 * u + c) * kH + 16 * q + 4 * g);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 4; q += 2) acc = mfma32(a[q], a[q + 1], h1[q][t], h1[q + 1][t], acc);
      // lane holds Z2[s = 16 t + c][h2 = 16 u + 4 g + r]
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int h2 = 16 * u + 4 * g + r;
        part[t] += P[kOffW3 + h2] * silu(acc[r] + P[kOffB2 + h2]);
      }
    }
  }
  // Sum the 4 lane groups g (hidden units) of each sample.
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    float v = part[t];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    out[t] = v + P[kOffB3];
  }
}

}  // namespace rhllc_mlp
