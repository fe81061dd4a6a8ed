// FP32 values with a small integer extension, for the mixed path's stored
// FP64 quantities: hi = (float)x plus a signed BITS-bit count e of
// 2^(E - 150 - (BITS - 1)) units, E the biased exponent of hi, i.e.
// 2^-(BITS-1) of hi's ulp for normal hi.  x ~ hi + e 2^(E-150-(BITS-1)) to
// 2^-(23+BITS) relative (the remainder is rounded; |x - hi| <= ulp/2).
//
// Used for
//  * the face flux S*(D, m, E - D): float4 (D, mx, my, mz) + uint2 (bits of
//    E - D, four 8-bit counts for mx, my, mz, E - D) = 24 bytes, 2^-31
//    relative for m and E - D.  FP32 alone is not enough for those: the
//    momentum flux p n S of a cell's faces cancels around the cell while the
//    roundings do not, and at Lorentz factors ~10 c2p amplifies an error in E
//    by W^2 E/p (~1e3-1e4) into the pressure.  D in FP32 leaves the errors
//    unchanged, since its rounding reaches D and E alike (pack below);
//  * the cell record's s - D (s = rho h W^2) and D (8-bit counts): they must match
//    the record's FP32 v and p far better than FP32 rounding does.
// Host and device use the same functions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>

namespace rhllc_ext {

__host__ __device__ __forceinline__ uint32_t f32_bits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}
__host__ __device__ __forceinline__ float bits_f32(uint32_t u) {
  float x;
  memcpy(&x, &u, 4);
  return x;
}

// hi = (float)x; e = the remainder in units, rounded (the value 2^(BITS-1)
// that |x - hi| = ulp/2 can give is stored as 2^(BITS-1) - 1).
template <int BITS>
__host__ __device__ __forceinline__ float split(double x, uint32_t &e) {
  constexpr int kTop = (1 << (BITS - 1)) - 1;
  const float hi = (float)x;
  const int E = (int)((f32_bits(hi) >> 23) & 0xFFu);
  int q = (int)rint(ldexp(x - (double)hi, 150 + (BITS - 1) - E));
  q = q < kTop ? q : kTop;
  e = (uint32_t)q & ((1u << BITS) - 1u);
  return hi;
}

template <int BITS>
__host__ __device__ __forceinline__ double join(float hi, uint32_t e) {
  constexpr uint32_t kSign = 1u << (BITS - 1);
  const int E = (int)((f32_bits(hi) >> 23) & 0xFFu);
  const int q = (int)((e & ((1u << BITS) - 1u)) ^ kSign) - (int)kSign;
  return (double)hi + ldexp((double)q, E - 150 - (BITS - 1));
}

struct Packed {
  float4 a;
  uint2 b;
};

// Face flux record.  The energy slot holds the flux of tau = E - D (the
// caller forms it without going through the E flux), and unpack returns E as
// (FP32 D) + tau: the rounding of the D flux then enters D and E alike and
// leaves E - D alone.  In a cold gas (p / rho ~ 1e-6) tau is a millionth of
// E, and an E rebuilt from separately rounded D and E fluxes carried the FP32
// D rounding into tau amplified ~1e6 (transverse-momentum L1 4e-4 against
// FP64; 3e-6 with tau).
__host__ __device__ __forceinline__ Packed pack(double d, double mx, double my, double mz, double tau) {
  Packed o;
  uint32_t e1, e2, e3, e4;
  o.a.x = (float)d;
  o.a.y = split<8>(mx, e1);
  o.a.z = split<8>(my, e2);
  o.a.w = split<8>(mz, e3);
  o.b.x = f32_bits(split<8>(tau, e4));
  o.b.y = e1 | (e2 << 8) | (e3 << 16) | (e4 << 24);
  return o;
}

__host__ __device__ __forceinline__ void unpack(const float4 &a, const uint2 &b, double out[5]) {
  out[0] = a.x;
  out[1] = join<8>(a.y, b.y);
  out[2] = join<8>(a.z, b.y >> 8);
  out[3] = join<8>(a.w, b.y >> 16);
  out[4] = (double)a.x + join<8>(bits_f32(b.x), b.y >> 24);
}

}  // namespace rhllc_ext
