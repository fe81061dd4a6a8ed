// Packed FP64 5-vectors of the mixed path: the face flux and the conserved
// state U.  Each component is stored as an FP32 value hi plus a signed 6-bit
// count e of 2^(E-156) units, E the biased exponent of hi (1/64 of hi's ulp
// for normal hi), so x = hi + e 2^(E-156) holds x to 2^-30 relative.
// FP32 alone (2^-24) is not enough for either: the momentum flux p n S of a
// cell's four faces cancels while their roundings do not, and U accumulates
// one rounding per step.  Record: float4 (c0..c3) + uint2 (bits of c4,
// e0 | e1 << 6 | e2 << 12 | e3 << 18 | e4 << 24) = 24 bytes.
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

// hi = (float)x; *e = the 6-bit remainder field.
__host__ __device__ __forceinline__ float split(double x, uint32_t &e) {
  const float hi = (float)x;
  const int E = (int)((f32_bits(hi) >> 23) & 0xFFu);
  double q = rint(ldexp(x - (double)hi, 156 - E));
  q = fmin(fmax(q, -31.0), 31.0);  // NaN remainders give -31: hi is NaN/inf then
  e = (uint32_t)((int)q & 63);
  return hi;
}

__host__ __device__ __forceinline__ double join(float hi, uint32_t e6) {
  const int E = (int)((f32_bits(hi) >> 23) & 0xFFu);
  const int e = (int)((e6 & 63u) ^ 32u) - 32;
  return (double)hi + ldexp((double)e, E - 156);
}

struct Packed {
  float4 a;
  uint2 b;
};

__host__ __device__ __forceinline__ Packed pack(double c0, double c1, double c2, double c3, double c4) {
  Packed o;
  uint32_t e0, e1, e2, e3, e4;
  o.a.x = split(c0, e0);
  o.a.y = split(c1, e1);
  o.a.z = split(c2, e2);
  o.a.w = split(c3, e3);
  o.b.x = f32_bits(split(c4, e4));
  o.b.y = e0 | (e1 << 6) | (e2 << 12) | (e3 << 18) | (e4 << 24);
  return o;
}

__host__ __device__ __forceinline__ void unpack(const float4 &a, const uint2 &b, double out[5]) {
  out[0] = join(a.x, b.y);
  out[1] = join(a.y, b.y >> 6);
  out[2] = join(a.z, b.y >> 12);
  out[3] = join(a.w, b.y >> 18);
  out[4] = join(bits_f32(b.x), b.y >> 24);
}

}  // namespace rhllc_ext
