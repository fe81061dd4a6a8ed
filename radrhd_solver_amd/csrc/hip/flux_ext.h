// Storage format of the mixed path's face flux: each component as an FP32
// value hi plus, for the momentum and energy, one signed byte e counting
// 2^(E-158) units, E the biased exponent of hi (1/256 of hi's ulp for normal
// hi).  x = hi + e 2^(E-158) holds x to 2^-33 relative -- FP32 alone (2^-24)
// leaves a visible error in the momentum update wherever the pressure term
// p n S of the four faces cancels around a cell while their roundings do not.
// Record per face: float4 (D, mx, my, mz) + uint2 (bits of E, bytes e of
// mx, my, mz, E).  Host and device use the same functions.
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

// hi = (float)x; *e = the byte of the remainder.
__host__ __device__ __forceinline__ float split(double x, uint32_t &byte) {
  const float hi = (float)x;
  const int E = (int)((f32_bits(hi) >> 23) & 0xFFu);
  double q = rint(ldexp(x - (double)hi, 158 - E));
  q = fmin(fmax(q, -127.0), 127.0);  // NaN remainders give -127: hi is NaN/inf then
  byte = (uint32_t)(uint8_t)(int8_t)(int)q;
  return hi;
}

__host__ __device__ __forceinline__ double join(float hi, uint32_t byte) {
  const int E = (int)((f32_bits(hi) >> 23) & 0xFFu);
  return (double)hi + ldexp((double)(int8_t)(uint8_t)byte, E - 158);
}

struct Packed {
  float4 a;  // D, mx, my, mz
  uint2 b;   // bits of E, extension bytes (mx | my << 8 | mz << 16 | E << 24)
};

__host__ __device__ __forceinline__ Packed pack(double d, double mx, double my, double mz, double E) {
  Packed o;
  uint32_t e0, e1, e2, e3;
  o.a.x = (float)d;
  o.a.y = split(mx, e0);
  o.a.z = split(my, e1);
  o.a.w = split(mz, e2);
  o.b.x = f32_bits(split(E, e3));
  o.b.y = e0 | (e1 << 8) | (e2 << 16) | (e3 << 24);
  return o;
}

__host__ __device__ __forceinline__ void unpack(const float4 &a, const uint2 &b, double out[5]) {
  out[0] = a.x;
  out[1] = join(a.y, b.y & 0xFFu);
  out[2] = join(a.z, (b.y >> 8) & 0xFFu);
  out[3] = join(a.w, (b.y >> 16) & 0xFFu);
  out[4] = join(bits_f32(b.x), b.y >> 24);
}

}  // namespace rhllc_ext
