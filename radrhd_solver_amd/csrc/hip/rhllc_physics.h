// Pointwise RHLLC physics for gfx950 (device functions).
//
// Each function is written from the reference's algorithm and keeps its
// operation order, so the double instantiation is bitwise equal to the
// reference compiled with strict IEEE arithmetic (the library is built with
// -ffp-contract=off; every divide and sqrt is IEEE correctly rounded).  The
// float instantiation is the mixed-precision face path.
//
//   p2c          rhllc::GetConvValueStab   solvers/rhllc/src/rhllc_flux_stab.cpp:10-38
//   pressure_fix rhllc::PhysPressureFix    solvers/rhllc/src/rhllc_flux_stab.cpp:40-121
//   c2p          rhllc::GetPhysValueStab   solvers/rhllc/src/rhllc_flux_stab.cpp:123-220
//   rotation     GetRotationMatrix         lib/geometry/src/linear_alg.cpp:3-42
//   hllc_flux    rhllc::GetFluxStab        solvers/rhllc/src/rhllc_flux_stab.cpp:231-431
//
// EOS (template parameter, include/hd_config.h:16-20 selects it at compile
// time in the reference):
//   kIdeal  ideal gas, h = 1 + Gg p/rho, Gg = gamma/(gamma - 1), with a runtime
//           gamma (the reference's kGamma1 is a constexpr 5/3,
//           lib/global_consts.h:20-21);
//   kTaub   Taub-Mathews, h = 2.5 theta + sqrt(2.25 theta^2 + 1), theta = p/rho
//           (the reference's EOS TAUB branches; its shipped default).
//
// Sums of three products follow the reference's Eigen 3.4 evaluation on
// x86-64 (2-double packets): dot products (a0+a1)+a2; rows 0 and 1 of T*v
// (a0+a1)+a2, row 2 a0+(a1+a2); rows of T^T*v (a0+a1)+a2.
#pragma once

#include <hip/hip_runtime.h>

#include "flux_ext.h"

namespace rhllc_dev {

constexpr int kMaxIter = 40;  // rhllc_flux_stab.cpp:8

enum Eos : int { kIdeal = 0, kTaub = 1 };

template <typename R>
struct State {
  R d, v[3], p;  // W = (rho, v, p) or U = (D, m, E)
};

template <typename R>
__device__ __forceinline__ R smax(R a, R b) {  // std::max, NaN behaviour included
  return (a < b) ? b : a;
}
template <typename R>
__device__ __forceinline__ R smin(R a, R b) {  // std::min
  return (b < a) ? b : a;
}
template <typename R>
__device__ __forceinline__ R dot3(const R *a, const R *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
// The same with fused multiply-adds, for the mixed path: a function's
// `fp contract(fast)` does not reach into dot3, which the FP64 path needs
// unfused (three products and two sums, against one product and two FMAs).
__device__ __forceinline__ double dot3c(const double *a, const double *b) {
  return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
}
// a b + c for code shared by both policies: fused with A = Fast, the product
// and the sum rounded separately with A = Ieee (the reference's expression).
template <typename A>
__device__ __forceinline__ double mad(double a, double b, double c) {
  if constexpr (A::kFast)
    return fma(a, b, c);
  else
    return a * b + c;
}

// Arithmetic policies.  Ieee: correctly rounded divide and square root (the
// parity path; every expression keeps the reference's form).  Fast: hardware
// reciprocal / reciprocal square root refined by Newton steps.  Measured on
// gfx950 over 2^20 inputs spanning 2^+-100: v_rcp_f64 is within 4.6e-8
// relative, 2.2e-15 after one step and 1.1e-16 after two; v_rsq_f64 5.2e-8,
// 4.2e-15 after one step.  rcp takes two steps: one leaves 1/x up to 2e-15
// below the true value, and c2p's 1/(1 - m^2/alpha^2) must not drop below 1
// for m^2/alpha^2 ~ 1e-15 (the reference's lor2 < 1 test).  The quotient and
// the root take one step and one correction of the result, which brings both
// to within about an ulp.
struct Ieee {
  static constexpr bool kFast = false;
  __device__ static __forceinline__ double div(double a, double b) { return a / b; }
  __device__ static __forceinline__ double rcp(double x) { return 1.0 / x; }
  __device__ static __forceinline__ double sqrt_(double x) { return sqrt(x); }
  __device__ static __forceinline__ double rsqrt(double x) { return 1.0 / sqrt(x); }
};
struct Fast {
  static constexpr bool kFast = true;
  __device__ static __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
  }
  __device__ static __forceinline__ double div(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);  // one correction of the quotient
  }
  __device__ static __forceinline__ double sqrt_(double x) {
    // y ~ 1/sqrt(x), one Newton step, then s = x*y with one correction.
    // Negative or NaN x gives NaN through v_rsq_f64; x = 0 (rsq = inf) is
    // returned as is.
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    const double s = x * y;
    const double r = fma(fma(-s, s, x), 0.5 * y, s);
    return x == 0.0 ? x : r;
  }
  __device__ static __forceinline__ double rsqrt(double x) {  // x > 0
    const double y = __builtin_amdgcn_rsq(x);
    return y * fma(-0.5 * x * y, y, 1.5);
  }
};

// The mixed face work's policy (also the M1 closure of the mixed radiation
// source): one Newton step on the hardware reciprocal and reciprocal root, no
// correction of quotients and roots.  Relative errors stay below 4.2e-15 (the
// one-step figures above), far below the FP32 rounding of the cell records
// the flux is formed from; c2p keeps Fast.
struct FaceFast {
  static constexpr bool kFast = true;
  __device__ static __forceinline__ double rcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
  }
  __device__ static __forceinline__ double div(double a, double b) { return a * rcp(b); }
  // x = 0 returns 0, as Fast does, without a branch: the root is taken of
  // x + 1e-300 (equal to x from x = 1e-284 on) and multiplied by x.
  __device__ static __forceinline__ double sqrt_(double x) {
    const double xs = x + 1e-300;
    double y = __builtin_amdgcn_rsq(xs);
    y = y * fma(-0.5 * xs * y, y, 1.5);
    return x * y;
  }
  __device__ static __forceinline__ double rsqrt(double x) { return Fast::rsqrt(x); }
  // The hardware approximations alone (within 5.2e-8 and 4.6e-8 relative),
  // for the signal speeds: the HLLC flux depends on them smoothly, and two
  // equal states give the physical flux whatever they are.
  __device__ static __forceinline__ double sqrt_rough(double x) {  // x = 0 returns 0, as sqrt_
    return x * __builtin_amdgcn_rsq(x + 1e-300);
  }
  __device__ static __forceinline__ double rcp_rough(double x) { return __builtin_amdgcn_rcp(x); }
};

template <typename R>
struct Rot {
  R t[9];  // row-major, rows (n, t1, t2)
};

template <typename R, typename A = Ieee>
__device__ __forceinline__ Rot<R> rotation(R n0, R n1, R n2) {
  Rot<R> T;
#pragma unroll
  for (int i = 0; i < 9; ++i) T.t[i] = R(0);
  const R eps = R(1e-12);
  if (fabs(n2 * n2 - R(1)) > eps) {
    T.t[0] = n0;
    T.t[1] = n1;
    T.t[2] = n2;
    const R sqr = A::sqrt_(R(1) - n2 * n2);
    const R sqr_inv = A::rcp(sqr);
    T.t[3] = -n1 * sqr_inv;
    T.t[4] = n0 * sqr_inv;
    T.t[6] = -n0 * n2 * sqr_inv;
    T.t[7] = -n1 * n2 * sqr_inv;
    T.t[8] = sqr;
  } else if (n2 > R(0)) {
    T.t[2] = R(1);
    T.t[4] = R(1);
    T.t[6] = R(-1);
  } else {
    T.t[2] = R(-1);
    T.t[4] = R(-1);
    T.t[6] = R(1);
  }
  return T;
}

// r = T v
template <typename R>
__device__ __forceinline__ void rot_fwd(const Rot<R> &T, const R *v, R *r) {
  const R r0 = (T.t[0] * v[0] + T.t[1] * v[1]) + T.t[2] * v[2];
  const R r1 = (T.t[3] * v[0] + T.t[4] * v[1]) + T.t[5] * v[2];
  const R r2 = T.t[6] * v[0] + (T.t[7] * v[1] + T.t[8] * v[2]);
  r[0] = r0, r[1] = r1, r[2] = r2;
}
// r = T^T v
template <typename R>
__device__ __forceinline__ void rot_back(const Rot<R> &T, const R *v, R *r) {
  const R r0 = (T.t[0] * v[0] + T.t[3] * v[1]) + T.t[6] * v[2];
  const R r1 = (T.t[1] * v[0] + T.t[4] * v[1]) + T.t[7] * v[2];
  const R r2 = (T.t[2] * v[0] + T.t[5] * v[1]) + T.t[8] * v[2];
  r[0] = r0, r[1] = r1, r[2] = r2;
}

// The wall of boundary code -2 outside GEOMETRY_TYPE Cone (rhllc_utils.cpp:
// 95-110): the normal velocity mirrored, v - 2 (v.n) n, with v.n summed as
// (v0 n0 + v1 n1) + v2 n2.  The reference writes it as T^T diag(-1, 1, 1) T v
// with T = GetRotationMatrix(n), forming T^T by `T = T.transpose()`; Eigen
// rejects that assignment without NDEBUG (aliasing), and with NDEBUG it
// overwrites T's lower triangle with its upper one before the second product.
// This is the reflection that code intends (equal to it up to rounding).
template <typename R>
__device__ __forceinline__ void wall_mirror(R n0, R n1, R n2, R *v) {
  const R vn = (v[0] * n0 + v[1] * n1) + v[2] * n2;
  const R t = vn + vn;
  v[0] = v[0] - t * n0, v[1] = v[1] - t * n1, v[2] = v[2] - t * n2;
}

// Conserved from primitive; clamps |v| >= 1 to 0.9999 and rewrites W.v as the reference does.
template <typename R, typename A = Ieee, int EOS = kIdeal>
__device__ __forceinline__ void p2c(R Gg, State<R> &W, State<R> &U) {
  R h;
  if (EOS == kIdeal) {
    h = R(1) + A::div(Gg * W.p, W.d);
  } else {
    const R theta = A::div(W.p, W.d);
    h = R(2.5) * theta + A::sqrt_(R(2.25) * theta * theta + R(1));
  }
  R g = dot3(W.v, W.v);
  if (g >= R(1)) {
    const R beta_fix = R(0.9999);
    g = A::div(beta_fix, A::sqrt_(g));
    W.v[0] *= g, W.v[1] *= g, W.v[2] *= g;
    g = beta_fix * beta_fix;
  }
  g = A::rcp(R(1) - g);
  const R scrh = W.d * h * g;
  g = A::sqrt_(g);
  U.d = W.d * g;
  U.v[0] = scrh * W.v[0], U.v[1] = scrh * W.v[1], U.v[2] = scrh * W.v[2];
  U.p = scrh - W.p;
}

// Secant fallback at p = 1e-12 (kMinPressure); rewrites U.E on success.  Returns 0 or 2.
// D h at pressure p lor: ideal D + plor Gg; Taub-Mathews 2.5 plor + sqrt(2.25 plor^2 + D^2).
template <int EOS>
__device__ __forceinline__ double fix_Dh(double Gg, double D, double plor) {
  return EOS == kIdeal ? D + plor * Gg : 2.5 * plor + sqrt(2.25 * plor * plor + D * D);
}
template <int EOS = kIdeal>
__device__ __forceinline__ int pressure_fix(double Gg, State<double> &U, State<double> &W) {
  const double p = 1e-12;
  const double D = U.d;
  const double m2 = dot3(U.v, U.v);
  const double m = sqrt(m2);
  const double umax = m / D;
  double u0 = umax;
  double lor = sqrt(1.0 + u0 * u0);
  double plor = p * lor;
  const double alpha = EOS == kIdeal ? Gg : 2.5;
  double Dh = fix_Dh<EOS>(Gg, D, plor);
  double f0 = m / Dh - u0;
  double u1 = (-D + sqrt(D * D + 4.0 * m * alpha * p)) / (2.0 * alpha * p);
  int done = 0;
  double f1, du;
  for (int k = 1; k < kMaxIter; k++) {
    lor = sqrt(1.0 + u1 * u1);
    plor = p * lor;
    Dh = fix_Dh<EOS>(Gg, D, plor);
    f1 = m / Dh - u1;
    if (done == 1) break;
    du = (u1 - u0) / (f1 - f0) * f1;
    u0 = u1;
    f0 = f1;
    u1 -= du;
    u1 = smin(u1, umax);
    u1 = smax(u1, 0.0);
    if (fabs(f1) < 1.e-9) done = 1;
  }
  if (done != 1) return 2;
  lor = sqrt(1.0 + u1 * u1);
  plor = p * lor;
  Dh = fix_Dh<EOS>(Gg, D, plor);
  W.d = U.d / lor;
  W.p = p;
  U.p = Dh * lor - p;
  f0 = 1.0 / (Dh * lor);
  W.v[0] = U.v[0] * f0, W.v[1] = U.v[1] * f0, W.v[2] = U.v[2] * f0;
  return 0;
}

// Newton c2p.  guess < 0 selects the reference's initial pressure
// max(|m| - E, 0, pmin); otherwise the start is max(guess, pmin).
// Returns 0 or the reference's codes 1..4; 5 = the reference's DIE_IF(lor2 < 1).
// On failure W is left unchanged.  *iters = Newton iterations done.
// A = Ieee is the reference's arithmetic, operation for operation.  The mixed
// path (A = Fast) runs the same iteration with the Fast divide/sqrt, fused
// multiply-adds, and these changes:
//  * per iteration one reciprocal root y = 1/sqrt(alpha^2 - m^2) gives
//    Gamma = alpha y (clamped to >= 1: rounding takes it to 1 - ulp for gas
//    at rest) and m^2 Gamma^3/alpha^3 = m^2 y^3; the reference's lor2 < 1
//    failure becomes alpha^2 - m^2 <= 0, the same states;
//  * for the ideal gas f and f' are formed without 1/D (D tau = Gamma);
//  * it also stops once |dp| < 1e-6 p.  Newton's next correction is then
//    K dp^2 with K = |f''/2f'|, and K p <= 23 over the surrogate's training
//    box (rho 1e-4..1e2, p/rho 1e-6..1e2, W <= 50, gamma 4/3 and 5/3), so the
//    pressure is within 2.3e-11 p of the root and |f|/(E + p) <= ~1e-11 --
//    the reference's own stop (|dp| < 1e-11 E) usually costs one more
//    iteration to confirm what the previous one already reached.
template <typename A = Ieee, int EOS = kIdeal>
__device__ __forceinline__ int c2p(double Gg, const State<double> &U, State<double> &W, double guess, int &iters) {
  const double tol = 1.e-11;
  const double eps2 = 1.e-12;
  const double D = U.d;
  const double E = U.p;
  const double m2 = A::kFast ? dot3c(U.v, U.v) : dot3(U.v, U.v);
  iters = 0;
  double m, pmin;
  if (!A::kFast) {
    const double Q = E - A::sqrt_(m2 + D * D);
    if (Q < 0.0) return 1;
    m = A::sqrt_(m2);
    pmin = A::sqrt_(A::div(m2, 1.0 - eps2)) - E;
  } else {
#pragma clang fp contract(fast)
    // E < sqrt(m^2 + D^2) without the root; sqrt(m^2/(1 - eps2)) = m / sqrt(1 - eps2).
    // |m| to one Newton step (4.2e-15): it only places the floor pmin.
    if (!(E >= 0.0 && E * E >= m2 + D * D)) return 1;
    m = FaceFast::sqrt_(m2);
    pmin = m * 1.0000000000005 - E;
  }
  double p = smax(m - E, 0.0);
  const double D_1 = A::rcp(D);
  if (guess >= 0.0) p = guess;
  p = smax(p, pmin);
  double alpha, alpha2, lor2, lor, tau, theta, h, dh_dp, dh_dtau, yp = 0.0, dp, dyp;
  int iter;
  for (iter = 0; iter < kMaxIter; iter++) {
    alpha = E + p;
    alpha2 = alpha * alpha;
    double mlor3;  // m^2 Gamma^3 / alpha^3
    if (!A::kFast) {
      lor2 = 1.0 - A::div(m2, alpha2);
      lor2 = A::rcp(lor2);
      if (lor2 < 1.0) {
        iters = iter + 1;
        return 5;
      }
      lor = A::sqrt_(lor2);
      mlor3 = A::div(m2 * lor2 * lor, alpha2 * alpha);
    } else {
      // One refined reciprocal root y = 1/sqrt(alpha^2 - m^2) gives
      // Gamma = alpha y (clamped to >= 1: rounding takes it to 1 - ulp for gas
      // at rest) and m^2 Gamma^3 / alpha^3 = m^2 y^3.
      const double q = fma(alpha, alpha, -m2);
      if (!(q > 0.0)) {
        iters = iter + 1;
        return 5;
      }
      const double y = A::rsqrt(q);
      lor = smax(alpha * y, 1.0);
      lor2 = lor * lor;
      mlor3 = m2 * (y * y) * y;
    }
    if (A::kFast && EOS == kIdeal) {
      // The same f and f' for the ideal gas without 1/D: D tau = Gamma, so
      // D h = D + g with g = Gg p Gamma, D Gamma dh_dp = Gg Gamma^2 and
      // Gamma dh_dtau = g.
      const double g = Gg * p * lor;
      const double Dh = D + g;
      yp = fma(Dh, lor, -E) - p;
      dyp = fma(-mlor3, Dh + g, fma(Gg * lor, lor, -1.0));
    } else {
      tau = lor * D_1;
      theta = p * tau;
      if (EOS == kIdeal) {
        h = mad<A>(Gg, theta, 1.0);  // 1 + Gg theta
        dh_dp = Gg * tau;
        dh_dtau = Gg * p;
      } else {
        h = 2.5 * theta + A::sqrt_(2.25 * theta * theta + 1.0);
        const double scrh = A::div(5.0 * h - 8.0 * theta, 2.0 * h - 5.0 * theta);
        dh_dp = tau * scrh;
        dh_dtau = p * scrh;
      }
      // yp = D h lor - E - p, dyp = D lor dh_dp - mlor3 (lor dh_dtau + D h) - 1
      const double Dh = D * h;
      yp = mad<A>(Dh, lor, -E) - p;
      dyp = mad<A>(-mlor3, mad<A>(lor, dh_dtau, Dh), D * lor * dh_dp) - 1.0;
    }
    // Fast: the quotient to one Newton step of the reciprocal (4e-15): the
    // iteration's fixed point is yp = 0 whatever the accuracy of dyp.
    dp = A::kFast ? yp * FaceFast::rcp(dyp) : A::div(yp, dyp);
    p -= dp;
    if (p < pmin) p = pmin;
    if (fabs(dp) < tol * E) break;
    if (A::kFast && fabs(dp) < 1e-6 * p) break;
  }
  iters = iter < kMaxIter ? iter + 1 : kMaxIter;
  if (p < 0.0) return 2;
  if (isnan(p)) return 4;
  if (iter >= kMaxIter || fabs(A::div(yp, E + p)) > 1.e-4 || p < (m - E)) return 3;
  if (!A::kFast) {
    alpha = E + p;
    alpha2 = alpha * alpha;
    lor2 = A::div(alpha2, alpha2 - m2);
    lor = A::sqrt_(lor2);
    tau = lor * D_1;
    W.d = A::rcp(tau);
    W.p = p;
    const double den = U.p + p;
    W.v[0] = A::div(U.v[0], den), W.v[1] = A::div(U.v[1], den), W.v[2] = A::div(U.v[2], den);
  } else {
#pragma clang fp contract(fast)
    // rho = D / Gamma = D sqrt(alpha^2 - m^2) / alpha, v = m / alpha, each to
    // one Newton step (4.2e-15), far below their FP32 storage.
    alpha = E + p;
    const double ia = FaceFast::rcp(alpha);
    W.d = D * FaceFast::sqrt_(smax(alpha * alpha - m2, 0.0)) * ia;
    W.p = p;
    W.v[0] = U.v[0] * ia, W.v[1] = U.v[1] * ia, W.v[2] = U.v[2] * ia;
  }
  return 0;
}

// |f(p)| / (E + p) of the c2p equation at the recovered state (used for the
// residual check of the mixed path; not part of the reference).
template <int EOS = kIdeal>
__device__ __forceinline__ double c2p_residual(double Gg, const State<double> &U, double p) {
  const double D = U.d, E = U.p;
  const double m2 = dot3(U.v, U.v);
  const double alpha = E + p;
  const double lor = sqrt(1.0 / (1.0 - m2 / (alpha * alpha)));
  const double theta = p * lor / D;
  const double h = EOS == kIdeal ? 1.0 + Gg * theta : 2.5 * theta + sqrt(2.25 * theta * theta + 1.0);
  return fabs(D * h * lor - E - p) / alpha;
}

template <typename R, typename A = Ieee>
__device__ __forceinline__ void signal_speed(R cs2, R Vn, R V2, R &cmin, R &cmax) {
  const R Vn2 = Vn * Vn;
  const R Vt2 = V2 - Vn2;
  const R sroot = A::sqrt_(cs2 * (R(1) - Vn2 - Vt2 * cs2) * (R(1) - V2));
  cmax = A::div(Vn * (R(1) - cs2) + sroot, R(1) - V2 * cs2);
  cmin = A::div(Vn * (R(1) - cs2) - sroot, R(1) - V2 * cs2);
}

// HLLC flux through a face with unit normal n (pointing out of the left cell)
// and area S.  F receives S * flux in the global frame; returns max(lL, lR).
template <typename R, typename A = Ieee, int EOS = kIdeal>
__device__ __forceinline__ R hllc_flux(R G1, R Gg, const State<R> &cUL, const State<R> &cUR, const State<R> &cWL,
                                       const State<R> &cWR, R n0, R n1, R n2, R S, State<R> &Fout) {
  const Rot<R> T = rotation<R, A>(n0, n1, n2);
  State<R> UL = cUL, UR = cUR, WL = cWL, WR = cWR;
  rot_fwd(T, cUL.v, UL.v);
  rot_fwd(T, cUR.v, UR.v);
  rot_fwd(T, cWL.v, WL.v);
  rot_fwd(T, cWR.v, WR.v);
  const R *VL = WL.v, *VR = WR.v;
  const R dL = WL.d, dR = WR.d, pL = WL.p, pR = WR.p;
  const R VVL = dot3(VL, VL), VVR = dot3(VR, VR);
  R hL, hR, csL2, csR2;
  if (EOS == kIdeal) {
    hL = R(1) + A::div(Gg * pL, dL);
    hR = R(1) + A::div(Gg * pR, dR);
    csL2 = A::div(G1 * pL, dL * hL);
    csR2 = A::div(G1 * pR, dR * hR);
  } else {
    const R thL = A::div(pL, dL), thR = A::div(pR, dR);
    hL = R(2.5) * thL + A::sqrt_(R(2.25) * thL * thL + R(1));
    hR = R(2.5) * thR + A::sqrt_(R(2.25) * thR * thR + R(1));
    csL2 = A::div(A::div(thL, R(3) * hL) * (R(5) * hL - R(8) * thL), hL - thL);
    csR2 = A::div(A::div(thR, R(3) * hR) * (R(5) * hR - R(8) * thR), hR - thR);
  }
  R cmaxL, cmaxR, cminL, cminR;
  signal_speed<R, A>(csL2, VL[0], VVL, cminL, cmaxL);
  signal_speed<R, A>(csR2, VR[0], VVR, cminR, cmaxR);
  const R lL = smin(cminL, cminR);
  const R lR = smax(cmaxL, cmaxR);
  State<R> F;
  if (lL >= R(0)) {
    F.d = UL.d * VL[0];
    F.v[0] = UL.v[0] * VL[0], F.v[1] = UL.v[1] * VL[0], F.v[2] = UL.v[2] * VL[0];
    F.p = UL.v[0];
    F.v[0] += pL;
  } else if (lR <= R(0)) {
    F.d = UR.d * VR[0];
    F.v[0] = UR.v[0] * VR[0], F.v[1] = UR.v[1] * VR[0], F.v[2] = UR.v[2] * VR[0];
    F.p = UR.v[0];
    F.v[0] += pR;
  } else {
    // The star state is taken on the side of the contact speed us; both
    // sides share one code path with (lam, U, V, p, F) of that side.
    const R FLd = UL.d * VL[0], FLv0 = UL.v[0] * VL[0], FLv1 = UL.v[1] * VL[0], FLv2 = UL.v[2] * VL[0];
    const R FLp = UL.v[0];
    const R FRd = UR.d * VR[0], FRv0 = UR.v[0] * VR[0], FRv1 = UR.v[1] * VR[0], FRv2 = UR.v[2] * VR[0];
    const R FRp = UR.v[0];
    const R AL = lL * UL.p - FLp;
    const R AR = lR * UR.p - FRp;
    const R BL = lL * UL.v[0] - FLv0 - pL;
    const R BR = lR * UR.v[0] - FRv0 - pR;
    const R a = AR * lL - AL * lR;
    const R b = AL + BL * lR - AR - BR * lL;
    const R c = BR - BL;
    const R sgn = b < R(0) ? R(-1) : R(1);  // SIGN, lib/global_def.h:62
    const R scrh = R(-0.5) * (b + sgn * A::sqrt_(b * b - R(4) * a * c));
    const R us = A::div(c, scrh);
    const R ps = A::div(AL * us - BL, R(1) - us * lL);
    const bool left = us >= R(0);
    const R lam = left ? lL : lR;
    const State<R> &Us = left ? UL : UR;
    const R Vn = left ? VL[0] : VR[0];
    const R ph = left ? pL : pR;
    const R Fd = left ? FLd : FRd, Fv0 = left ? FLv0 : FRv0, Fv1 = left ? FLv1 : FRv1, Fv2 = left ? FLv2 : FRv2;
    const R Fp = left ? FLp : FRp;
    const R dif = A::rcp(lam - us);
    R sd = Us.d * (lam - Vn) * dif;
    R sv0 = (lam * (Us.p + ps) - Us.v[0]) * us * dif;
    R sv1 = Us.v[1] * (lam - Vn) * dif;
    R sv2 = Us.v[2] * (lam - Vn) * dif;
    R sp = Us.p + A::div(sv0 - Us.v[0], lam);
    sd -= Us.d, sv0 -= Us.v[0], sv1 -= Us.v[1], sv2 -= Us.v[2], sp -= Us.p;
    sd *= lam, sv0 *= lam, sv1 *= lam, sv2 *= lam, sp *= lam;
    F.d = Fd + sd;
    F.v[0] = Fv0 + sv0, F.v[1] = Fv1 + sv1, F.v[2] = Fv2 + sv2;
    F.p = Fp + sp;
    F.v[0] += ph;
  }
  Fout.d = F.d;
  Fout.p = F.p;
  rot_back(T, F.v, Fout.v);
  Fout.d *= S, Fout.v[0] *= S, Fout.v[1] *= S, Fout.v[2] *= S, Fout.p *= S;
  return smax(lL, lR);
}

// ---------------------------------------------------------------- mixed path
// Face flux of the mixed path: FP32 primitives and FP64 geometry in, FP64
// arithmetic with the Fast policy and FMA contraction.  Same HLLC as
// hllc_flux (same signal speeds, same quadratic for the contact speed, same
// side selection), written without the rotation matrix: in the face frame
// every tangential flux component is its tangential momentum times one factor
// c_t, so the global momentum flux is c_t m + (F_n - c_t m_n) n, with F_n the
// normal component.  Only v.n and m.n of each side are needed.
// What the face work reads of a cell on the mixed path, stored by the cell
// work as two 16-byte halves: (vx, vy, vz, p) in FP32, and (sigma, D) in FP64
// with sigma = s - D, s = rho h W^2 (m = s v, E = s - p), D = rho W.  sigma
// and D are formed in FP64 from the FP32-rounded primitives, so they match
// the stored v and p: at Lorentz factors ~10 an s rounded on its own
// (inconsistent with v, p by 6e-8) raised the transverse-momentum L1 of a
// shear layer from 1e-6 to 2e-5.  The record holds sigma rather than s
// because E - D = sigma - p is the thermal and kinetic energy: in a gas at
// T ~ 1e4 K (p / rho c^2 ~ 1e-9) it is far below the resolution of an FP32 s.
// Forming these once per cell, instead of once per face side, removes the
// p2c divides and roots from the face work.  The sound speed is not stored:
// a face side forms it from rho h = s (1 - v^2) (ideal gas) or from
// rho = D sqrt(1 - v^2) (Taub-Mathews).
struct MixedIn {
  double v[3], p, s, D;
  double sigma;  // s - D
};

// The record of an FP64 primitive state (|v| < 1).
template <int EOS = kIdeal>
__device__ __forceinline__ void mixed_record(double Gg, const State<double> &W, float4 &va, double2 &sg) {
#pragma clang fp contract(fast)
  using A = FaceFast;  // 4.2e-15: sigma and D match the FP32 v and p far below their rounding
  va = make_float4((float)W.v[0], (float)W.v[1], (float)W.v[2], (float)W.p);
  const double rho = (float)W.d, p = va.w;
  const double v[3] = {va.x, va.y, va.z};
  const double vv = dot3c(v, v);
  const double lor = A::rsqrt(1.0 - vv);  // one reciprocal root: 4.2e-15
  const double lor2 = lor * lor;
  double rhm1;  // rho (h - 1)
  if (EOS == kIdeal) {
    rhm1 = Gg * p;
  } else {
    const double th = A::div(p, rho);
    const double r = A::sqrt_(2.25 * th * th + 1.0);
    rhm1 = 2.5 * p + rho * A::div(2.25 * th * th, r + 1.0);  // rho (h - 1) without the cancellation
  }
  // sigma = s - D = rho W (h W - 1) = W (rho (h - 1) W + rho (W - 1)), W - 1 = W^2 v^2 / (W + 1)
  sg.x = lor * (rhm1 * lor + rho * A::div(lor2 * vv, lor + 1.0));
  sg.y = rho * lor;
}

__device__ __forceinline__ MixedIn mixed_in(const float4 &va, const double2 &sg) {
  MixedIn o;
  o.v[0] = va.x, o.v[1] = va.y, o.v[2] = va.z, o.p = va.w;
  o.sigma = sg.x;
  o.D = sg.y;
  o.s = o.D + o.sigma;
  return o;
}

struct MixedSide {
  double p, v[3];           // primitives (v after the clamp)
  double D, E, s;           // conserved D, E and s (m = s v)
  double sigma;             // s - D
  double vn, mn;            // v.n, m.n (n the unit normal)
  double cmin, cmax;        // signal speeds
};

// sv = S n, rS = 1/S.
template <int EOS>
__device__ __forceinline__ void mixed_side(double gamma, const MixedIn &w, const double *sv, double rS,
                                           MixedSide &o) {
#pragma clang fp contract(fast)
  using A = FaceFast;
  o.p = w.p;
  o.v[0] = w.v[0], o.v[1] = w.v[1], o.v[2] = w.v[2];
  double V2 = dot3c(o.v, o.v);
  if (__builtin_expect(V2 >= 1.0, 0)) {
    // GetConvValueStab's clamp, for a velocity that rounded to |v| >= 1 in
    // FP32 (Lorentz factors above ~3000); s and D keep their FP64 values.
    const double g = 0.9999 * A::rsqrt(V2);
    o.v[0] *= g, o.v[1] *= g, o.v[2] *= g;
    V2 = 0.9999 * 0.9999;
  }
  o.s = w.s;
  o.D = w.D;
  o.sigma = w.sigma;
  o.E = w.D + (w.sigma - w.p);  // E - D = sigma - p, formed from the two small terms
  o.vn = dot3c(o.v, sv) * rS;
  o.mn = o.s * o.vn;
  double cs2;  // sound speed^2, to the hardware reciprocal's 5e-8 (it enters only the signal speeds)
  if (EOS == kIdeal) {
    cs2 = gamma * o.p * A::rcp_rough(o.s * (1.0 - V2));  // gamma p / (rho h)
  } else {
    const double th = o.p * A::rcp_rough(o.D * A::sqrt_rough(1.0 - V2));  // p / rho
    const double h = 2.5 * th + A::sqrt_rough(2.25 * th * th + 1.0);
    cs2 = th * (5.0 * h - 8.0 * th) * A::rcp_rough(3.0 * h * (h - th));
  }
  const double vn2 = o.vn * o.vn;
  const double sroot = A::sqrt_rough(cs2 * (1.0 - vn2 - (V2 - vn2) * cs2) * (1.0 - V2));
  const double r = A::rcp_rough(1.0 - V2 * cs2);
  const double a = o.vn * (1.0 - cs2);
  o.cmax = (a + sroot) * r;
  o.cmin = (a - sroot) * r;
}

// The face geometry is the vector S n (sx, sy, sz), exact to 2^-30 of its size,
// so a cell's four vectors close to that; the area S and v.n = v.(S n)/S are
// recovered from it (the frame-free form relies on |n| = 1: m = m_n n + m_t),
// and the flux's normal part is scaled by S n directly.  The contact solve is
// evaluated for every face and its result taken only in the subsonic case
// (lL < 0 < lR), so the face needs no branch.
// F receives S * flux with F.p the flux of E - D (not of E); returns
// max(lL, lR) as hllc_flux does.
// wall: the right state is the left one with the normal velocity mirrored
// (boundary code -2 outside GEOMETRY_TYPE Cone), v - 2 (v.n) n.
template <int EOS>
__device__ __forceinline__ double mixed_face_flux(double gamma, const MixedIn &WL, const MixedIn &WR_in, double sx,
                                                  double sy, double sz, State<double> &F, bool wall = false) {
#pragma clang fp contract(fast)
  using A = FaceFast;
  const double sv[3] = {sx, sy, sz};
  const double q = dot3c(sv, sv);
  // 1/S; a zero-area face (q = 0) takes a finite rS and S = 0, so a zero flux
  const double rS = A::rsqrt(q + 1e-300);
  const double S = q * rS;
  MixedSide L, R;
  MixedIn WR = WR_in;
  if (__builtin_expect(wall, 0)) wall_mirror(sx * rS, sy * rS, sz * rS, WR.v);
  mixed_side<EOS>(gamma, WL, sv, rS, L);
  mixed_side<EOS>(gamma, WR, sv, rS, R);
  const double lL = smin(L.cmin, R.cmin);
  const double lR = smax(L.cmax, R.cmax);
  // contact solve (rhllc_flux_stab.cpp:362-421)
  const double AL = lL * L.E - L.mn;
  const double AR = lR * R.E - R.mn;
  const double BL = lL * L.mn - L.mn * L.vn - L.p;
  const double BR = lR * R.mn - R.mn * R.vn - R.p;
  const double a = AR * lL - AL * lR;
  const double b = AL + BL * lR - AR - BR * lL;
  const double c = BR - BL;
  const double sgn = b < 0.0 ? -1.0 : 1.0;
  const double scrh = -0.5 * (b + sgn * A::sqrt_(b * b - 4.0 * a * c));
  const double us = A::div(c, scrh);
  const double ps = A::div(AL * us - BL, 1.0 - us * lL);
  const bool sup = lL >= 0.0 || lR <= 0.0;
  const bool left = sup ? lL >= 0.0 : us >= 0.0;
  const double vn = left ? L.vn : R.vn, mn = left ? L.mn : R.mn, p = left ? L.p : R.p;
  const double E = left ? L.E : R.E, D = left ? L.D : R.D, s = left ? L.s : R.s;
  const double sig = left ? L.sigma : R.sigma;
  const double v0 = left ? L.v[0] : R.v[0], v1 = left ? L.v[1] : R.v[1], v2 = left ? L.v[2] : R.v[2];
  const double lam = left ? lL : lR;
  const double dif = A::rcp(lam - us);
  const double ms = (lam * (E + ps) - mn) * us * dif;  // star normal momentum
  const double ct = sup ? vn : vn + lam * ((lam - vn) * dif - 1.0);
  const double Fn = sup ? mn * vn + p : mn * vn + lam * (ms - mn) + p;
  const double ctS = ct * S;
  const double cmS = ctS * s;       // S c_t m = S c_t s v
  const double cn = Fn - ct * mn;   // normal remainder (the pressure-like part), times S n below
  F.d = D * ctS;
  F.v[0] = cmS * v0 + cn * sx;
  F.v[1] = cmS * v1 + cn * sy;
  F.v[2] = cmS * v2 + cn * sz;
  // F_E - F_D from the small terms only: sigma v_n outside the fan, and in the
  // star region m*_n - D* u* = u* (lam (sigma - p + p*) - sigma v_n) / (lam - u*)
  // (D* = D (lam - v_n) / (lam - u*)); forming F_E and F_D and subtracting
  // would leave E - D with the rounding of D.  (F_E = m*_n in the star region.)
  const double Ftau = sup ? sig * vn : us * dif * (lam * (sig - p + ps) - sig * vn);
  F.p = Ftau * S;
  return smax(lL, lR);
}

// Radiation 4-momentum source, solvers/RadRHD/src/radRHD_utils.cpp:8-158 with
// the M1 Eddington tensor (ISOTROPIC_INTENSITY == 1 branch).  EddTensor(a, b) is evaluated once per
// component from shared subexpressions; every component keeps the reference's
// operation order, so the values are bitwise those of nine separate calls
// (A = Ieee; the mixed path runs the same sequence with the Fast policy).
template <typename A = Ieee>
__device__ __forceinline__ void edd_matrix(double U, const double *F, double D[3][3]) {
  double fm = dot3(F, F);
  if (fm < 1e-150 || U < 1e-150) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) D[a][b] = a == b ? 0.33333333333333 : 0.0;
    return;
  }
  fm = A::sqrt_(fm);
  double n[3];
  if (!A::kFast) {
    n[0] = A::div(F[0], fm), n[1] = A::div(F[1], fm), n[2] = A::div(F[2], fm);
  } else {
    const double r = A::rcp(fm);
    n[0] = F[0] * r, n[1] = F[1] * r, n[2] = F[2] * r;
  }
  if (fm > U) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) D[a][b] = n[a] * n[b];
    return;
  }
  double f2 = A::div(fm, U);
  f2 = f2 * f2;
  const double chi_2 = A::div(1.5 + 2.0 * f2, 5.0 + 2.0 * A::sqrt_(4.0 - 3.0 * f2));
  const double c = 3.0 * chi_2 - 0.5;
  const double dg = 0.5 - chi_2;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double e = c * n[a] * n[b];
      if (a == b) e += dg;
      D[a][b] = e;
    }
}

// Physical constants of the source, lib/global_consts.h:12-83 (CGS branch)
// evaluated as the reference's constexpr chain, and the three logarithms its
// header generator writes into log_global_consts.h (resources/parser_header.cpp,
// 16 significant digits), which full_phys_data_t::Init uses through LOG().
namespace radc {
constexpr double kPI = 3.1415926535897932384626433832795;
constexpr double kC_Light = 3 * 1e10;
constexpr double kM_hydrogen = 1.6735575 * 1e-24;
constexpr double k_boltzmann = 1.3807 * 1e-16;
constexpr double kStefanBoltzmann = 5.670374419 * 1e-5;
constexpr double kSigma_thomson = 6.65210 * 1e-25;
constexpr double kGravity = 6.6726 * 1e-8;
constexpr double kDist = 1e13;
constexpr double kVelocity = kC_Light;
constexpr double kMass = (kVelocity * kVelocity * kDist) / kGravity;
constexpr double kTime = kDist / kVelocity;
constexpr double kDensity = kMass / (kDist * kDist * kDist);
constexpr double kPressure = kMass / (kDist * kTime * kTime);
constexpr double kRadiation = kMass / (kTime * kTime * kTime);
constexpr double log_kDensity = 4.904385174253596;
constexpr double log_kStefanBoltzmann4 = -8.391375953193897;
constexpr double log_kDist = 29.93360620892259;
}  // namespace radc

// GetTemperature (lib/physics/include/gas_state.h:22-27): T in K of (rho, p) in code units.
__device__ __forceinline__ double gas_temperature(double rho, double prs) {
  constexpr double g = (radc::kM_hydrogen / radc::k_boltzmann) * radc::kPressure / radc::kDensity;
  return g * prs / rho;
}
// B_Plank(T) (lib/physics/src/plunk.cpp:87-90): sigma T^4 / pi.
__device__ __forceinline__ double b_planck(double T) {
  const double T2 = T * T;
  return (T2 * T2) * radc::kStefanBoltzmann / radc::kPI;
}

// The common tail of GetRadSource and GetRadSourceOpt (radRHD_utils.cpp:
// 74-100 and 131-157) once gamma, gamma^2 and u^2 are known.  D(j, i) is the
// M1 Eddington tensor EddTensor(j, i, Ur, Fr) (ISOTROPIC_INTENSITY == 1), or,
// when the caller passes the radiation pressure tensor Tr (row-major,
// Tr(j, i) = Tr9[3 j + i]), D(j, i) = Tr(j, i) as the reference is compiled
// (radRHD_config.h:8, ISOTROPIC_INTENSITY 0; radRHD_utils.cpp:80-86, 137-143),
// multiplied by Ur below as there.
//
// The mixed path (A = Fast) with the M1 tensor evaluates the same source in a
// compact form.  With D = c n n^T + dg I and n = F/|F|, u.D = cf (u.F) F + dg u
// and u.D.u = cf (u.F)^2 + dg u.u, cf = c/|F|^2, so no unit vector, |F| or
// tensor is formed, and each spatial component is G_i = a u_i + b F_i with
//   a = q - rho Ur (k_t dg + k_s (gamma^2 + uuD)) + 2 rho gamma k_s (u.F),
//   b = rho k_t (gamma - Ur cf (u.F)),
// the reference's terms regrouped.  The closure's cases compare squares:
// F.F < 1e-150 (as the reference does) or Ur < 1e-150: cf = 0, dg =
// 0.33333333333333; |F|^2 > Ur^2: cf = 1/|F|^2, dg = 0; else chi of
// f^2 = |F|^2 / Ur^2 as EddTensor has it.
template <typename A>
__device__ __forceinline__ void rad_source_tail(const State<double> &W, double Ur, const double *Fr, double abs_op,
                                                double scat_op, double q, double gamma, double gamma2, double u2,
                                                double *G, const float *Tr9 = nullptr) {
  const double tot_op = abs_op + scat_op;
  const double rho = W.d;
  const double u[3] = {gamma * W.v[0], gamma * W.v[1], gamma * W.v[2]};
  const double uF = A::kFast ? dot3c(u, Fr) : dot3(u, Fr);
  if (A::kFast && !Tr9) {
#pragma clang fp contract(fast)
    using B = FaceFast;  // one Newton step on each hardware estimate (4.2e-15)
    const double fm2 = dot3c(Fr, Fr);
    double cf = 0.0, dg = 0.33333333333333;
    if (!(fm2 < 1e-150 || Ur < 1e-150)) {
      const double ifm2 = B::rcp(fm2);
      const double U2 = Ur * Ur;
      if (fm2 > U2) {
        cf = ifm2, dg = 0.0;
      } else {
        const double f2 = fm2 * B::rcp(U2);
        const double chi_2 = B::div(1.5 + 2.0 * f2, 5.0 + 2.0 * B::sqrt_(4.0 - 3.0 * f2));
        cf = (3.0 * chi_2 - 0.5) * ifm2, dg = 0.5 - chi_2;
      }
    }
    const double cuF = cf * uF;
    const double uuD = cuF * uF + dg * u2;
    const double rhogamma = rho * gamma;
    const double rUr = rho * Ur;
    G[0] = q * gamma + rhogamma * (abs_op - scat_op * (u2 + uuD)) * Ur - rho * (abs_op - scat_op * (u2 + gamma2)) * uF;
    const double a = q - rUr * (tot_op * dg + scat_op * (gamma2 + uuD)) + 2.0 * rhogamma * scat_op * uF;
    const double b = rho * tot_op * (gamma - Ur * cuF);
#pragma unroll
    for (int i = 0; i < 3; i++) G[i + 1] = a * u[i] + b * Fr[i];
    return;
  }
  double uuD = 0.;
  double uD[3];
  double Dm[3][3];
  if (Tr9) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) Dm[a][b] = (double)Tr9[3 * a + b];
  } else {
    edd_matrix<A>(Ur, Fr, Dm);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    uD[i] = 0.;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double D = Dm[j][i];  // EddTensor(j, i, Ur, Fr) or Tr(j, i)
      uD[i] += u[j] * D;
      uuD += u[i] * u[j] * D;
    }
  }
  const double rhogamma = rho * gamma;
  G[0] = q * gamma + rhogamma * (abs_op - scat_op * (u2 + uuD)) * Ur - rho * (abs_op - scat_op * (u2 + gamma2)) * uF;
#pragma unroll
  for (int i = 0; i < 3; i++)
    G[i + 1] = q * u[i] - rho * (tot_op * uD[i] + scat_op * u[i] * (gamma2 + uuD)) * Ur +
               rhogamma * (tot_op * Fr[i] + scat_op * 2.0 * uF * u[i]);
}

// GetRadSource (radRHD_utils.cpp:44-101): given opacities; the thermal
// emission q = -kappa_a rho B(T(rho, p)) / kRadiation when `emission`.
template <typename A = Ieee>
__device__ __forceinline__ void rad_source(const State<double> &W, double Ur, const double *Fr, double abs_op,
                                           double scat_op, double *G, bool emission = false,
                                           const float *Tr9 = nullptr) {
  double gamma = A::kFast ? dot3c(W.v, W.v) : dot3(W.v, W.v);
  double u2 = gamma;
  double gamma2;
  if (!A::kFast) {
    gamma2 = A::rcp(1.0 - gamma);
    u2 *= gamma2;
    gamma = A::sqrt_(gamma2);
  } else {  // gamma from one reciprocal root
    gamma = A::rsqrt(1.0 - gamma);
    gamma2 = gamma * gamma;
    u2 *= gamma2;
  }
  const double q = emission ? -abs_op * W.d * b_planck(gas_temperature(W.d, W.p)) / radc::kRadiation : 0.0;
  rad_source_tail<A>(W, Ur, Fr, abs_op, scat_op, q, gamma, gamma2, u2, G, Tr9);
}

// The reference's TableFunc (lib/grids/include/solvers_struct.h:24-44) over
// (ln D, ln T): ln Lambda at [i*ny + j].
struct OpacityTable {
  const double *data;
  int nx, ny;
  double min_x, max_x, step_x, min_y, max_y, step_y;
};

// TableFunc::operator() (lib/grids/src/solvers_struct.cpp:10-58): the
// arguments are clamped below and x above; the node is taken at round(x),
// round(y) (nearest integers, then truncated to a node index); above max_y the
// last two nodes of the row are extended linearly in exp(y).  The indices are
// also clamped at 0, which the reference leaves to its tables (round(min) >= min).
__device__ __forceinline__ double table_lookup(const OpacityTable &t, double x, double y) {
  if (x < t.min_x) x = t.min_x;
  if (y < t.min_y) y = t.min_y;
  if (x > t.max_x) x = t.max_x;
  int i = (int)((round(x) - t.min_x) / t.step_x);
  i = i < t.nx - 1 ? i : t.nx - 1;
  i = i > 0 ? i : 0;
  if (y > t.max_y) {
    const double y1 = exp(t.data[t.ny * i + (t.ny - 1)]);
    const double y0 = exp(t.data[t.ny * i + (t.ny - 2)]);
    const double x1 = exp(t.max_y);
    const double x0 = exp(t.max_y - t.step_y);
    const double k = (y0 - y1) / (x0 - x1);
    const double m = (-x1 * y0 + x0 * y1) / (x0 - x1);
    return log(k * exp(y) + m);
  }
  int j = (int)((round(y) - t.min_y) / t.step_y);
  j = j < t.ny - 1 ? j : t.ny - 1;
  j = j > 0 ? j : 0;
  return t.data[t.ny * i + j];
}

// full_phys_data_t::Init(const flux_t *) (lib/grids/src/solvers_struct.cpp:264-281):
// temperature, speed, Lorentz factor and the opacities of a cell's state.
struct CellRad {
  double T, vel, lorenz, alpha, betta;
};
template <typename A = Ieee>
__device__ __forceinline__ CellRad cell_rad(const OpacityTable &t, const State<double> &W) {
  CellRad c;
  c.T = gas_temperature(W.d, W.p);
  const double logT = log(c.T);
  const double vel2 = dot3(W.v, W.v);
  c.vel = A::sqrt_(vel2);
  c.lorenz = A::rcp(A::sqrt_(1. - vel2));
  const double L = table_lookup(t, log(W.d) + radc::log_kDensity, logT);
  const double log_alpha = L - (radc::log_kStefanBoltzmann4 + 4 * logT) + radc::log_kDist;
  c.alpha = exp(log_alpha);
  c.betta = (radc::kSigma_thomson / radc::kM_hydrogen * radc::kDist) * W.d * radc::kDensity;
  return c;
}

// GetRadSourceOpt (radRHD_utils.cpp:103-158): the cell's own opacities and
// temperature, gamma from its Lorentz factor, thermal emission always on.
template <typename A = Ieee>
__device__ __forceinline__ void rad_source_opt(const State<double> &W, const CellRad &c, double Ur, const double *Fr,
                                               double *G, const float *Tr9 = nullptr) {
  const double gamma = c.lorenz;
  const double gamma2 = gamma * gamma;
  const double u2 = gamma2 * c.vel * c.vel;
  const double q = -c.alpha * W.d * b_planck(c.T) / radc::kRadiation;
  rad_source_tail<A>(W, Ur, Fr, c.alpha, c.betta, q, gamma, gamma2, u2, G, Tr9);
}

}  // namespace rhllc_dev
