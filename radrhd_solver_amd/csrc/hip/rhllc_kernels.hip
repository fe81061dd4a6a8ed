// gfx950 kernels of the RHLLC step.
//
// One step = three launches on one stream (solvers/rhllc/src/rhllc_3d_stab.cpp:10-65
// plus rhllc_utils.cpp:15-23):
//   k_face  one thread per face: gather left/right states, HLLC flux, write
//           the face flux, fold the signal speed into per-slot maxima;
//   k_cell  one thread per cell: gather its 4 face fluxes in slot order with
//           their signs, U -= tau/V * sum, c2p (Newton, FP64) with the
//           PhysPressureFix fallback, optional radiation source + second c2p;
//   k_dt    one workgroup: max over the slots, tau = CFL*h_min/speed, t += tau.
// No float atomics: face fluxes are written once and read by a fixed-order
// gather, so the step is deterministic; the speed maximum is order-free.
#include "rhllc_internal.h"
#include "flux_ext.h"
#include "mlp.h"
#include "rhllc_physics.h"

namespace rhllc_dev {

// Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a
// contiguous range of logical blocks so neighbouring faces/cells (close in the
// Hilbert numbering) are gathered through the same L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblk) {
  const uint32_t x = bid & 7u, k = bid >> 3;
  const uint32_t q = nblk >> 3, r = nblk & 7u;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = v < w ? w : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Element i of a device array, addressed as the array's (uniform) base plus a
// 32-bit byte offset: the compiler issues that as one SGPR base and one VGPR
// offset, where a 64-bit index needs four to five VALU instructions of
// address arithmetic per access.  Every array of a handle stays below 4 GiB:
// cells < 2^25 and faces < 2^27 (rhllc_create), elements of at most 32 bytes.
template <typename T>
__device__ __forceinline__ T ld(const T *base, uint32_t i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(T *base, uint32_t i, const T &v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T)) = v;
}

__device__ __forceinline__ void load_U(const Dev &d, int c, State<double> &U) {
  const double4 a = ld(d.U4, c);
  U.d = a.x, U.v[0] = a.y, U.v[1] = a.z, U.v[2] = a.w, U.p = ld(d.UE, c);
}
__device__ __forceinline__ void load_W(const Dev &d, int c, State<double> &W) {
  const double4 a = ld(d.W4, c);
  W.d = a.x, W.v[0] = a.y, W.v[1] = a.z, W.v[2] = a.w, W.p = ld(d.Wp, c);
}
// The FP32 primitives of the mixed path, widened (rho = D sqrt(1 - v^2)).
__device__ __forceinline__ void load_W32(const Dev &d, int c, State<double> &W) {
  const MixedIn r = mixed_in(ld(d.Wva, c), ld(d.Wsg, c));
  W.v[0] = r.v[0], W.v[1] = r.v[1], W.v[2] = r.v[2], W.p = r.p;
  W.d = r.D * sqrt(1.0 - dot3(W.v, W.v));
}
template <int EOS>
__device__ __forceinline__ void store_W32(const Dev &d, float4 *va_out, double2 *sg_out, int c,
                                          const State<double> &W) {
  float4 va;
  double2 sg;
  mixed_record<EOS>(d.Gg, W, va, sg);
  st(va_out, c, va);
  st(sg_out, c, sg);
}

}  // namespace rhllc_dev
// Wavefront reduction of the ROCm device library (ockl), declared by the HIP
// headers only under HIP_ENABLE_EXTRA_WARP_SYNC_TYPES.
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(unsigned long long);
namespace rhllc_dev {

// Per-block max of non-negative speeds into one of kSpeedSlots slots.
// Non-negative doubles order like their bit patterns (a NaN above all), so the
// maximum is an exact unsigned 64-bit maximum.  Every thread leaves its speed
// in LDS; after the barrier the first wave alone takes the maximum of four
// each and the DPP reduction of the ROCm device library, so the other waves
// skip the cross-lane steps.
__device__ __forceinline__ void fold_speed(const Dev &d, double speed) {
  __shared__ unsigned long long part[kBlock];
  part[threadIdx.x] = (unsigned long long)__double_as_longlong(speed);
  __syncthreads();
  if (threadIdx.x < 64) {
    unsigned long long m = part[threadIdx.x];
#pragma unroll
    for (int k = 1; k < kBlock / 64; ++k) {
      const unsigned long long v = part[threadIdx.x + 64 * k];
      m = m < v ? v : m;
    }
    m = __ockl_wfred_max_u64(m);
    if (threadIdx.x == 0) {
      unsigned long long *slot =
          reinterpret_cast<unsigned long long *>(d.speed_slots + (size_t)(blockIdx.x % kSpeedSlots) * kSlotStride);
      atomicMax(slot, m);
    }
  }
}

// The mixed path's face vector S n from its 16-byte form: the FP32 vector
// plus k Q per component, k the signed 7-bit fields (bits 0-20) of the fourth
// word and Q = 2^(E - 30), 2^E the power of two at or below the largest
// |component|.  k Q is exact and so is the sum (at most 31 significant bits),
// so this is the vector the host encoded, to 2^-30 of its largest component.
__device__ __forceinline__ void unpack_sn(const float4 &g, double s[3]) {
  const float m = fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fabsf(g.z));
  const double Q = (double)__int_as_float(__float_as_int(m) & 0x7F800000) * 0x1p-30;
  const int w = __float_as_int(g.w);
  s[0] = fma((double)((w << 25) >> 25), Q, (double)g.x);
  s[1] = fma((double)((w << 18) >> 25), Q, (double)g.y);
  s[2] = fma((double)((w << 11) >> 25), Q, (double)g.z);
}

// What a face reads before its cells: the two cell indices and the geometry
// (FP64 path: n and S; mixed path: the vector S n in its 16-byte form).
template <bool MIXED>
struct FaceIn {
  int2 lr;
  typename std::conditional<MIXED, float4, double4>::type g;
};
template <bool MIXED>
__device__ __forceinline__ FaceIn<MIXED> face_in(const Dev &d, int i) {
  FaceIn<MIXED> o;
  o.lr = ld(d.face_lr, i);
  if constexpr (MIXED)
    o.g = ld(d.face_Sn, i);
  else
    o.g = ld(d.face_nS, i);
  return o;
}
// Entry i of the tiled step's face arrays, a face of tile t.  On the mixed
// path the cell in the tile is bits 21-28 of the vector's fourth word (bit 29
// set when it is the right cell) and face_far holds the other side.
template <bool MIXED>
__device__ __forceinline__ FaceIn<MIXED> tile_face_in(const Dev &d, int t, int i) {
  if constexpr (!MIXED) {
    return face_in<MIXED>(d, i);
  } else {
    FaceIn<MIXED> o;
    o.g = ld(d.face_Sn, i);
    const int far = ld(d.face_far, i);
    const unsigned int w = (unsigned int)__float_as_int(o.g.w);
    const int nc = t * kTileCells + (int)((w >> 21) & 0xFFu);
    const bool right = (w >> 29) & 1u;
    o.lr = right ? make_int2(far, nc) : make_int2(nc, far);
    return o;
  }
}

// One face: gather of the two cells' state at the start of the step, HLLC
// flux into F (S * flux in the global frame; on the mixed path F.p is the
// flux of tau = E - D); returns max(lL, lR).
template <bool MIXED, int EOS>
__device__ __forceinline__ double face_flux(const Dev &d, const FaceIn<MIXED> &fi, State<double> &F) {
  const int2 lr = fi.lr;
  if constexpr (!MIXED) {
    const double4 nS = fi.g;
    // lr.y: the neighbour, a ghost state (n_cells, n_cells + 1) or ~lr.x on a wall
    const int r = lr.y >= 0 ? lr.y : ~lr.y;
    State<double> UL, WL, UR, WR;
    load_U(d, lr.x, UL);
    load_W(d, lr.x, WL);
    load_U(d, r, UR);
    load_W(d, r, WR);
    if (__builtin_expect(lr.y < 0, 0)) {
      wall_mirror(nS.x, nS.y, nS.z, UR.v);
      wall_mirror(nS.x, nS.y, nS.z, WR.v);
    }
    return hllc_flux<double, Ieee, EOS>(d.gamma, d.Gg, UL, UR, WL, WR, nS.x, nS.y, nS.z, nS.w, F);
  } else {
    // FP32 cell records, FP64 geometry and arithmetic with the fast
    // divide/sqrt policy.
    double Sn[3];
    unpack_sn(fi.g, Sn);
    const int r = lr.y >= 0 ? lr.y : ~lr.y;  // the neighbour, a ghost state, or the cell on a wall
    const MixedIn WL = mixed_in(ld(d.Wva, lr.x), ld(d.Wsg, lr.x));
    const MixedIn WR = mixed_in(ld(d.Wva, r), ld(d.Wsg, r));
    return mixed_face_flux<EOS>(d.gamma, WL, WR, Sn[0], Sn[1], Sn[2], F, lr.y < 0);
  }
}

// Face flux stored for the two-kernel step: FP64 (D, m) + E on the FP64
// path; the packed 24-byte record (flux_ext.h) on the mixed path.
template <bool MIXED, int EOS>
__device__ __forceinline__ double face_flux(const Dev &d, int i, State<double> &F) {
  return face_flux<MIXED, EOS>(d, face_in<MIXED>(d, i), F);
}

template <bool MIXED, int EOS>
__device__ __forceinline__ double face_flux_store(const Dev &d, int i) {
  State<double> F;
  const double s = face_flux<MIXED, EOS>(d, i, F);
  const double speed = smax(0.0, s);  // rhllc_3d_stab.cpp:18 -- per-thread max starts at 0
  if (!MIXED) {
    st(d.F4, i, make_double4(F.d, F.v[0], F.v[1], F.v[2]));
    st(d.FE, i, F.p);
  } else {
    const rhllc_ext::Packed pk = rhllc_ext::pack(F.d, F.v[0], F.v[1], F.v[2], F.p);
    st(d.F4f, i, pk.a);
    st(d.FXf, i, pk.b);
  }
  return speed;
}

// kFacesPerThread faces per thread, block-strided so loads and stores stay
// coalesced.
template <bool MIXED, int EOS>
__global__ __launch_bounds__(kBlock) void k_face(Dev d) {
  const int base = (int)(xcd_block(blockIdx.x, gridDim.x) * (kBlock * kFacesPerThread) + threadIdx.x);
  double speed = 0.0;
#pragma unroll
  for (int k = 0; k < kFacesPerThread; ++k) {
    const int i = base + k * kBlock;
    if (i < d.n_faces) speed = smax(speed, face_flux_store<MIXED, EOS>(d, i));
  }
  fold_speed(d, speed);
}

// The step's face fluxes in FP64 (energy slot = E flux) for a download; on
// the mixed path the same arithmetic as the step, E flux = D flux + tau flux.
template <bool MIXED, int EOS>
__global__ __launch_bounds__(kBlock) void k_face_fp64_out(Dev d, double4 *F4, double *FE) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n_faces) return;
  State<double> F;
  (void)face_flux<MIXED, EOS>(d, i, F);
  F4[i] = make_double4(F.d, F.v[0], F.v[1], F.v[2]);
  FE[i] = MIXED ? F.d + F.p : F.p;
}

template <bool MIXED>
__device__ __forceinline__ void load_flux(const Dev &d, int f, State<double> &F) {
  if (!MIXED) {
    const double4 a = ld(d.F4, f);
    F.d = a.x, F.v[0] = a.y, F.v[1] = a.z, F.v[2] = a.w, F.p = ld(d.FE, f);
  } else {
    double o[5];
    rhllc_ext::unpack(ld(d.F4f, f), ld(d.FXf, f), o);
    F.d = o[0], F.v[0] = o[1], F.v[1] = o[2], F.v[2] = o[3], F.p = o[4];
  }
}

// Counters of rare events: one atomic per wave that saw the event.
__device__ __forceinline__ void count_if(unsigned long long *ctr, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(ctr, (unsigned long long)__popcll(m));
}

// Surrogate start: every lane of the wave takes part (MFMA), dead lanes with
// a dummy state; returns the predicted pressure of this lane's cell.
__device__ __forceinline__ double mlp_guess(const Dev &d, bool live, const State<double> &U) {
  using namespace rhllc_mlp;
  const float *norm = d.mlpP + kOffNorm;
  float xl[3] = {0.f, 0.f, 0.f};
  double p0 = 0.0;
  if (live) features(norm, d.Gg, U.d, sqrt(dot3(U.v, U.v)), U.p, xl, p0);
  const int l = threadIdx.x & 63;
  float x[4][3];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int f = 0; f < 3; ++f) x[t][f] = __shfl(xl[f], 16 * t + (l & 15), 64);
  float out[4];
  forward_wave(d.mlpP, d.mlpW2b, x, out);
  return predicted_pressure(norm, out[l >> 4], p0);
}

// Surrogate start only where the previous pressure is stale: the cell's
// tau = E - D changed by more than this fraction in the step.  Measured on
// the same states (a 48^3-hex blast, steps 1-60): where tau moved by 5-20%,
// Newton takes 2.0-2.5 iterations from the surrogate and 3.0-3.4 from the
// previous pressure, above 20% 2.1-2.6 against 3.7-4.9; between 2% and 5%
// the two are within a quarter iteration either way, and below 0.5% the
// previous pressure converges in one iteration where the surrogate needs two.
constexpr double kMlpGate = 0.05;

// Where a cell's four face fluxes come from.  GlobalFlux: the arrays the
// face kernel of the two-kernel step wrote (face index | bit 31 when the cell
// is id_r).  TileFlux: the LDS of the tiled step, FP64 (D, m0), (m1, m2), E
// per slot (16-bit slot | bit 15 when the cell is id_r); on the mixed path the
// energy slot holds the tau = E - D flux, so E's flux is D's plus tau's, as
// the packed record's unpack gives it.
template <bool MIXED>
struct GlobalFlux {
  static constexpr bool kTau = false;  // F.p is the E flux
  const Dev &d;
  int4 cf;
  __device__ __forceinline__ bool neg(int j) const { return (&cf.x)[j] < 0; }
  __device__ __forceinline__ void load(int j, State<double> &F) const {
    load_flux<MIXED>(d, (&cf.x)[j] & 0x7FFFFFFF, F);
  }
};
template <bool MIXED>
struct TileFlux {
  // Mixed path: F.p is the flux of tau = E - D; the cell sums it apart and
  // adds the D sum once (E flux = D flux + tau flux).
  static constexpr bool kTau = MIXED;
  const double2 *A, *B;
  const double *C;
  uint32_t sv[4];
  __device__ __forceinline__ bool neg(int j) const { return (sv[j] & 0x8000u) != 0; }
  __device__ __forceinline__ void load(int j, State<double> &F) const {
    const uint32_t k = sv[j] & 0x7FFFu;
    const double2 a = A[k], b = B[k];
    F.d = a.x, F.v[0] = a.y, F.v[1] = b.x, F.v[2] = b.y;
    F.p = C[k];
  }
};

// Gather of the four face fluxes in slot order with their signs
// (rhllc_3d_stab.cpp:30-45) and U -= tau/V * sum.  Also returns the change of
// tau = E - D (for the surrogate gate) and, on the mixed path, the cell's
// FP32 pressure (the c2p start).  Both paths add each face's flux to its two
// cells with opposite signs, so the totals change only by the boundary fluxes
// (to rounding).
template <bool MIXED, typename Src>
__device__ __forceinline__ void update_U(const Dev &d, int c, double tau, const Src &src, State<double> &U,
                                         double &dtau, float &p32) {
  State<double> s;
  s.d = 0.0, s.v[0] = 0.0, s.v[1] = 0.0, s.v[2] = 0.0, s.p = 0.0;
  // s +- F as fma(+-1, F, s): the product is exact, so each sum is rounded
  // once, as the reference's += and -= are (signed zeros included), without
  // a branch on the sign.
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    State<double> F;
    src.load(j, F);
    const double sg = src.neg(j) ? -1.0 : 1.0;
    s.d = fma(sg, F.d, s.d), s.v[0] = fma(sg, F.v[0], s.v[0]), s.v[1] = fma(sg, F.v[1], s.v[1]);
    s.v[2] = fma(sg, F.v[2], s.v[2]), s.p = fma(sg, F.p, s.p);
  }
  if constexpr (Src::kTau) s.p += s.d;
  load_U(d, c, U);
  double k;
  if (MIXED) {
    p32 = ld(d.Wva, c).w;
    k = tau * ld(d.cell_Vinv, c);  // V k = tau to rounding: the totals of V U telescope
  } else {
    k = tau / ld(d.cell_V, c);
  }
  s.d *= k, s.v[0] *= k, s.v[1] *= k, s.v[2] *= k, s.p *= k;
  U.d -= s.d, U.v[0] -= s.v[0], U.v[1] -= s.v[1], U.v[2] -= s.v[2], U.p -= s.p;
  dtau = s.p - s.d;
}

// The radiation source of one cell and the c2p after it, from the cell's
// state W after the hydro update; U and W are updated in place.
//   radRHD_main.cpp:117-132: U += (ds*tau)*G(W), ds = 1, then c2p (code ignored).
//   With the opacity table: AddRadFlux (rhllc_3d_mpi.cpp:509-533), GetRadSourceOpt
//   with the cell's full_phys_data_t::Init of W, then HllcConvToPhys
//   (rhllc_3d_mpi.cpp:488-505: c2p, PhysPressureFix on failure).
template <bool MIXED, int EOS>
__device__ __forceinline__ void apply_rad(const Dev &d, int c, double tau, State<double> &U, State<double> &W,
                                          int &iters, bool &fixed, bool &fix_failed) {
  using A = typename std::conditional<MIXED, Fast, Ieee>::type;
  const float4 fu = ld(d.rad_FU, c);
  const double Fr[3] = {(double)fu.x, (double)fu.y, (double)fu.z};
  const float *Tr9 = d.rad_T ? d.rad_T + 9 * (size_t)c : nullptr;
  double G[4];
  if (d.tab) {
    const OpacityTable t{d.tab,        d.tab_nx,    d.tab_ny,    d.tab_min_x, d.tab_max_x,
                         d.tab_step_x, d.tab_min_y, d.tab_max_y, d.tab_step_y};
    rad_source_opt<A>(W, cell_rad<A>(t, W), (double)fu.w, Fr, G, Tr9);
  } else {
    const float2 kk = ld(d.rad_k, c);
    rad_source<A>(W, (double)fu.w, Fr, (double)kk.x, (double)kk.y, G, d.rad_emission != 0, Tr9);
  }
  const double ds = 1;
  U.p += ds * tau * G[0];
  U.v[0] += ds * tau * G[1];
  U.v[1] += ds * tau * G[2];
  U.v[2] += ds * tau * G[3];
  int it = 0;
  int code = c2p<A, EOS>(d.Gg, U, W, MIXED ? W.p : -1.0, it);
  iters += it;
  if (d.tab && code) {
    // HllcConvToPhys: the reference start, then PhysPressureFix of the state
    // that c2p left (the pre-source W).
    if (MIXED) {
      it = 0;
      code = c2p<A, EOS>(d.Gg, U, W, -1.0, it);
      iters += it;
    }
    if (code) {
      fixed = true;
      fix_failed = fix_failed || pressure_fix<EOS>(d.Gg, U, W) != 0;
    }
  }
}

// Everything after the update for one cell: c2p from `guess` with the
// fallback chain, PhysPressureFix, the optional radiation source with its
// second c2p, the stores and the counters.  Called by all lanes of a wave
// (ballots) and, with STATS, by all threads of the block (histogram).
//   first fallback: the previous pressure (when the start was the surrogate's)
//   then the reference start (all variants but kFp64Ref), then PhysPressureFix
template <int VARIANT, bool RAD, bool STATS, int EOS>
__device__ __forceinline__ void finish_cell(const Dev &d, int c, bool live, State<double> U, double guess,
                                            bool from_mlp, double tau) {
  constexpr bool MIXED = VARIANT >= kMixedPrev;
  using A = typename std::conditional<MIXED, Fast, Ieee>::type;
  int iters = 0, code = 0;
  bool fell_back = false, fixed = false, fix_failed = false;
  double resid = 0.0;
  State<double> W;
  if (live) {
    // c2p writes W only when it succeeds.
    code = c2p<A, EOS>(d.Gg, U, W, guess, iters);
    if (code && from_mlp) {
      int it2 = 0;
      fell_back = true;
      code = c2p<A, EOS>(d.Gg, U, W, (double)ld(d.Wva, c).w, it2);
      iters += it2;
    }
    if (code && VARIANT != kFp64Ref) {
      int it2 = 0;
      fell_back = true;
      code = c2p<A, EOS>(d.Gg, U, W, -1.0, it2);
      iters += it2;
    }
    if (code) {
      // PhysPressureFix receives the cell's previous phys_val (rhllc_3d_stab.cpp:49-52).
      if (!MIXED)
        load_W(d, c, W);
      else
        load_W32(d, c, W);
      fixed = true;
      fix_failed = pressure_fix<EOS>(d.Gg, U, W) != 0;
    }
    if (RAD) apply_rad<MIXED, EOS>(d, c, tau, U, W, iters, fixed, fix_failed);
    if (STATS) resid = c2p_residual<EOS>(d.Gg, U, W.p);
    st(d.U4o, c, make_double4(U.d, U.v[0], U.v[1], U.v[2]));
    st(d.UEo, c, U.p);
    if (!MIXED) {
      st(d.W4o, c, make_double4(W.d, W.v[0], W.v[1], W.v[2]));
      st(d.Wpo, c, W.p);
    } else {
      store_W32<EOS>(d, d.Wvao, d.Wsgo, c, W);
    }
  }
  if (__ballot(live && (code != 0 || fixed || fell_back)) != 0) {  // rare: skip eight ballots otherwise
    count_if(&d.sc->c2p_fail[1], live && code == 1);
    count_if(&d.sc->c2p_fail[2], live && code == 2);
    count_if(&d.sc->c2p_fail[3], live && code == 3);
    count_if(&d.sc->c2p_fail[4], live && code == 4);
    count_if(&d.sc->c2p_fail[5], live && code == 5);
    count_if(&d.sc->fix_calls, fixed);
    count_if(&d.sc->fix_fail, fix_failed);
    count_if(&d.sc->fallback, fell_back);
  }
  if (STATS) {
    __shared__ unsigned int hist[kHistBins];
    __shared__ unsigned long long isum[kBlock / 64];
    __shared__ double rmax[kBlock / 64];
    for (int b = threadIdx.x; b < kHistBins; b += kBlock) hist[b] = 0;
    __syncthreads();
    if (live) atomicAdd(&hist[iters < kHistBins ? iters : kHistBins - 1], 1u);
    const unsigned long long ws = wave_sum_u64(live ? (unsigned long long)iters : 0ull);
    const double wr = wave_max(resid);
    if ((threadIdx.x & 63) == 0) isum[threadIdx.x >> 6] = ws, rmax[threadIdx.x >> 6] = wr;
    __syncthreads();
    for (int b = threadIdx.x; b < kHistBins; b += kBlock)
      if (hist[b]) atomicAdd(&d.sc->hist[b], (unsigned long long)hist[b]);
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      double r = 0.0;
      for (int k = 0; k < kBlock / 64; ++k) t += isum[k], r = r < rmax[k] ? rmax[k] : r;
      atomicAdd(&d.sc->newton_iters, t);
      atomicMax(&d.sc->max_residual_bits, (unsigned long long)__double_as_longlong(r));
    }
    __syncthreads();  // the shared arrays are reused if the caller loops
  }
}

// The surrogate-start variant defers cells whose tau changed by more than
// kMlpGate: they are appended to d.mlp_list (one atomic per wave), their
// updated U is stored, and k_cell_mlp finishes them on full waves.  Returns
// whether this lane's cell stays live here.  Called by all lanes of a wave.
__device__ __forceinline__ bool defer_stale(const Dev &d, int c, bool live, const State<double> &U, double dtau) {
  const bool stale = live && !(fabs(dtau) <= kMlpGate * fabs(U.p - U.d));
  const unsigned long long m = __ballot(stale);
  if (m) {
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if (lane == first) base = atomicAdd(&d.sc->mlp_count, (unsigned int)__popcll(m));
    base = __shfl(base, first, 64);
    if (stale) {
      st(d.mlp_list, base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)), c);
      st(d.U4o, c, make_double4(U.d, U.v[0], U.v[1], U.v[2]));
      st(d.UEo, c, U.p);
      return false;
    }
  }
  return live;
}

// One thread per cell.  Start of the Newton iteration: the reference start
// (kFp64Ref), the FP64 previous pressure (kFp64Prev) or the FP32 one (mixed).
// kMixedMlp: stale cells go to k_cell_mlp (defer_stale).
template <int VARIANT, bool RAD, bool STATS, int EOS>
__global__ __launch_bounds__(kBlock) void k_cell(Dev d) {
  constexpr bool MIXED = VARIANT >= kMixedPrev;
  const int c = (int)(xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x);
  bool live = c < d.n_cells;
  const double tau = d.sc->tau;
  State<double> U;
  double dtau = 0.0;
  float p32 = -1.f;
  if (live) {
    const GlobalFlux<MIXED> src{d, ld(d.cell_faces, c)};
    update_U<MIXED>(d, c, tau, src, U, dtau, p32);
  }
  if (VARIANT == kMixedMlp) live = defer_stale(d, c, live, U, dtau);
  double guess = -1.0;
  if (VARIANT == kFp64Prev && live) guess = ld(d.Wp, c);
  if (MIXED && live) guess = (double)p32;
  finish_cell<VARIANT, RAD, STATS, EOS>(d, c, live, U, guess, false, tau);
}

// The tiled step, one workgroup per tile of kTileCells cells (Hilbert order).
// Face phase: every face of the tile's cells, own faces then extra faces,
// block-strided; the flux goes to LDS in FP64 and the speed into the per-step
// maximum.  Cell phase, after the barrier of the speed fold: one thread per
// cell sums its four fluxes from LDS in slot order with their signs and
// finishes the cell exactly as k_cell does.  The state the face phase gathers
// is the step's input buffer; the cell phase writes the output buffer, so no
// tile reads what another tile of the same step writes.  A face shared by two
// tiles is evaluated by both from the same inputs, to the same bits.
// At most 96 VGPRs (five waves a SIMD): without the floor the compiler takes
// 116 for the mixed variants and the kernel runs at four waves.
template <int VARIANT, bool RAD, bool STATS, int EOS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_tile(Dev d) {
  static_assert(kTileCells == kBlock, "one cell per thread in the cell phase");
  constexpr bool MIXED = VARIANT >= kMixedPrev;
  extern __shared__ double2 lds_flux[];
  // Every other step walks the tiles backwards, so a step starts on the tiles
  // (and the arrays' ends) the previous one touched last, which the
  // memory-side cache still holds.
  const int tx = (int)xcd_block(blockIdx.x, gridDim.x);
  const int t = d.tile_reverse ? (int)gridDim.x - 1 - tx : tx;
  const int4 tm = ld(d.tile_meta, t);
  const int nf = tm.y + tm.w;
  double2 *A = lds_flux, *B = lds_flux + nf;
  double *C = reinterpret_cast<double *>(lds_flux + 2 * nf);
  double speed = 0.0;
  for (int k = threadIdx.x; k < nf; k += kBlock) {
    // One index into one set of arrays (the extra faces' copies follow the
    // faces): a select between two sets of base pointers would hold them in
    // VGPRs through the loop.
    const FaceIn<MIXED> fi = tile_face_in<MIXED>(d, t, k < tm.y ? tm.x + k : d.n_faces + tm.z + (k - tm.y));
    State<double> F;
    speed = smax(speed, face_flux<MIXED, EOS>(d, fi, F));
    A[k] = make_double2(F.d, F.v[0]);
    B[k] = make_double2(F.v[1], F.v[2]);
    C[k] = F.p;
  }
  fold_speed(d, speed);  // its barrier also publishes the LDS fluxes
  const int c = t * kTileCells + (int)threadIdx.x;
  bool live = c < d.n_cells;
  const double tau = d.sc->tau;
  State<double> U;
  double dtau = 0.0;
  float p32 = -1.f;
  if (live) {
    const uint2 sl = ld(d.cell_slots, c);
    const TileFlux<MIXED> src{A, B, C, {sl.x & 0xFFFFu, sl.x >> 16, sl.y & 0xFFFFu, sl.y >> 16}};
    update_U<MIXED>(d, c, tau, src, U, dtau, p32);
  }
  if (VARIANT == kMixedMlp) live = defer_stale(d, c, live, U, dtau);
  double guess = -1.0;
  if (VARIANT == kFp64Prev && live) guess = ld(d.Wp, c);
  if (MIXED && live) guess = (double)p32;
  finish_cell<VARIANT, RAD, STATS, EOS>(d, c, live, U, guess, false, tau);
}

// Cells deferred by k_cell<kMixedMlp>: list entries [0, mlp_count), 64 per
// wave so the surrogate's MFMA pass runs on full waves.  A fixed grid loops
// over the list (the count is only known on the device); every block runs
// the same number of iterations, and all of them end when the list does.
// The surrogate is trained for the ideal gas only.
template <bool RAD, bool STATS>
__global__ __launch_bounds__(kBlock) void k_cell_mlp(Dev d) {
  const unsigned int n = d.sc->mlp_count;
  const double tau = d.sc->tau;
  for (unsigned int b0 = blockIdx.x * kBlock; b0 < n; b0 += gridDim.x * kBlock) {
    const unsigned int i = b0 + threadIdx.x;
    const bool live = i < n;
    const int c = live ? ld(d.mlp_list, i) : 0;
    State<double> U;
    U.d = 1.0, U.v[0] = 0.0, U.v[1] = 0.0, U.v[2] = 0.0, U.p = 2.0;  // dummy lanes
    if (live) load_U(d, c, U);
    const double g = mlp_guess(d, live, U);
    finish_cell<kMixedMlp, RAD, STATS, kIdeal>(d, c, live, U, g, true, tau);
  }
}

// tau of the next step from the max signal speed (rhllc_utils.cpp:15-23):
// tau = CFL*h_min / speed.  Also clears the slots for the next step.
__global__ __launch_bounds__(1024) void k_dt(Dev d) {
  __shared__ double part[1024 / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < kSpeedSlots; i += 1024) {
    double *slot = d.speed_slots + (size_t)i * kSlotStride;
    const double v = *slot;
    m = m < v ? v : m;
    *slot = 0.0;
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = part[0];
    for (int k = 1; k < 1024 / 64; ++k) s = s < part[k] ? part[k] : s;
    Scalars *sc = d.sc;
    sc->t += sc->tau;
    sc->last_speed = s;
    sc->tau = d.cfl_hmin / s;
    sc->steps += 1;
    sc->mlp_starts += sc->mlp_count;
    sc->mlp_count = 0;
  }
}

// HllcPhysToConv (rhllc_utils.cpp:25-33): U = p2c(W), W clamped in place.
template <int EOS>
__global__ __launch_bounds__(kBlock) void k_p2c_init(Dev d) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= d.n_cells) return;
  State<double> W, U;
  load_W(d, c, W);
  p2c<double, Ieee, EOS>(d.Gg, W, U);
  d.W4[c] = make_double4(W.d, W.v[0], W.v[1], W.v[2]);
  d.Wp[c] = W.p;
  d.U4[c] = make_double4(U.d, U.v[0], U.v[1], U.v[2]);
  d.UE[c] = U.p;
}

// Mixed path: the FP32 cell records from the FP64 primitives (after an
// upload or HllcPhysToConv).
template <int EOS>
__global__ __launch_bounds__(kBlock) void k_seed_mixed(Dev d) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= d.n_cells) return;
  State<double> W;
  load_W(d, c, W);
  store_W32<EOS>(d, d.Wva, d.Wsg, c, W);
}

// The ghost states of boundary codes -3 and -4 as entries n_cells and
// n_cells + 1 of the state arrays, input and output buffers alike (two threads).
template <int EOS>
__global__ void k_ghost_cell(Dev d, bool mixed) {
  const int k = threadIdx.x;
  if (k > 1) return;
  const int c = d.n_cells + k;
  const double *gw = k == 0 ? d.ghostW : d.ghost2W, *gu = k == 0 ? d.ghostU : d.ghost2U;
  State<double> W;
  W.d = gw[0], W.v[0] = gw[1], W.v[1] = gw[2], W.v[2] = gw[3], W.p = gw[4];
  if (mixed) {
    store_W32<EOS>(d, d.Wva, d.Wsg, c, W);
    store_W32<EOS>(d, d.Wvao, d.Wsgo, c, W);
    return;
  }
  const double4 u = make_double4(gu[0], gu[1], gu[2], gu[3]);
  const double4 w = make_double4(W.d, W.v[0], W.v[1], W.v[2]);
  d.U4[c] = u, d.UE[c] = gu[4], d.W4[c] = w, d.Wp[c] = W.p;
  d.U4o[c] = u, d.UEo[c] = gu[4], d.W4o[c] = w, d.Wpo[c] = W.p;
}

// Mixed path: FP64 primitives from the FP64 conserved state (for download).
template <int EOS>
__global__ __launch_bounds__(kBlock) void k_recover_W(Dev d) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= d.n_cells) return;
  State<double> U, W;
  load_U(d, c, U);
  load_W32(d, c, W);
  int it = 0;
  c2p<Ieee, EOS>(d.Gg, U, W, W.p, it);  // W keeps the FP32 values if c2p fails
  d.W4[c] = make_double4(W.d, W.v[0], W.v[1], W.v[2]);
  d.Wp[c] = W.p;
}

// The radiation source on its own, after a step without it: AddRadFlux then
// HllcConvToPhys (rhllc_3d_mpi.cpp:509-533, 488-505) for a caller that solves
// the transport between the hydro update and the source, as the reference's
// coupled loop does (radRHD_mpi_main.cpp:112-122).  The FP64 path reads the
// stored W; the mixed path first recovers the FP64 W from U (c2p from the
// stored FP32 pressure, as the fused step has it in registers).
template <bool MIXED, int EOS>
__global__ __launch_bounds__(kBlock) void k_add_rad(Dev d, double tau) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  const bool live = c < d.n_cells;
  bool fixed = false, fix_failed = false;
  if (live) {
    using A = typename std::conditional<MIXED, Fast, Ieee>::type;
    State<double> U, W;
    load_U(d, c, U);
    int iters = 0;
    if (!MIXED) {
      load_W(d, c, W);
    } else {
      load_W32(d, c, W);
      c2p<A, EOS>(d.Gg, U, W, W.p, iters);  // keeps the FP32 W if it fails
    }
    apply_rad<MIXED, EOS>(d, c, tau, U, W, iters, fixed, fix_failed);
    d.U4[c] = make_double4(U.d, U.v[0], U.v[1], U.v[2]);
    d.UE[c] = U.p;
    if (!MIXED) {
      d.W4[c] = make_double4(W.d, W.v[0], W.v[1], W.v[2]);
      d.Wp[c] = W.p;
    } else {
      store_W32<EOS>(d, d.Wva, d.Wsg, c, W);
    }
  }
  if (__ballot(fixed) != 0) {
    count_if(&d.sc->fix_calls, fixed);
    count_if(&d.sc->fix_fail, fix_failed);
  }
}

// Clears the counters and speed slots with plain stores (graph-safe).
__global__ void k_clear(Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kSpeedSlots) d.speed_slots[(size_t)i * kSlotStride] = 0.0;
  if (i == 0) {
    Scalars *sc = d.sc;
    sc->steps = 0;
    for (int k = 0; k < 6; ++k) sc->c2p_fail[k] = 0;
    sc->fix_calls = sc->fix_fail = sc->fallback = sc->newton_iters = sc->mlp_starts = 0;
    sc->mlp_count = 0;
    for (int k = 0; k < kHistBins; ++k) sc->hist[k] = 0;
    sc->max_residual_bits = 0;
  }
}

static inline unsigned int blocks_for(int64_t n) { return (unsigned int)((n + kBlock - 1) / kBlock); }
static inline unsigned int face_blocks(int64_t n) {
  return (unsigned int)((n + kBlock * kFacesPerThread - 1) / (kBlock * kFacesPerThread));
}

hipError_t launch_p2c_init(const Dev &d, hipStream_t s) {
  if (d.n_cells == 0) return hipSuccess;
  if (d.eos == kTaub)
    hipLaunchKernelGGL(k_p2c_init<kTaub>, dim3(blocks_for(d.n_cells)), dim3(kBlock), 0, s, d);
  else
    hipLaunchKernelGGL(k_p2c_init<kIdeal>, dim3(blocks_for(d.n_cells)), dim3(kBlock), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_face(const Dev &d, bool mixed, hipStream_t s) {
  if (d.n_faces == 0) return hipSuccess;
  const dim3 g(face_blocks(d.n_faces)), b(kBlock);
  if (d.eos == kTaub) {
    if (mixed)
      hipLaunchKernelGGL((k_face<true, kTaub>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((k_face<false, kTaub>), g, b, 0, s, d);
  } else {
    if (mixed)
      hipLaunchKernelGGL((k_face<true, kIdeal>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((k_face<false, kIdeal>), g, b, 0, s, d);
  }
  return hipGetLastError();
}

template <int V, int EOS>
static void launch_cell_v(const Dev &d, bool rad, bool stats, hipStream_t s) {
  const dim3 g(blocks_for(d.n_cells)), b(kBlock);
  if (rad) {
    if (stats)
      hipLaunchKernelGGL((k_cell<V, true, true, EOS>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((k_cell<V, true, false, EOS>), g, b, 0, s, d);
  } else {
    if (stats)
      hipLaunchKernelGGL((k_cell<V, false, true, EOS>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((k_cell<V, false, false, EOS>), g, b, 0, s, d);
  }
}

// The surrogate start (kMixedMlp) exists for the ideal gas only; rhllc_create
// rejects it with the Taub-Mathews EOS.
hipError_t launch_cell(const Dev &d, int variant, bool rad, bool stats, hipStream_t s) {
  if (d.n_cells == 0) return hipSuccess;
  const bool taub = d.eos == kTaub;
  switch (variant) {
    case kFp64Ref:
      taub ? launch_cell_v<kFp64Ref, kTaub>(d, rad, stats, s) : launch_cell_v<kFp64Ref, kIdeal>(d, rad, stats, s);
      break;
    case kFp64Prev:
      taub ? launch_cell_v<kFp64Prev, kTaub>(d, rad, stats, s) : launch_cell_v<kFp64Prev, kIdeal>(d, rad, stats, s);
      break;
    case kMixedPrev:
      taub ? launch_cell_v<kMixedPrev, kTaub>(d, rad, stats, s) : launch_cell_v<kMixedPrev, kIdeal>(d, rad, stats, s);
      break;
    case kMixedMlp:
      if (taub) return hipErrorInvalidValue;
      launch_cell_v<kMixedMlp, kIdeal>(d, rad, stats, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Grid of the list kernel: enough blocks to fill the chip; each loops over
// the list, whose length is read on the device.
hipError_t launch_cell_mlp(const Dev &d, bool rad, bool stats, hipStream_t s) {
  if (d.n_cells == 0) return hipSuccess;
  const unsigned int nb = blocks_for(d.n_cells) < 2048u ? blocks_for(d.n_cells) : 2048u;
  const dim3 g(nb), b(kBlock);
  if (rad) {
    if (stats)
      hipLaunchKernelGGL((k_cell_mlp<true, true>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((k_cell_mlp<true, false>), g, b, 0, s, d);
  } else {
    if (stats)
      hipLaunchKernelGGL((k_cell_mlp<false, true>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((k_cell_mlp<false, false>), g, b, 0, s, d);
  }
  return hipGetLastError();
}

template <int V, int EOS>
static void launch_tile_v(const Dev &d, bool rad, bool stats, dim3 g, size_t lds, hipStream_t s) {
  const dim3 b(kBlock);
  if (rad) {
    if (stats)
      hipLaunchKernelGGL((k_tile<V, true, true, EOS>), g, b, lds, s, d);
    else
      hipLaunchKernelGGL((k_tile<V, true, false, EOS>), g, b, lds, s, d);
  } else {
    if (stats)
      hipLaunchKernelGGL((k_tile<V, false, true, EOS>), g, b, lds, s, d);
    else
      hipLaunchKernelGGL((k_tile<V, false, false, EOS>), g, b, lds, s, d);
  }
}

// One workgroup per tile; LDS for the largest tile's fluxes (40 bytes a face).
hipError_t launch_tile(const Dev &d, bool mixed, int variant, bool rad, bool stats, int n_tiles, int lds_slots,
                       hipStream_t s) {
  if (d.n_cells == 0) return hipSuccess;
  if (n_tiles < 1 || lds_slots < 1 || lds_slots > kTileMaxSlots) return hipErrorInvalidValue;
  if (mixed != (variant >= kMixedPrev)) return hipErrorInvalidValue;
  const dim3 g((unsigned int)n_tiles);
  const size_t lds = (size_t)lds_slots * 40;
  const bool taub = d.eos == kTaub;
  switch (variant) {
    case kFp64Ref:
      taub ? launch_tile_v<kFp64Ref, kTaub>(d, rad, stats, g, lds, s) : launch_tile_v<kFp64Ref, kIdeal>(d, rad, stats, g, lds, s);
      break;
    case kFp64Prev:
      taub ? launch_tile_v<kFp64Prev, kTaub>(d, rad, stats, g, lds, s)
           : launch_tile_v<kFp64Prev, kIdeal>(d, rad, stats, g, lds, s);
      break;
    case kMixedPrev:
      taub ? launch_tile_v<kMixedPrev, kTaub>(d, rad, stats, g, lds, s)
           : launch_tile_v<kMixedPrev, kIdeal>(d, rad, stats, g, lds, s);
      break;
    case kMixedMlp:
      if (taub) return hipErrorInvalidValue;
      launch_tile_v<kMixedMlp, kIdeal>(d, rad, stats, g, lds, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_face_fp64_out(const Dev &d, bool mixed, double4 *F4, double *FE, hipStream_t s) {
  if (d.n_faces == 0) return hipSuccess;
  const dim3 g(blocks_for(d.n_faces)), b(kBlock);
  if (d.eos == kTaub) {
    if (mixed)
      hipLaunchKernelGGL((k_face_fp64_out<true, kTaub>), g, b, 0, s, d, F4, FE);
    else
      hipLaunchKernelGGL((k_face_fp64_out<false, kTaub>), g, b, 0, s, d, F4, FE);
  } else {
    if (mixed)
      hipLaunchKernelGGL((k_face_fp64_out<true, kIdeal>), g, b, 0, s, d, F4, FE);
    else
      hipLaunchKernelGGL((k_face_fp64_out<false, kIdeal>), g, b, 0, s, d, F4, FE);
  }
  return hipGetLastError();
}

hipError_t launch_dt(const Dev &d, hipStream_t s) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(1024), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_seed_mixed(const Dev &d, hipStream_t s) {
  if (d.n_cells == 0) return hipSuccess;
  if (d.eos == kTaub)
    hipLaunchKernelGGL(k_seed_mixed<kTaub>, dim3(blocks_for(d.n_cells)), dim3(kBlock), 0, s, d);
  else
    hipLaunchKernelGGL(k_seed_mixed<kIdeal>, dim3(blocks_for(d.n_cells)), dim3(kBlock), 0, s, d);
  return hipGetLastError();
}
hipError_t launch_ghost_cell(const Dev &d, bool mixed, hipStream_t s) {
  if (d.eos == kTaub)
    hipLaunchKernelGGL(k_ghost_cell<kTaub>, dim3(1), dim3(64), 0, s, d, mixed);
  else
    hipLaunchKernelGGL(k_ghost_cell<kIdeal>, dim3(1), dim3(64), 0, s, d, mixed);
  return hipGetLastError();
}
hipError_t launch_recover_W(const Dev &d, hipStream_t s) {
  if (d.n_cells == 0) return hipSuccess;
  if (d.eos == kTaub)
    hipLaunchKernelGGL(k_recover_W<kTaub>, dim3(blocks_for(d.n_cells)), dim3(kBlock), 0, s, d);
  else
    hipLaunchKernelGGL(k_recover_W<kIdeal>, dim3(blocks_for(d.n_cells)), dim3(kBlock), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_add_rad(const Dev &d, bool mixed, double tau, hipStream_t s) {
  if (d.n_cells == 0) return hipSuccess;
  const dim3 g(blocks_for(d.n_cells)), b(kBlock);
  const bool taub = d.eos == kTaub;
  if (mixed && taub)
    hipLaunchKernelGGL((k_add_rad<true, kTaub>), g, b, 0, s, d, tau);
  else if (mixed)
    hipLaunchKernelGGL((k_add_rad<true, kIdeal>), g, b, 0, s, d, tau);
  else if (taub)
    hipLaunchKernelGGL((k_add_rad<false, kTaub>), g, b, 0, s, d, tau);
  else
    hipLaunchKernelGGL((k_add_rad<false, kIdeal>), g, b, 0, s, d, tau);
  return hipGetLastError();
}

hipError_t launch_clear_scalars(const Dev &d, hipStream_t s) {
  hipLaunchKernelGGL(k_clear, dim3((kSpeedSlots + 255) / 256), dim3(256), 0, s, d);
  return hipGetLastError();
}

// ----------------------------------------------------------- pointwise evals
__device__ __forceinline__ void ld5(const double *p, State<double> &s) {
  s.d = p[0], s.v[0] = p[1], s.v[1] = p[2], s.v[2] = p[3], s.p = p[4];
}
__device__ __forceinline__ void st5(double *p, const State<double> &s) {
  p[0] = s.d, p[1] = s.v[0], p[2] = s.v[1], p[3] = s.v[2], p[4] = s.p;
}

template <bool FP32, int EOS>
__global__ void k_eval_flux(double gamma, int64_t n, const double *UL, const double *UR, const double *WL,
                            const double *WR, const double *nS4, double *F, double *speed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double Gg = gamma / (gamma - 1.0);
  State<double> ul, ur, wl, wr, f;
  ld5(UL + 5 * i, ul), ld5(UR + 5 * i, ur), ld5(WL + 5 * i, wl), ld5(WR + 5 * i, wr);
  if (!FP32) {
    speed[i] = hllc_flux<double, Ieee, EOS>(gamma, Gg, ul, ur, wl, wr, nS4[4 * i], nS4[4 * i + 1], nS4[4 * i + 2],
                                            nS4[4 * i + 3], f);
    st5(F + 5 * i, f);
  } else {
    // The mixed face path: FP32 primitives, FP64 geometry and fast arithmetic,
    // flux in the two-kernel step's packed record.
    // The cell records the mixed cell kernel would store for these states.
    State<double> ff;
    float4 la, ra;
    double2 ls, rs;
    mixed_record<EOS>(Gg, wl, la, ls);
    mixed_record<EOS>(Gg, wr, ra, rs);
    const double *g = nS4 + 4 * i;
    speed[i] = mixed_face_flux<EOS>(gamma, mixed_in(la, ls), mixed_in(ra, rs), g[0] * g[3], g[1] * g[3],
                                    g[2] * g[3], ff);
    const rhllc_ext::Packed pk = rhllc_ext::pack(ff.d, ff.v[0], ff.v[1], ff.v[2], ff.p);
    rhllc_ext::unpack(pk.a, pk.b, F + 5 * i);
  }
}

template <int EOS>
__global__ void k_eval_c2p(bool mixed, double gamma, int64_t n, const double *U, const double *guess, double *W,
                           int32_t *codes, int32_t *iters) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  State<double> u, w;
  ld5(U + 5 * i, u);
  ld5(W + 5 * i, w);
  int it = 0;
  const double Gg = gamma / (gamma - 1.0), g0 = guess ? guess[i] : -1.0;
  codes[i] = mixed ? c2p<Fast, EOS>(Gg, u, w, g0, it) : c2p<Ieee, EOS>(Gg, u, w, g0, it);
  iters[i] = it;
  st5(W + 5 * i, w);
}

template <int EOS>
__global__ void k_eval_p2c(double gamma, int64_t n, double *W, double *U) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  State<double> w, u;
  ld5(W + 5 * i, w);
  p2c<double, Ieee, EOS>(gamma / (gamma - 1.0), w, u);
  st5(W + 5 * i, w);
  st5(U + 5 * i, u);
}

template <int EOS>
__global__ void k_eval_fix(double gamma, int64_t n, double *U, double *W, int32_t *codes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  State<double> u, w;
  ld5(U + 5 * i, u);
  ld5(W + 5 * i, w);
  codes[i] = pressure_fix<EOS>(gamma / (gamma - 1.0), u, w);
  st5(U + 5 * i, u);
  st5(W + 5 * i, w);
}

template <typename A>
__global__ void k_eval_rad(int64_t n, const double *W, const double *Ur, const double *Fr3, const double *ka,
                           const double *ks, int emission, const float *Tr9, double *G) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  State<double> w;
  ld5(W + 5 * i, w);
  rad_source<A>(w, Ur[i], Fr3 + 3 * i, ka[i], ks[i], G + 4 * i, emission != 0, Tr9 ? Tr9 + 9 * i : nullptr);
}

__global__ void k_eval_rad_opt(OpacityTable t, int64_t n, const double *W, const double *Ur, const double *Fr3,
                               const float *Tr9, double *G, double *cell) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  State<double> w;
  ld5(W + 5 * i, w);
  const CellRad c = cell_rad(t, w);
  rad_source_opt(w, c, Ur[i], Fr3 + 3 * i, G + 4 * i, Tr9 ? Tr9 + 9 * i : nullptr);
  double *o = cell + 5 * i;
  o[0] = c.T, o[1] = c.vel, o[2] = c.lorenz, o[3] = c.alpha, o[4] = c.betta;
}

// Plain copy (calibration of byte counters and the streaming-rate reference).
// One float4 per thread over a grid that covers the array: on gfx950 this
// streams faster (6.15 TB/s at 2 GiB) than grid-stride loops of a fixed grid,
// with one or four loads in flight per thread (4.5-4.8 TB/s).
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

hipError_t launch_copy(const void *src, void *dst, int64_t n16, hipStream_t s) {
  if (n16 <= 0) return hipSuccess;
  const int64_t blocks = (n16 + 255) / 256;
  if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy, dim3((unsigned int)blocks), dim3(256), 0, s, (const float4 *)src, (float4 *)dst, n16);
  return hipGetLastError();
}

static inline dim3 eval_grid(int64_t n) { return dim3((unsigned int)((n + 255) / 256)); }

template <int EOS>
static void eval_flux_e(bool fp32, double gamma, int64_t n, const double *UL, const double *UR, const double *WL,
                        const double *WR, const double *nS4, double *F, double *speed, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL((k_eval_flux<true, EOS>), eval_grid(n), dim3(256), 0, s, gamma, n, UL, UR, WL, WR, nS4, F,
                       speed);
  else
    hipLaunchKernelGGL((k_eval_flux<false, EOS>), eval_grid(n), dim3(256), 0, s, gamma, n, UL, UR, WL, WR, nS4, F,
                       speed);
}
hipError_t launch_eval_flux(bool fp32, int eos, double gamma, int64_t n, const double *UL, const double *UR,
                            const double *WL, const double *WR, const double *nS4, double *F, double *speed,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (eos == kTaub)
    eval_flux_e<kTaub>(fp32, gamma, n, UL, UR, WL, WR, nS4, F, speed, s);
  else
    eval_flux_e<kIdeal>(fp32, gamma, n, UL, UR, WL, WR, nS4, F, speed, s);
  return hipGetLastError();
}
hipError_t launch_eval_c2p(bool mixed, int eos, double gamma, int64_t n, const double *U, const double *guess,
                           double *W, int32_t *codes, int32_t *iters, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (eos == kTaub)
    hipLaunchKernelGGL(k_eval_c2p<kTaub>, eval_grid(n), dim3(256), 0, s, mixed, gamma, n, U, guess, W, codes, iters);
  else
    hipLaunchKernelGGL(k_eval_c2p<kIdeal>, eval_grid(n), dim3(256), 0, s, mixed, gamma, n, U, guess, W, codes, iters);
  return hipGetLastError();
}
hipError_t launch_eval_p2c(int eos, double gamma, int64_t n, double *W, double *U, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (eos == kTaub)
    hipLaunchKernelGGL(k_eval_p2c<kTaub>, eval_grid(n), dim3(256), 0, s, gamma, n, W, U);
  else
    hipLaunchKernelGGL(k_eval_p2c<kIdeal>, eval_grid(n), dim3(256), 0, s, gamma, n, W, U);
  return hipGetLastError();
}
hipError_t launch_eval_fix(int eos, double gamma, int64_t n, double *U, double *W, int32_t *codes, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (eos == kTaub)
    hipLaunchKernelGGL(k_eval_fix<kTaub>, eval_grid(n), dim3(256), 0, s, gamma, n, U, W, codes);
  else
    hipLaunchKernelGGL(k_eval_fix<kIdeal>, eval_grid(n), dim3(256), 0, s, gamma, n, U, W, codes);
  return hipGetLastError();
}
hipError_t launch_eval_rad(bool mixed, int64_t n, const double *W, const double *Ur, const double *Fr3,
                           const double *ka, const double *ks, int emission, const float *Tr9, double *G,
                           hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (mixed)
    hipLaunchKernelGGL(k_eval_rad<Fast>, eval_grid(n), dim3(256), 0, s, n, W, Ur, Fr3, ka, ks, emission, Tr9, G);
  else
    hipLaunchKernelGGL(k_eval_rad<Ieee>, eval_grid(n), dim3(256), 0, s, n, W, Ur, Fr3, ka, ks, emission, Tr9, G);
  return hipGetLastError();
}

hipError_t launch_eval_rad_opt(int nx, int ny, const double *hdr, const double *table, int64_t n, const double *W,
                               const double *Ur, const double *Fr3, const float *Tr9, double *G, double *cell,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  const OpacityTable t{table, nx, ny, hdr[0], hdr[1], hdr[2], hdr[3], hdr[4], hdr[5]};
  hipLaunchKernelGGL(k_eval_rad_opt, eval_grid(n), dim3(256), 0, s, t, n, W, Ur, Fr3, Tr9, G, cell);
  return hipGetLastError();
}

}  // namespace rhllc_dev
