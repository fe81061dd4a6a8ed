// Launchers of the surrogate-MLP kernels (mlp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhllc_mlp {

hipError_t launch_gen(double gamma, uint64_t seed, const uint32_t *step, int n, const float *norm, float *X, float *Y,
                      float *OK, hipStream_t s);
hipError_t launch_grad(const float *P, const short *W2b, const short *W2Tb, const float *X, const float *Y,
                       const float *OK, float inv_n, int n, float *grad, float *loss, hipStream_t s);
hipError_t launch_adam(float *P, float *M, float *V, short *W2b, short *W2Tb, const float *grad, int nwaves,
                       const uint32_t *step, int total_steps, float lr0, const float *loss, float *loss_hist,
                       hipStream_t s);
hipError_t launch_tick(uint32_t *step, hipStream_t s);
// out[p] = sum_w part[w * np + p] in w order (FP32).
hipError_t launch_sum_partials(const float *part, int nw, int np, float *out, hipStream_t s);
hipError_t launch_pack(const float *P, short *W2b, short *W2Tb, hipStream_t s);
hipError_t launch_eval(const float *P, const short *W2b, int64_t n, const double *U, double *p_out, hipStream_t s);
hipError_t launch_selftest(float *out, hipStream_t s);

}  // namespace rhllc_mlp
