// k_pcg_res.hip -- PCG with the Schur system resident on chip (gfx950):
// one workgroup (512 or 256 lanes) per bin of instances, the body in
// pcg_res_body.h.
#include "common.h"
#include "pcg_res_body.h"

namespace mpc {

namespace {

using namespace resident;

template <bool STAIR, int NT>
__global__ void __launch_bounds__(NT) k_pcg_res(Batch b, const uint4 *__restrict__ items,
                                                const uint32_t *__restrict__ bin_off, uint32_t bin0,
                                                const float *__restrict__ S, const float *__restrict__ Pinv,
                                                const float *__restrict__ gamma, float *__restrict__ lambda,
                                                uint32_t max_iter, float tol, const uint32_t *__restrict__ status,
                                                uint32_t *__restrict__ iters_out, uint32_t *__restrict__ flag_out,
                                                float *__restrict__ eta_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  body<STAIR, NT>(threadIdx.x, bin0 + blockIdx.x, smem_raw, b, items, bin_off, S, Pinv, gamma, lambda, max_iter,
                  tol, status, iters_out, flag_out, eta_out);
}

template <int NT>
hipError_t launch_shape(const Batch &b, const uint4 *items, const uint32_t *bin_off, uint32_t bin0, uint32_t n_bins,
                        const float *S, const float *Pinv, const float *gamma, float *lambda, uint32_t max_iter,
                        float tol, const uint32_t *status, uint32_t *iters, uint32_t *flag, float *eta, bool stair,
                        hipStream_t st) {
  static bool attr_set = false;
  const size_t smem = sizeof(ResShared<NT>);
  if (!attr_set) {
    // allow more than the default dynamic LDS; the launch reports any error
    (void)hipFuncSetAttribute((const void *)k_pcg_res<false, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem);
    (void)hipFuncSetAttribute((const void *)k_pcg_res<true, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem);
    (void)hipGetLastError();
    attr_set = true;
  }
  if (stair)
    hipLaunchKernelGGL((k_pcg_res<true, NT>), dim3(n_bins), dim3(NT), smem, st, b, items, bin_off, bin0, S, Pinv,
                       gamma, lambda, max_iter, tol, status, iters, flag, eta);
  else
    hipLaunchKernelGGL((k_pcg_res<false, NT>), dim3(n_bins), dim3(NT), smem, st, b, items, bin_off, bin0, S, Pinv,
                       gamma, lambda, max_iter, tol, status, iters, flag, eta);
  return hipGetLastError();
}

}  // namespace

uint32_t pcg_res_groups(uint32_t N) {
  const uint32_t groups = (N + RK - 1) / RK;
  return (groups + 3u) & ~3u;  // whole waves
}
uint32_t pcg_res_bin_groups(uint32_t nt) { return nt / 16u; }
uint32_t pcg_res_max_knots(uint32_t nt) { return nt / 16u * (uint32_t)RK; }
uint32_t pcg_res_max_items() { return MAXI; }

// bins bin0 .. bin0 + n_bins - 1 of the bin table, one workgroup of nt
// lanes (512 or 256) each
hipError_t launch_pcg_res(const Batch &b, const uint4 *items, const uint32_t *bin_off, uint32_t bin0, uint32_t n_bins,
                          uint32_t nt, const float *S, const float *Pinv, const float *gamma, float *lambda,
                          uint32_t max_iter, float tol, const uint32_t *status, uint32_t *iters,
                          uint32_t *flag, float *eta, bool stair, hipStream_t st) {
  if (n_bins == 0) return hipSuccess;
  if (nt == 512)
    return launch_shape<512>(b, items, bin_off, bin0, n_bins, S, Pinv, gamma, lambda, max_iter, tol, status, iters,
                             flag, eta, stair, st);
  if (nt == 256)
    return launch_shape<256>(b, items, bin_off, bin0, n_bins, S, Pinv, gamma, lambda, max_iter, tol, status, iters,
                             flag, eta, stair, st);
  return hipErrorInvalidValue;
}

}  // namespace mpc
