// k_tune.hip -- the synthetic Schur system on which the launch-plan
// autotuner times the batched PCG (gfx950).  Every knot gets the block row
// S(b,b-1) = I (b > 0), S(b,b) = -2.05 I, S(b,b+1) = I (b < N-1), the
// block-Jacobi Pinv(b,b) = -I / 2.05, gamma_b = 1 and lambda_b = 0: a
// negative definite system (condition number about 81) on which a fixed
// number of PCG iterations stays finite, so every plan runs the same
// arithmetic for the same time per iteration.  Layouts as the Schur kernels
// write them: per block row [left | diag | right], 14 x 14 column-major.
#include "common.h"

namespace mpc {

namespace {

__global__ void __launch_bounds__(256) k_tune_system(Batch b, float *__restrict__ S, float *__restrict__ Pinv,
                                                     float *__restrict__ gamma, float *__restrict__ lambda) {
  const uint32_t gk = blockIdx.x;
  const KnotRef kr = locate(b, gk);
  float *Sb = S + (size_t)S_STRIDE * gk;
  float *Pb = Pinv + (size_t)S_STRIDE * gk;
  for (uint32_t j = threadIdx.x; j < (uint32_t)S_STRIDE; j += blockDim.x) {
    const uint32_t blk = j / NXX, e = j % NXX, r = e % NX, c = e / NX;
    const bool diag = r == c;
    float s = 0.f, p = 0.f;
    if (blk == 0) s = diag && kr.k > 0 ? 1.f : 0.f;
    else if (blk == 1) s = diag ? -2.05f : 0.f, p = diag ? -1.f / 2.05f : 0.f;
    else s = diag && kr.k + 1 < kr.N ? 1.f : 0.f;
    Sb[j] = s;
    Pb[j] = p;
  }
  if (threadIdx.x < NX) {
    gamma[(size_t)NX * gk + threadIdx.x] = 1.f;
    lambda[(size_t)NX * gk + threadIdx.x] = 0.f;
  }
}

}  // namespace

hipError_t launch_tune_system(const Batch &b, float *S, float *Pinv, float *gamma, float *lambda, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tune_system, dim3(b.total_knots), dim3(256), 0, st, b, S, Pinv, gamma, lambda);
  return hipGetLastError();
}

}  // namespace mpc
