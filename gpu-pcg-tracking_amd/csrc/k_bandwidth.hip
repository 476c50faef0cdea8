// k_bandwidth.hip -- the bandwidth probes that the per-phase roofline of the
// batched SQP iteration is priced against, measured on the device itself
// (gfx950): an HBM stream copy and an LDS read loop.
//
//   k_bw_copy: dst = src with 128-bit loads and stores, grid-stride; the
//              bytes moved are 2 x 4 n (one read, one write per float).
//   k_bw_lds:  every 256-lane workgroup fills a 16 KB LDS block, then each
//              lane reads reps x 4 128-bit words of it (consecutive lanes,
//              consecutive words: no bank conflicts) and folds their bits
//              with 3-input integer adds, so VALU issue stays well below
//              the LDS rate; the bytes read are n_wg x 256 x reps x 64.
// The folded sums are written out (one uint per lane) so that no read can
// be dropped, and a test checks them against their closed form.
#include "common.h"

namespace mpc {

namespace {

constexpr uint32_t BW_NT = 256;
constexpr uint32_t LDS_WORDS = 1024;  // 128-bit words per workgroup (16 KB)

__global__ void __launch_bounds__(BW_NT) k_bw_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                   size_t n4) {
  const size_t stride = (size_t)gridDim.x * BW_NT;
  for (size_t i = (size_t)blockIdx.x * BW_NT + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(BW_NT) k_bw_lds(uint32_t *__restrict__ out, uint32_t reps) {
  __shared__ uint4 buf[LDS_WORDS];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < LDS_WORDS; i += BW_NT) buf[i] = make_uint4(4u * i, 4u * i + 1u, 4u * i + 2u, 4u * i + 3u);
  __syncthreads();
  uint32_t a = 0, b = 0;
  for (uint32_t r = 0; r < reps; ++r) {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint4 v = buf[(t + BW_NT * k + r) & (LDS_WORDS - 1u)];
      a = a + v.x + v.y;
      b = b + v.z + v.w;
    }
  }
  out[(size_t)blockIdx.x * BW_NT + t] = a ^ b;
}

}  // namespace

hipError_t launch_bw_copy(const float *src, float *dst, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (n % 4 != 0 || (reinterpret_cast<uintptr_t>(src) & 15u) || (reinterpret_cast<uintptr_t>(dst) & 15u))
    return hipErrorInvalidValue;
  const size_t n4 = n / 4;
  const size_t want = (n4 + BW_NT - 1) / BW_NT;
  const uint32_t grid = (uint32_t)(want < 65536 ? want : 65536);
  hipLaunchKernelGGL(k_bw_copy, dim3(grid), dim3(BW_NT), 0, st, reinterpret_cast<const float4 *>(src),
                     reinterpret_cast<float4 *>(dst), n4);
  return hipGetLastError();
}

hipError_t launch_bw_lds(uint32_t *out, uint32_t n_wg, uint32_t reps, hipStream_t st) {
  if (n_wg == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bw_lds, dim3(n_wg), dim3(BW_NT), 0, st, out, reps);
  return hipGetLastError();
}

}  // namespace mpc
