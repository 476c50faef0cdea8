// direct_static.hip -- mpc_direct_solve_static: the direct Schur solve of
// mpc_direct_solve with its factor workspace in a file-scope device array
// instead of a caller- or library-allocated buffer, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops).  No handle: the entry point allocates nothing, neither device nor
// host memory.
//
// The array `scratch` holds the 588 N floats the block LDL^T of -S needs
// (E_b, D_b^-1 and a vector per knot, the S layout) for N up to
// MPC_DIRECT_STATIC_MAX_KNOTS.  Every GPU the library runs on holds its own
// copy at its own address until the process ends.  Host code never passes
// the array's name to a kernel (in host code it is a stand-in symbol, not a
// device address): the first call on a device, with that device current,
// looks the address up with hipGetSymbolAddress and keeps it per device.
//
// Calls on one device share the array: two calls that overlap in time on
// one GPU (different streams, or graphs replayed concurrently) race on it.
// The kernels write every workspace word they read within the call (the
// factor sweep writes E_b and D_b^-1 of every knot before the substitution
// reads them, the residual kernel writes every knot's vector slot before
// the refinement reads it), so nothing left by an earlier call is read and
// nothing needs clearing.
#include <mutex>

#include "common.h"
#include "mpcpcg.h"

namespace {

constexpr uint32_t kMaxKnots = MPC_DIRECT_STATIC_MAX_KNOTS;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[(size_t)mpc::S_STRIDE * kMaxKnots];

std::mutex g_mu;
float *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (nullptr on failure)
float *scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[dev] = static_cast<float *>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" int mpc_direct_solve_static(uint32_t N, const float *d_S, const float *d_gamma, float *d_lambda,
                                       uint32_t refine_steps, mpc_stream_t stream) {
  if (N < 2 || N > kMaxKnots || !d_S || !d_gamma || !d_lambda || refine_steps > 16) return MPC_ERR_ARG;
  float *work = scratch_on_current_device();
  if (!work) return MPC_ERR_HIP;
  const mpc::Batch b{nullptr, nullptr, 1u, N};
  if (mpc::launch_direct_solve(b, nullptr, 1, d_S, d_gamma, nullptr, work, d_lambda, refine_steps, true, nullptr,
                               nullptr, (hipStream_t)stream) != hipSuccess)
    return MPC_ERR_HIP;
  return MPC_OK;
}
