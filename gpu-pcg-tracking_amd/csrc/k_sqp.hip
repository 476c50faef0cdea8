// k_sqp.hip -- per-instance SQP bookkeeping kernels: merit reductions,
// line-search acceptance, rho rule and the trajectory update, all on the
// device so a batched SQP iteration is a fixed sequence of launches.
//
// Reference: the host logic of sqpSolve (toplevel.cuh:384-444) and the
// cublasSaxpy update (toplevel.cuh:424-430):
//   choose the smallest p with merit_p strictly below the current merit;
//   failure:  drho = max(drho * f, f); rho = max(rho * drho, rho_min);
//             rho > rho_max -> abort the solve, rho = rho_reset
//   success:  alpha = -2^-p; drho = min(drho / f, 1 / f);
//             rho = max(rho * drho, rho_min); xu += alpha dz; merit = min
// Sums over knots use a fixed-order tree, so they are reproducible
// (the reference's initial merit uses atomicAdd, merit.cuh:142).
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// sum over the instance's knots of knot_merits[slot * T + K0 + k] for
// slots 0..NP-1; result visible to every thread in out[NP]
template <int NP>
__device__ __forceinline__ void reduce_slots(const float *__restrict__ km, uint32_t T, uint32_t K0,
                                             uint32_t N, float *red, float *out) {
  float part[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) part[p] = 0.f;
  for (uint32_t k = threadIdx.x; k < N; k += NT)
#pragma unroll
    for (int p = 0; p < NP; ++p) part[p] += km[(size_t)p * T + K0 + k];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const float v = wave_sum(part[p]);
    if ((threadIdx.x & 63) == 0) red[p * (NT / 64) + w] = v;
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += red[threadIdx.x * (NT / 64) + i];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(NT) k_merit_reduce(Batch b, const float *__restrict__ km, int np,
                                                     float *__restrict__ out) {
  __shared__ float red[NT / 64];
  const uint32_t inst = blockIdx.x, slot = blockIdx.y;
  const uint32_t K0 = inst_first(b, inst), N = inst_len(b, inst);
  float part = 0.f;
  for (uint32_t k = threadIdx.x; k < N; k += NT) part += km[(size_t)slot * b.total_knots + K0 + k];
  part = wave_sum(part);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    out[(size_t)inst * np + slot] = s;
  }
}

__global__ void __launch_bounds__(NT) k_batch_init(Batch b, const float *__restrict__ xu,
                                                   const float *__restrict__ km, InstState s) {
  __shared__ float red[NT / 64];
  __shared__ float m0[1];
  const uint32_t inst = blockIdx.x;
  const uint32_t K0 = inst_first(b, inst), N = inst_len(b, inst);
  reduce_slots<1>(km, b.total_knots, K0, N, red, m0);
  const float *xi = xu + (size_t)NXU * K0 - (size_t)NU * inst;
  if (threadIdx.x < NX) s.xs[(size_t)NX * inst + threadIdx.x] = xi[threadIdx.x];
  if (threadIdx.x == 0) {
    s.merit0[inst] = m0[0];
    s.drho[inst] = 1.f;
    s.status[inst] = ST_ACTIVE;
    s.sqp_done[inst] = 0;
    if (inst == 0 && s.t0) *s.t0 = (unsigned long long)wall_clock64();
  }
}

__global__ void __launch_bounds__(NT) k_ls_update(Batch b, Params p, const float *__restrict__ km,
                                                  const float *__restrict__ dz, float *__restrict__ xu,
                                                  InstState s, uint32_t *__restrict__ pcg_log,
                                                  const uint32_t *__restrict__ pcg_iters_now,
                                                  const uint32_t *__restrict__ pcg_flag_now,
                                                  uint32_t it, uint32_t log_stride) {
  __shared__ float red[8 * (NT / 64)];
  __shared__ float merits[8];
  __shared__ float alpha_s[1];
  __shared__ uint32_t status_s[1];
  const uint32_t inst = blockIdx.x;
  // the instance's values, read up front: independent addresses, one round
  // trip (this workgroup is their only writer below)
  const uint32_t st0 = s.status[inst];
  const uint32_t K0 = inst_first(b, inst), N = inst_len(b, inst);
  uint32_t pit = 0u, pf = 0u, done = 0u;
  float m0 = 0.f, rho0 = 0.f, drho0 = 0.f;
  if (threadIdx.x == 0) {
    pit = pcg_iters_now ? pcg_iters_now[inst] : 0u;
    pf = pcg_flag_now ? pcg_flag_now[inst] : 0u;
    m0 = s.merit0[inst];
    rho0 = s.rho[inst];
    drho0 = s.drho[inst];
    done = s.sqp_done[inst];
  }
  if (!(st0 & ST_ACTIVE)) return;  // uniform over the workgroup
  reduce_slots<8>(km, b.total_knots, K0, N, red, merits);
  if (threadIdx.x == 0) {
    uint32_t st = st0;
    if (pcg_log) pcg_log[(size_t)inst * log_stride + it] = pit;
    if (pf == PCG_FLAG_FAULT) st |= ST_PCG_FAULT;
    else if (pf) st |= ST_PCG_MAXITER;
    // the time budget, checked where the reference checks it before its
    // line-search decision (toplevel.cuh:381): past it, the iteration ends
    // without its step and is not counted
    if (p.budget_ticks && s.t0 && (unsigned long long)wall_clock64() - *s.t0 > p.budget_ticks)
      st = (st | ST_TIME_BUDGET) & ~ST_ACTIVE;
    s.status[inst] = st;
    status_s[0] = st;
  }
  __syncthreads();
  if (!(status_s[0] & ST_ACTIVE)) return;  // uniform: written by thread 0 before the barrier
  if (threadIdx.x == 0) {
    uint32_t st = status_s[0];
    float minm = m0;
    int step = -1;
    for (int i = 0; i < 8; ++i)
      if (merits[i] < minm) {
        minm = merits[i];
        step = i;
      }
    float rho = rho0, drho = drho0;
    float alpha = 0.f;
    if (step < 0) {
      drho = fmaxf(drho * p.rho_factor, p.rho_factor);
      rho = fmaxf(rho * drho, p.rho_min);
      if (rho > p.rho_max) {
        st |= ST_RHO_ABORT;
        st &= ~ST_ACTIVE;
        rho = p.rho_reset;
      }
      if (!isfinite(merits[0])) st |= ST_NONFINITE;
    } else {
      alpha = -1.f / (float)(1u << step);
      drho = fminf(drho / p.rho_factor, 1.f / p.rho_factor);
      rho = fmaxf(rho * drho, p.rho_min);
      s.merit0[inst] = minm;
    }
    if (st & ST_PCG_FAULT) {  // the step came from an abandoned solve: not taken
      st &= ~ST_ACTIVE;
      alpha = 0.f;
    }
    s.rho[inst] = rho;
    s.drho[inst] = drho;
    s.sqp_done[inst] = done + 1u;
    s.status[inst] = st;
    alpha_s[0] = alpha;
  }
  __syncthreads();
  const float alpha = alpha_s[0];
  if (alpha != 0.f) {
    const size_t off = (size_t)NXU * K0 - (size_t)NU * inst;
    const uint32_t len = NXU * N - NU;
    for (uint32_t i = threadIdx.x; i < len; i += NT) xu[off + i] += alpha * dz[off + i];
  }
}

__global__ void k_axpy(uint32_t n, float alpha, const float *__restrict__ x, float *__restrict__ y) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += alpha * x[i];
}

}  // namespace

hipError_t launch_merit_reduce(const Batch &b, const float *knot_merits, int np, float *out,
                               hipStream_t st) {
  if (b.n_inst == 0 || np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_merit_reduce, dim3(b.n_inst, np), dim3(NT), 0, st, b, knot_merits, np, out);
  return hipGetLastError();
}

hipError_t launch_batch_init(const Batch &b, const float *xu, const float *knot_merits, InstState s,
                             hipStream_t st) {
  if (b.n_inst == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_init, dim3(b.n_inst), dim3(NT), 0, st, b, xu, knot_merits, s);
  return hipGetLastError();
}

hipError_t launch_ls_update(const Batch &b, const Params &p, const float *knot_merits,
                            const float *dz, float *xu, InstState s, uint32_t *pcg_iter_log,
                            const uint32_t *pcg_iters_now, const uint32_t *pcg_flag_now,
                            uint32_t it, uint32_t log_stride, hipStream_t st) {
  if (b.n_inst == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ls_update, dim3(b.n_inst), dim3(NT), 0, st, b, p, knot_merits, dz, xu, s,
                     pcg_iter_log, pcg_iters_now, pcg_flag_now, it, log_stride);
  return hipGetLastError();
}

hipError_t launch_axpy(uint32_t n, float alpha, const float *x, float *y, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, st, n, alpha, x, y);
  return hipGetLastError();
}

}  // namespace mpc
