// k_pcg.hip -- preconditioned conjugate gradient on the block-tridiagonal
// Schur system S lambda = gamma, one workgroup per instance (gfx950).
//
// Mapping: a workgroup of NG 16-lane groups; lane i < 14 of group g owns
// row i of block rows b = g, g + NG, g + 2 NG, ... (at most R of them), and
// keeps lambda, r, p, S p of those rows in registers.  The search
// direction p is exchanged through LDS (S p needs p_{b-1}, p_b, p_{b+1});
// r through LDS inside the group only (the preconditioner is block
// diagonal).  Dot products: wave64 butterfly (DPP) + a fixed-order sum of
// the per-wave partials, so results are bitwise reproducible.
//
// Algorithm (GPU-PCG pcg<T,14,N>, call site toplevel.cuh:143-157; the
// submodule source is absent from the reference snapshot), with z = M r for
// the block-Jacobi M = D^-1 or (stair) M r = D^-1 (2 r - S D^-1 r):
//   r = gamma - S lambda, z = Pinv r, p = z, eta = r.z
//   exit if |eta| < tol (checked once before the loop as well)
//   loop: y = S p; alpha = eta / p.y; lambda += alpha p; r -= alpha y;
//         z = Pinv r; eta' = r.z; iters++; exit if |eta'| < tol;
//         p = z + (eta'/eta) p; eta = eta'
// S blocks are read in the reference layout ([left | diag | right] per
// block row, column-major 14x14); only the diagonal blocks of Pinv are read.
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NW>
__device__ __forceinline__ float block_sum(float v, float *red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < NW; ++i) s += red[i];
  return s;
}

// row i of (B x) for a column-major 14x14 block in global memory
__device__ __forceinline__ float blk_row_dot(const float *__restrict__ blk, int i, const float *x) {
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < NX; ++j) acc += blk[i + NX * j] * x[j];
  return acc;
}
// row i of (B^T x): column i of the column-major block (S(b,b+1) = S(b+1,b)^T
// is read from the next block row's left block)
__device__ __forceinline__ float blk_col_dot(const float *__restrict__ blk, int i, const float *x) {
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < NX; ++j) acc += blk[j + NX * i] * x[j];
  return acc;
}

template <int NG, int R, bool STAIR>
__global__ void __launch_bounds__(NG * 16) k_pcg(Batch b, const float *__restrict__ S,
                                                 const float *__restrict__ Pinv,
                                                 const float *__restrict__ gamma,
                                                 float *__restrict__ lambda, uint32_t max_iter,
                                                 float tol, const uint32_t *__restrict__ status,
                                                 const uint32_t *__restrict__ inst_list,
                                                 uint32_t *__restrict__ iters_out,
                                                 uint32_t *__restrict__ flag_out,
                                                 float *__restrict__ eta_out) {
  constexpr int NT = NG * 16;
  constexpr int NW = NT / 64;
  extern __shared__ float smem[];
  const uint32_t inst = inst_list ? inst_list[blockIdx.x] : blockIdx.x;
  if (status && !(status[inst] & ST_ACTIVE)) return;  // uniform over the workgroup
  const uint32_t K0 = inst_first(b, inst);
  const uint32_t N = inst_len(b, inst);
  float *pv = smem;                 // [14 N] search direction / lambda staging
  float *rv = smem + NX * N;        // [14 N] residual
  float *red = rv + NX * N;         // [2][NW] reduction slots (double buffered)
  const int grp = threadIdx.x >> 4, i = threadIdx.x & 15;
  const bool lane_ok = i < NX;
  const float *Sg = S + (size_t)S_STRIDE * K0;
  const float *Pg = Pinv + (size_t)S_STRIDE * K0;
  float lam[R], r[R], p[R], y[R];

  // S x for my rows, x staged in pv
  auto smv = [&](float *out) {
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const uint32_t bk = grp + NG * t;
      float acc = 0.f;
      if (lane_ok && bk < N) {
        const float *row = Sg + (size_t)S_STRIDE * bk;
        acc = blk_row_dot(row + NXX, i, pv + NX * bk);
        if (bk > 0) acc += blk_row_dot(row, i, pv + NX * (bk - 1));
        if (bk + 1 < N) acc += blk_col_dot(row + S_STRIDE, i, pv + NX * (bk + 1));
      }
      out[t] = acc;
    }
  };
  // z = D^-1 v for my rows (v staged in rv), returns the r.z partial
  auto jacobi = [&](float *z) {
    float part = 0.f;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const uint32_t bk = grp + NG * t;
      float acc = 0.f;
      if (lane_ok && bk < N) {
        acc = blk_row_dot(Pg + (size_t)S_STRIDE * bk + NXX, i, rv + NX * bk);
        part += r[t] * acc;
      }
      z[t] = acc;
    }
    return part;
  };
  // z = M r (r staged in rv), returns the r.z partial.  Stair: M r =
  // D^-1 (2 r - S D^-1 r); w = D^-1 r is staged in pv (p lives in
  // registers), 2 r - S w in rv.  Ends with rv, pv free to overwrite after
  // the caller's next barrier.
  auto precond = [&](float *z) {
    if constexpr (!STAIR) {
      return jacobi(z);
    } else {
      float wv[R], y2[R];
      jacobi(wv);
#pragma unroll
      for (int t = 0; t < R; ++t) {
        const uint32_t bk = grp + NG * t;
        if (lane_ok && bk < N) pv[NX * bk + i] = wv[t];
      }
      __syncthreads();
      smv(y2);
      __syncthreads();  // every read of rv and pv done
#pragma unroll
      for (int t = 0; t < R; ++t) {
        const uint32_t bk = grp + NG * t;
        if (lane_ok && bk < N) rv[NX * bk + i] = 2.f * r[t] - y2[t];
      }
      __syncthreads();
      return jacobi(z);
    }
  };

#pragma unroll
  for (int t = 0; t < R; ++t) {
    const uint32_t bk = grp + NG * t;
    lam[t] = (lane_ok && bk < N) ? lambda[(size_t)NX * (K0 + bk) + i] : 0.f;
    if (lane_ok && bk < N) pv[NX * bk + i] = lam[t];
  }
  __syncthreads();
  smv(y);
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const uint32_t bk = grp + NG * t;
    r[t] = (lane_ok && bk < N) ? gamma[(size_t)NX * (K0 + bk) + i] - y[t] : 0.f;
    if (lane_ok && bk < N) rv[NX * bk + i] = r[t];
  }
  __syncthreads();
  float z[R];
  float eta = block_sum<NW>(precond(z), red);
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const uint32_t bk = grp + NG * t;
    p[t] = z[t];
    if (lane_ok && bk < N) pv[NX * bk + i] = p[t];
  }
  __syncthreads();
  uint32_t it = 0;
  // tol < 0: exit relative to the first residual, |eta| < -tol |eta_0|
  const float tol_i = tol < 0.f ? fmaxf(-tol * fabsf(eta), FLT_MIN) : tol;
  bool conv = fabsf(eta) < tol_i;
  float eta_last = eta;  // of the last exit test
  const uint32_t cap = pcg_cap(max_iter, N);
  int slot = 1;
  while (!conv && it < cap) {
    smv(y);
    float part = 0.f;
#pragma unroll
    for (int t = 0; t < R; ++t) part += p[t] * y[t];
    const float pAp = block_sum<NW>(part, red + NW * slot);
    slot ^= 1;
    const float alpha = eta / pAp;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const uint32_t bk = grp + NG * t;
      lam[t] += alpha * p[t];
      r[t] -= alpha * y[t];
      if (lane_ok && bk < N) rv[NX * bk + i] = r[t];
    }
    __syncthreads();
    const float eta_new = block_sum<NW>(precond(z), red + NW * slot);
    eta_last = eta_new;
    slot ^= 1;
    ++it;
    if (fabsf(eta_new) < tol_i) {
      conv = true;
      break;
    }
    const float beta = eta_new / eta;
    eta = eta_new;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const uint32_t bk = grp + NG * t;
      p[t] = z[t] + beta * p[t];
      if (lane_ok && bk < N) pv[NX * bk + i] = p[t];
    }
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const uint32_t bk = grp + NG * t;
    if (lane_ok && bk < N) lambda[(size_t)NX * (K0 + bk) + i] = lam[t];
  }
  if (threadIdx.x == 0) {
    if (iters_out) iters_out[inst] = it;
    if (flag_out) flag_out[inst] = conv ? 0u : 1u;
    if (eta_out) eta_out[inst] = fabsf(eta_last);
  }
}

template <int NG, int R>
hipError_t launch_pcg_t(const Batch &b, uint32_t max_n, const float *S, const float *Pinv,
                        const float *gamma, float *lambda, uint32_t max_iter, float tol,
                        const uint32_t *status, const uint32_t *inst_list, uint32_t n_list,
                        uint32_t *iters, uint32_t *flag, float *eta, bool stair, hipStream_t st) {
  const size_t smem = sizeof(float) * (2 * NX * (size_t)max_n + 2 * (NG * 16 / 64));
  const uint32_t grid = inst_list ? n_list : b.n_inst;
  if (grid == 0) return hipSuccess;
  if (stair)
    hipLaunchKernelGGL((k_pcg<NG, R, true>), dim3(grid), dim3(NG * 16), smem, st, b, S, Pinv, gamma, lambda,
                       max_iter, tol, status, inst_list, iters, flag, eta);
  else
    hipLaunchKernelGGL((k_pcg<NG, R, false>), dim3(grid), dim3(NG * 16), smem, st, b, S, Pinv, gamma, lambda,
                       max_iter, tol, status, inst_list, iters, flag, eta);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_pcg_stream(const Batch &b, uint32_t max_n, const float *S, const float *Pinv,
                             const float *gamma, float *lambda, uint32_t max_iter, float tol,
                             const uint32_t *status, const uint32_t *inst_list, uint32_t n_list,
                             uint32_t *iters, uint32_t *flag, float *eta, bool stair, hipStream_t st) {
  if (b.n_inst == 0) return hipSuccess;
#define MPC_PCG_ARGS \
  b, max_n, S, Pinv, gamma, lambda, max_iter, tol, status, inst_list, n_list, iters, flag, eta, stair, st
  // 16 groups (256 lanes) up to N = 256, then 32 / 64 groups; R = rows per group.
  if (max_n <= 16) return launch_pcg_t<16, 1>(MPC_PCG_ARGS);
  if (max_n <= 32) return launch_pcg_t<16, 2>(MPC_PCG_ARGS);
  if (max_n <= 64) return launch_pcg_t<16, 4>(MPC_PCG_ARGS);
  if (max_n <= 128) return launch_pcg_t<16, 8>(MPC_PCG_ARGS);
  if (max_n <= 256) return launch_pcg_t<16, 16>(MPC_PCG_ARGS);
  if (max_n <= 512) return launch_pcg_t<32, 16>(MPC_PCG_ARGS);
#undef MPC_PCG_ARGS
  return hipErrorInvalidValue;
}

}  // namespace mpc
