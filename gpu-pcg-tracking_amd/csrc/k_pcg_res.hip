// k_pcg_res.hip -- PCG with the Schur system resident on chip (gfx950).
//
// Same algorithm and exit rule as k_pcg.hip (GPU-PCG pcg<T,14,N>, call
// site toplevel.cuh:143-157), but every instance's matrix is read from HBM
// once per solve instead of once per iteration:
//   - off-diagonal blocks S(b,b-1) in LDS (196 floats per knot), read as
//     rows for S(b,b-1) p_{b-1} and as columns for S(b+1,b)^T p_{b+1};
//   - row i of the diagonal block S(b,b) and of Pinv(b,b) in the registers
//     of lane i (2 x 14 floats per knot and lane).
// A 1024-lane workgroup holds up to 128 knots (R = 2 knots per 16-lane
// group, 64 groups).  Several instances share a workgroup ("bin"): each
// instance owns a whole number of waves, so its dot products are
// wave64 butterflies plus a fixed-order sum over its own waves, and its
// arithmetic does not depend on which bin it is packed into.  Instances in
// a bin iterate in lockstep; a converged instance idles until the bin is
// done.  Per iteration: 3 workgroup barriers (p exchange, p.Sp, r.z); the
// residual exchange for the block-diagonal preconditioner stays inside a
// 16-lane group (one wave).
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int NT = 1024;
constexpr int NGRP = NT / 16;  // 64 groups
constexpr int NWAV = NT / 64;  // 16 waves
constexpr int RK = 2;          // knots per group
constexpr int KNOTS = NGRP * RK;
constexpr int MAXI = 8;        // instances per bin

struct ResShared {
  float L[KNOTS * NXX];  // S(b,b-1), column-major per knot slot
  float pv[KNOTS * NX];
  float rv[KNOTS * NX];
  float red[2][NWAV];
  uint32_t act[MAXI];
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum of v over the waves [w0, w0 + nw) of the workgroup, fixed order
__device__ __forceinline__ float seg_sum(float v, float *red, int w0, int nw) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float s = 0.f;
  for (int w = 0; w < nw; ++w) s += red[w0 + w];
  return s;
}

__global__ void __launch_bounds__(NT) k_pcg_res(Batch b, const uint4 *__restrict__ items,
                                                const uint32_t *__restrict__ bin_off,
                                                const float *__restrict__ S,
                                                const float *__restrict__ Pinv,
                                                const float *__restrict__ gamma,
                                                float *__restrict__ lambda, uint32_t max_iter,
                                                float tol, const uint32_t *__restrict__ status,
                                                uint32_t *__restrict__ iters_out,
                                                uint32_t *__restrict__ flag_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  ResShared &sh = *reinterpret_cast<ResShared *>(smem_raw);
  const int tid = threadIdx.x, g = tid >> 4, i = tid & 15;

  // my instance in this bin (items: {inst, first group, groups, -}); a null
  // item table means one instance (0) starting at group 0
  uint32_t i0 = 0, ni = 1;
  if (items) {
    i0 = bin_off[blockIdx.x];
    ni = bin_off[blockIdx.x + 1] - i0;
  }
  int slot = -1;
  uint32_t inst = 0, g0 = 0, ng = NGRP;
  // a lone instance uses the same group count as inside a bin
  const uint32_t ng1 = ((inst_len(b, 0) + RK - 1) / RK + 3u) & ~3u;
  for (uint32_t j = 0; j < ni; ++j) {
    const uint4 it = items ? items[i0 + j] : make_uint4(0u, 0u, ng1, 0u);
    if ((uint32_t)g >= it.y && (uint32_t)g < it.y + it.z) {
      slot = (int)j;
      inst = it.x;
      g0 = it.y;
      ng = it.z;
    }
  }
  const bool mine = slot >= 0;
  const uint32_t N = mine ? inst_len(b, inst) : 0u;
  const uint32_t K0 = mine ? inst_first(b, inst) : 0u;
  const bool live = mine && !(status && !(status[inst] & ST_ACTIVE));
  const int lg = (int)(g - g0);
  const int w0 = (int)(g0 >> 2), nw = (int)(ng >> 2);
  const uint32_t lbase = g0 * RK;  // first knot slot of the instance
  const bool lane_ok = live && i < NX;

  // ---- load: L blocks into LDS, D / Pinv rows into registers
  if (live) {
    const uint32_t nthr = ng * 16, t0 = (uint32_t)(tid - (int)g0 * 16);
    for (uint32_t idx = t0; idx < N * NXX; idx += nthr) {
      const uint32_t bk = idx / NXX, e = idx - bk * NXX;
      sh.L[(lbase + bk) * NXX + e] = S[(size_t)S_STRIDE * (K0 + bk) + e];
    }
  }
  float D[RK][NX], P[RK][NX];
  float lam[RK], r[RK], p[RK], y[RK], z[RK];
  bool has[RK];
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    const uint32_t bk = (uint32_t)lg + ng * t;
    has[t] = lane_ok && bk < N;
    const float *Sd = S + (size_t)S_STRIDE * (K0 + bk) + NXX;
    const float *Pd = Pinv + (size_t)S_STRIDE * (K0 + bk) + NXX;
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      D[t][j] = has[t] ? Sd[i + NX * j] : 0.f;
      P[t][j] = has[t] ? Pd[i + NX * j] : 0.f;
    }
    lam[t] = has[t] ? lambda[(size_t)NX * (K0 + bk) + i] : 0.f;
    if (has[t]) sh.pv[(lbase + bk) * NX + i] = lam[t];
  }
  __syncthreads();

  // y = S x with x staged in sh.pv
  auto smv = [&]() {
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      float acc = 0.f;
      if (has[t]) {
        const uint32_t bk = (uint32_t)lg + ng * t, ks = lbase + bk;
        const float *x0 = sh.pv + ks * NX;
#pragma unroll
        for (int j = 0; j < NX; ++j) acc += D[t][j] * x0[j];
        if (bk > 0) {
          const float *Lb = sh.L + ks * NXX;
          const float *xm = x0 - NX;
#pragma unroll
          for (int j = 0; j < NX; ++j) acc += Lb[i + NX * j] * xm[j];
        }
        if (bk + 1 < N) {
          const float *Ln = sh.L + (ks + 1) * NXX;
          const float *xp = x0 + NX;
#pragma unroll
          for (int j = 0; j < NX; ++j) acc += Ln[NX * i + j] * xp[j];
        }
      }
      y[t] = acc;
    }
  };
  // z = Pinv r (r staged in sh.rv by this group), returns r.z partial
  auto precond = [&]() {
    float part = 0.f;
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      float acc = 0.f;
      if (has[t]) {
        const uint32_t ks = lbase + (uint32_t)lg + ng * t;
        const float *rr = sh.rv + ks * NX;
#pragma unroll
        for (int j = 0; j < NX; ++j) acc += P[t][j] * rr[j];
        part += r[t] * acc;
      }
      z[t] = acc;
    }
    return part;
  };
  auto stage_r = [&]() {
#pragma unroll
    for (int t = 0; t < RK; ++t)
      if (has[t]) sh.rv[(lbase + (uint32_t)lg + ng * t) * NX + i] = r[t];
    wave_sync();
  };

  // ---- prologue: r = gamma - S lambda, z = Pinv r, p = z, eta = r.z
  smv();
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    const uint32_t bk = (uint32_t)lg + ng * t;
    r[t] = has[t] ? gamma[(size_t)NX * (K0 + bk) + i] - y[t] : 0.f;
  }
  stage_r();
  float eta = seg_sum(precond(), sh.red[0], w0, nw);
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    p[t] = z[t];
    if (has[t]) sh.pv[(lbase + (uint32_t)lg + ng * t) * NX + i] = p[t];
  }
  uint32_t it = 0;
  bool conv = !live || fabsf(eta) < tol;

  for (;;) {
    if (mine && lg == 0 && i == 0) sh.act[slot] = (!conv && it < max_iter) ? 1u : 0u;
    __syncthreads();  // p (and the flags) visible to the workgroup
    uint32_t any = 0;
    for (uint32_t j = 0; j < ni; ++j) any |= sh.act[j];
    if (!any) break;
    const bool run = live && !conv && it < max_iter;
    smv();
    float part = 0.f;
#pragma unroll
    for (int t = 0; t < RK; ++t) part += p[t] * y[t];
    const float pAp = seg_sum(run ? part : 0.f, sh.red[1], w0, nw);
    if (run) {
      const float alpha = eta / pAp;
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        lam[t] += alpha * p[t];
        r[t] -= alpha * y[t];
      }
    }
    stage_r();
    const float rz = precond();
    const float eta_new = seg_sum(run ? rz : 0.f, sh.red[0], w0, nw);
    if (run) {
      ++it;
      if (fabsf(eta_new) < tol) {
        conv = true;
      } else {
        const float beta = eta_new / eta;
        eta = eta_new;
#pragma unroll
        for (int t = 0; t < RK; ++t) {
          p[t] = z[t] + beta * p[t];
          if (has[t]) sh.pv[(lbase + (uint32_t)lg + ng * t) * NX + i] = p[t];
        }
      }
    }
  }

  if (!live) return;
#pragma unroll
  for (int t = 0; t < RK; ++t)
    if (has[t]) lambda[(size_t)NX * (K0 + (uint32_t)lg + ng * t) + i] = lam[t];
  if (lg == 0 && i == 0) {
    if (iters_out) iters_out[inst] = it;
    if (flag_out) flag_out[inst] = conv ? 0u : 1u;
  }
}

}  // namespace

uint32_t pcg_res_groups(uint32_t N) {
  const uint32_t groups = (N + RK - 1) / RK;
  return (groups + 3u) & ~3u;  // whole waves
}
uint32_t pcg_res_max_knots() { return KNOTS; }
uint32_t pcg_res_max_items() { return MAXI; }

hipError_t launch_pcg_res(const Batch &b, const uint4 *items, const uint32_t *bin_off, uint32_t n_bins,
                          const float *S, const float *Pinv, const float *gamma, float *lambda,
                          uint32_t max_iter, float tol, const uint32_t *status, uint32_t *iters,
                          uint32_t *flag, hipStream_t st) {
  if (n_bins == 0) return hipSuccess;
  static bool attr_set = false;
  const size_t smem = sizeof(ResShared);
  if (!attr_set) {
    // allow more than the default dynamic LDS; the launch reports any error
    (void)hipFuncSetAttribute((const void *)k_pcg_res, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    (void)hipGetLastError();
    attr_set = true;
  }
  hipLaunchKernelGGL(k_pcg_res, dim3(n_bins), dim3(NT), smem, st, b, items, bin_off, S, Pinv, gamma, lambda,
                     max_iter, tol, status, iters, flag);
  return hipGetLastError();
}

}  // namespace mpc
