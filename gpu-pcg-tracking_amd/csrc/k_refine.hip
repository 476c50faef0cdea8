// k_refine.hip -- mixed-precision iterative refinement of the KKT solves of
// the differentiable QP layer (gfx950).
//
// The layer's KKT system K w = b, w = [dz; lambda], is formed in fp32 and
// solved through the fp32 Schur / PCG pipeline.  The entries of K are
// accurate to fp32 rounding, but the solve loses digits with the
// conditioning of S (rho = 1e-3 and control weights near 1e-4 leave
// relative errors of 1e-3..1e-2 in w).  One refinement step:
//   r = b - K w        fp64 accumulation over the fp32 entries of K and w
//   s = max |r|        per instance; r / s is the right-hand side
//   gamma of r / s     with the S, Pinv and G~^-1 already formed
//   PCG, dz recovery   -> dw,   w += s dw   (fp64 add, fp32 store)
// Each step contracts the error of w by about the inner solve's relative
// accuracy, down to the fp32 storage of w.
//
// Blocks (C stores -A_k, -B_k; G~ = G + rho I; constraint row 0 is dx_0 = 0):
//   r_x(k)   = q_k - Q~_k dx_k - lambda_k + A_k^T lambda_{k+1}
//   r_u(k)   = r_k - R~_k du_k + B_k^T lambda_{k+1}               (k < N-1)
//   r_c(0)   = -dx_0
//   r_c(k+1) = c_{k+1} - dx_{k+1} + A_k dx_k + B_k du_k
// and gamma of a right-hand side [g; c] honours c_0:
//   gamma_b = c_b - (Q~_b^-1 q_b - A Q~^-1 q_{b-1} - B R~^-1 r_{b-1}),  gamma_0 = c_0 - Q~_0^-1 q_0.
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int GROUP = 16;
constexpr int GROUPS_PER_WG = 4;
constexpr int WG = GROUP * GROUPS_PER_WG;
constexpr int NT_NORM = 256;

struct ResLds {
  float dx[16], du[8], l1[16];
};

// KKT residual of w = [dz (xu layout); lambda (14 per knot)] into rz (21 per
// knot, control part of the last knot 0) and rc (14 per knot)
__global__ void __launch_bounds__(WG) k_kkt_residual(Batch b, const float *__restrict__ G,
                                                     const float *__restrict__ C, const float *__restrict__ g,
                                                     const float *__restrict__ cvec,
                                                     const float *__restrict__ rho_dev,
                                                     const float *__restrict__ dz, const float *__restrict__ lam,
                                                     float *__restrict__ rz, float *__restrict__ rc) {
  __shared__ ResLds lds[GROUPS_PER_WG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t gk_raw = blockIdx.x * GROUPS_PER_WG + grp;
  const bool valid = gk_raw < b.total_knots;
  const uint32_t gk = valid ? gk_raw : b.total_knots - 1;
  const KnotRef kr = locate(b, gk);
  const bool has_next = kr.k + 1 < kr.N;
  const double rho = rho_dev[kr.inst];
  const float *dzk = dz + (size_t)NXU * kr.K0 - (size_t)NU * kr.inst + (size_t)NXU * kr.k;
  ResLds &L = lds[grp];
  if (lane < NX) {
    L.dx[lane] = dzk[lane];
    L.l1[lane] = has_next ? lam[(size_t)NX * (gk + 1) + lane] : 0.f;
  }
  if (lane < NU) L.du[lane] = has_next ? dzk[NX + lane] : 0.f;
  __syncthreads();
  if (!valid) return;
  const float *Q = G + (size_t)G_STRIDE * gk, *R = Q + NXX;
  const float *Ca = C + (size_t)C_STRIDE * gk, *Cb = Ca + NXX;
  if (lane < NX) {
    double ax = (double)g[(size_t)NXU * gk + lane] - (double)lam[(size_t)NX * gk + lane] - rho * L.dx[lane];
#pragma unroll
    for (int j = 0; j < NX; ++j) ax -= (double)Q[lane + NX * j] * L.dx[j];
    if (has_next) {
      double ac = (cvec ? (double)cvec[(size_t)NX * (gk + 1) + lane] : 0.0) - (double)dzk[NXU + lane];
#pragma unroll
      for (int l = 0; l < NX; ++l) {
        ax -= (double)Ca[l + NX * lane] * L.l1[l];
        ac -= (double)Ca[lane + NX * l] * L.dx[l];
      }
#pragma unroll
      for (int l = 0; l < NU; ++l) ac -= (double)Cb[lane + NX * l] * L.du[l];
      rc[(size_t)NX * (gk + 1) + lane] = (float)ac;
    }
    rz[(size_t)NXU * gk + lane] = (float)ax;
    if (kr.k == 0) rc[(size_t)NX * gk + lane] = -L.dx[lane];
  }
  if (lane < NU) {
    double au = 0.0;
    if (has_next) {
      au = (double)g[(size_t)NXU * gk + NX + lane] - rho * L.du[lane];
#pragma unroll
      for (int j = 0; j < NU; ++j) au -= (double)R[lane + NU * j] * L.du[j];
#pragma unroll
      for (int l = 0; l < NX; ++l) au -= (double)Cb[l + NX * lane] * L.l1[l];
    }
    rz[(size_t)NXU * gk + NX + lane] = (float)au;
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// per instance: s = max |[rz; rc]| (1 if zero), then rz /= s, rc /= s
__global__ void __launch_bounds__(NT_NORM) k_rhs_normalise(Batch b, float *__restrict__ rz, float *__restrict__ rc,
                                                           float *__restrict__ scale) {
  __shared__ float part[NT_NORM / 64];
  const uint32_t inst = blockIdx.x;
  const uint32_t K0 = inst_first(b, inst), N = inst_len(b, inst);
  float *z = rz + (size_t)NXU * K0, *c = rc + (size_t)NX * K0;
  const uint32_t nz = NXU * N, nc = NX * N;
  float m = 0.f;
  for (uint32_t i = threadIdx.x; i < nz; i += NT_NORM) m = fmaxf(m, fabsf(z[i]));
  for (uint32_t i = threadIdx.x; i < nc; i += NT_NORM) m = fmaxf(m, fabsf(c[i]));
  m = wave_max(m);
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x / 64] = m;
  __syncthreads();
  float s = part[0];
#pragma unroll
  for (int w = 1; w < NT_NORM / 64; ++w) s = fmaxf(s, part[w]);
  if (!(s > 0.f)) s = 1.f;
  const float inv = 1.f / s;
  for (uint32_t i = threadIdx.x; i < nz; i += NT_NORM) z[i] *= inv;
  for (uint32_t i = threadIdx.x; i < nc; i += NT_NORM) c[i] *= inv;
  if (threadIdx.x == 0) scale[inst] = s;
}

struct GamLds {
  float q0[16], q1[16], r0[8], t[16], tu[8];
};

// gamma of the right-hand side [g; c] (c_0 included) from G~^-1 and C
__global__ void __launch_bounds__(WG) k_form_gamma(Batch b, const float *__restrict__ Ginv,
                                                   const float *__restrict__ C, const float *__restrict__ g,
                                                   const float *__restrict__ cvec, float *__restrict__ gamma) {
  __shared__ GamLds lds[GROUPS_PER_WG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t gk_raw = blockIdx.x * GROUPS_PER_WG + grp;
  const bool valid = gk_raw < b.total_knots;
  const uint32_t gk = valid ? gk_raw : b.total_knots - 1;
  const KnotRef kr = locate(b, gk);
  const bool row0 = kr.k == 0;
  const uint32_t kp = row0 ? gk : gk - 1;  // previous knot (own knot for row 0, unused there)
  GamLds &L = lds[grp];
  if (lane < NX) {
    L.q1[lane] = g[(size_t)NXU * gk + lane];
    L.q0[lane] = g[(size_t)NXU * kp + lane];
  }
  if (lane < NU) L.r0[lane] = g[(size_t)NXU * kp + NX + lane];
  __syncthreads();
  const float *Q1i = Ginv + (size_t)G_STRIDE * gk;
  const float *Q0i = Ginv + (size_t)G_STRIDE * kp, *R0i = Q0i + NXX;
  float a1 = 0.f;
  if (lane < NX) {
    float t = 0.f;
#pragma unroll
    for (int l = 0; l < NX; ++l) {
      a1 += Q1i[lane + NX * l] * L.q1[l];
      t += Q0i[lane + NX * l] * L.q0[l];
    }
    L.t[lane] = t;
  }
  if (lane < NU) {
    float tu = 0.f;
    if (!row0) {
#pragma unroll
      for (int l = 0; l < NU; ++l) tu += R0i[lane + NU * l] * L.r0[l];
    }
    L.tu[lane] = tu;
  }
  __syncthreads();
  if (!valid || lane >= NX) return;
  const float cb = cvec ? cvec[(size_t)NX * gk + lane] : 0.f;
  if (row0) {
    gamma[(size_t)NX * gk + lane] = cb - a1;
    return;
  }
  // A t = -C_A t, B tu = -C_B tu (C of the previous knot)
  const float *Ca = C + (size_t)C_STRIDE * kp, *Cb = Ca + NXX;
  float a2 = 0.f, a3 = 0.f;
#pragma unroll
  for (int l = 0; l < NX; ++l) a2 -= Ca[lane + NX * l] * L.t[l];
#pragma unroll
  for (int l = 0; l < NU; ++l) a3 -= Cb[lane + NX * l] * L.tu[l];
  gamma[(size_t)NX * gk + lane] = cb - (a1 - a2 - a3);
}

// w += s dw: dz (xu layout) and lambda (14 per knot), fp64 add
__global__ void k_refine_update(Batch b, const float *__restrict__ scale, const float *__restrict__ ddz,
                                const float *__restrict__ dlam, float *__restrict__ dz, float *__restrict__ lam) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= b.total_knots * NXU) return;
  const uint32_t gk = idx / NXU, j = idx - gk * NXU;
  const KnotRef kr = locate(b, gk);
  const double s = scale[kr.inst];
  if (j < NX || kr.k + 1 < kr.N) {
    const size_t o = (size_t)NXU * kr.K0 - (size_t)NU * kr.inst + (size_t)NXU * kr.k + j;
    dz[o] = (float)((double)dz[o] + s * (double)ddz[o]);
  }
  if (j < NX) {
    const size_t o = (size_t)NX * gk + j;
    lam[o] = (float)((double)lam[o] + s * (double)dlam[o]);
  }
}

}  // namespace

hipError_t launch_kkt_residual(const Batch &b, const float *G, const float *C, const float *g, const float *c,
                               const float *rho_dev, const float *dz, const float *lam, float *rz, float *rc,
                               hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kkt_residual, dim3((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG), 0, st,
                     b, G, C, g, c, rho_dev, dz, lam, rz, rc);
  return hipGetLastError();
}

hipError_t launch_rhs_normalise(const Batch &b, float *rz, float *rc, float *scale, hipStream_t st) {
  if (b.n_inst == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rhs_normalise, dim3(b.n_inst), dim3(NT_NORM), 0, st, b, rz, rc, scale);
  return hipGetLastError();
}

hipError_t launch_form_gamma(const Batch &b, const float *Ginv, const float *C, const float *g, const float *c,
                             float *gamma, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_form_gamma, dim3((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG), 0, st, b,
                     Ginv, C, g, c, gamma);
  return hipGetLastError();
}

hipError_t launch_refine_update(const Batch &b, const float *scale, const float *ddz, const float *dlam, float *dz,
                                float *lam, hipStream_t st) {
  const uint32_t n = b.total_knots * NXU;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_refine_update, dim3((n + 255) / 256), dim3(256), 0, st, b, scale, ddz, dlam, dz, lam);
  return hipGetLastError();
}

}  // namespace mpc
