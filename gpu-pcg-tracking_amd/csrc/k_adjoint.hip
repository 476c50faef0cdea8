// k_adjoint.hip -- differentiable QP layer kernels (gfx950): the tracking
// loss of a trajectory, the adjoint right-hand side and the per-knot terms
// of the parameter gradient.  One thread per knot (end-effector kinematics
// and its Jacobian in registers).
//
// The layer (the build's differentiation growth direction; the reference
// has no counterpart): at an iterate xu, the QP of the SQP step
//   K w = b,  K = [G~ C^T; C 0] (G~ = G + rho I),  w = [dz; lambda],  b = [g; c]
// is solved and the full-step solution x* = xu - dz is returned.  With
// L = sum_k |ee(q*_k) - ref_k|^2 and a = dL/dx*, the adjoint
// v = K^-1 [-a; 0] is the Schur solve of gamma(g := -a, c := 0) followed by
// the dz recovery, and for the cost parameters (per-dimension weights
// wx, wu and the reference points; C and c do not depend on them)
//   dL/dtheta = v_z^T (dg/dtheta - dG/dtheta dz)   (+ the direct ref term).
// Per knot k with cost state x~ (x_k; knot N-1 uses x_{N-2}, as the
// reference's terminal cost) and g_e = J^T q_ee (ee(x~) - ref_k):
//   dL/dwx_j += v_x,j (x~_j - dz_x,j),  dL/dwu_j += v_u,j (u_k,j - dz_u,j)  (k < N-1)
//   dL/dref_k,b += (v.d) - (v.d)(g_e.dz) - (v.g_e)(d.dz),  d = -q_ee J[b, :]
// (q parts only: the Gauss-Newton block is g_e g_e^T + diag(wx_q)).
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int NT = 128;

// end-effector position and Jacobian J[3 j + b] = d ee_b / d q_j
__device__ __forceinline__ void ee_jacobian(const float *q, float *ee, float *J) {
  float s[NQ], c[NQ], zax[21], org[21], R[9];
#pragma unroll
  for (int j = 0; j < NQ; ++j) sincosf(q[j], &s[j], &c[j]);
  fk_frames(s, c, zax, org, R);
  ee[0] = org[18];
  ee[1] = org[19];
  ee[2] = org[20];
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const float d0 = ee[0] - org[3 * j], d1 = ee[1] - org[3 * j + 1], d2 = ee[2] - org[3 * j + 2];
    const float z0 = zax[3 * j], z1 = zax[3 * j + 1], z2 = zax[3 * j + 2];
    J[3 * j + 0] = z1 * d2 - z2 * d1;
    J[3 * j + 1] = z2 * d0 - z0 * d2;
    J[3 * j + 2] = z0 * d1 - z1 * d0;
  }
}

// loss terms of the trajectory xu: per knot sum_b (ee_b - ref_b)^2 (into
// knot_loss[K0 + k]); dLdx (xu layout, fully written) and dLdref (6 per knot)
__global__ void __launch_bounds__(NT) k_qp_loss(Batch b, const float *__restrict__ xu,
                                                const float *__restrict__ ee_ref, float *__restrict__ knot_loss,
                                                float *__restrict__ dLdx, float *__restrict__ dLdref) {
  const uint32_t gk = blockIdx.x * NT + threadIdx.x;
  if (gk >= b.total_knots) return;
  const KnotRef kr = locate(b, gk);
  const size_t xo = (size_t)NXU * kr.K0 - (size_t)NU * kr.inst + (size_t)NXU * kr.k;
  const float *x = xu + xo;
  float q[NQ], ee[3], J[21];
#pragma unroll
  for (int j = 0; j < NQ; ++j) q[j] = x[j];
  ee_jacobian(q, ee, J);
  const float *ref = ee_ref + (size_t)6 * gk;
  float e[3], L = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    e[a] = ee[a] - ref[a];
    L += e[a] * e[a];
  }
  knot_loss[gk] = L;
  if (dLdx) {
    float *d = dLdx + xo;
#pragma unroll
    for (int j = 0; j < NQ; ++j) d[j] = 2.f * (e[0] * J[3 * j] + e[1] * J[3 * j + 1] + e[2] * J[3 * j + 2]);
    const int n = kr.k + 1 < kr.N ? NXU : NX;
    for (int j = NQ; j < n; ++j) d[j] = 0.f;
  }
  if (dLdref) {
#pragma unroll
    for (int a = 0; a < 3; ++a) dLdref[(size_t)6 * gk + a] = -2.f * e[a];
#pragma unroll
    for (int a = 3; a < 6; ++a) dLdref[(size_t)6 * gk + a] = 0.f;
  }
}

// workspace g (21 per knot) = -a (xu layout); the last knot's control part is 0
__global__ void __launch_bounds__(NT) k_pack_neg(Batch b, const float *__restrict__ a, float *__restrict__ gz) {
  const uint32_t idx = blockIdx.x * NT + threadIdx.x;
  if (idx >= b.total_knots * NXU) return;
  const uint32_t gk = idx / NXU, j = idx - gk * NXU;
  const KnotRef kr = locate(b, gk);
  const bool has = j < NX || kr.k + 1 < kr.N;
  gz[idx] = has ? -a[(size_t)NXU * kr.K0 - (size_t)NU * kr.inst + (size_t)NXU * kr.k + j] : 0.f;
}

// per-knot parameter-gradient terms (21 per knot: wx then wu) and the
// implicit part of dL/dref added into gref (6 per knot)
__global__ void __launch_bounds__(NT) k_adj_params(Batch b, Params p, const float *__restrict__ xu,
                                                   const float *__restrict__ ee_ref, const float *__restrict__ dz,
                                                   const float *__restrict__ v, float *__restrict__ knot_w,
                                                   float *__restrict__ gref) {
  const uint32_t gk = blockIdx.x * NT + threadIdx.x;
  if (gk >= b.total_knots) return;
  const KnotRef kr = locate(b, gk);
  const size_t io = (size_t)NXU * kr.K0 - (size_t)NU * kr.inst;
  const bool dyn = kr.k + 1 < kr.N;
  const uint32_t kc = dyn ? kr.k : ((p.flags & FLAG_LAST_KNOT_OWN_STATE) ? kr.k : kr.k - 1);
  const float *xc = xu + io + (size_t)NXU * kc;
  const float *vk = v + io + (size_t)NXU * kr.k, *dk = dz + io + (size_t)NXU * kr.k;
  float *out = knot_w + (size_t)NXU * gk;
  float vx[NX], dx[NX];
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    vx[j] = vk[j];
    dx[j] = dk[j];
    out[j] = vx[j] * (xc[j] - dx[j]);
  }
  const float *u = xu + io + (size_t)NXU * kr.k + NX;
#pragma unroll
  for (int j = 0; j < NU; ++j) out[NX + j] = dyn ? vk[NX + j] * (u[j] - dk[NX + j]) : 0.f;
  // reference points: through g_e = J^T q_ee (ee - ref) in q and in the rank-1 block
  float q[NQ], ee[3], J[21];
#pragma unroll
  for (int j = 0; j < NQ; ++j) q[j] = xc[j];
  ee_jacobian(q, ee, J);
  const float *ref = ee_ref + (size_t)6 * gk;
  const float w = p.q_ee;
  const float e0 = w * (ee[0] - ref[0]), e1 = w * (ee[1] - ref[1]), e2 = w * (ee[2] - ref[2]);
  float vge = 0.f, gedz = 0.f;
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const float ge = J[3 * j] * e0 + J[3 * j + 1] * e1 + J[3 * j + 2] * e2;
    vge += vx[j] * ge;
    gedz += ge * dx[j];
  }
#pragma unroll
  for (int bb = 0; bb < 3; ++bb) {
    float vd = 0.f, ddz = 0.f;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const float dge = -w * J[3 * j + bb];
      vd += vx[j] * dge;
      ddz += dge * dx[j];
    }
    gref[(size_t)6 * gk + bb] += vd - vd * gedz - vge * ddz;
  }
}

// out[inst * S + c] = sum over the instance's knots (in order) of vals[gk * S + c]
__global__ void k_reduce_knots(Batch b, const float *__restrict__ vals, uint32_t S, float *__restrict__ out) {
  const uint32_t inst = blockIdx.x, c = threadIdx.x;
  if (c >= S) return;
  const uint32_t K0 = inst_first(b, inst), N = inst_len(b, inst);
  float s = 0.f;
  for (uint32_t k = 0; k < N; ++k) s += vals[(size_t)(K0 + k) * S + c];
  out[(size_t)inst * S + c] = s;
}

__global__ void k_sub(uint32_t n, const float *__restrict__ a, const float *__restrict__ bb, float *__restrict__ o) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = a[i] - bb[i];
}

__global__ void k_fill(uint32_t n, float v, float *__restrict__ o) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = v;
}

}  // namespace

hipError_t launch_qp_loss(const Batch &b, const float *xu, const float *ee_ref, float *knot_loss, float *dLdx,
                          float *dLdref, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_qp_loss, dim3((b.total_knots + NT - 1) / NT), dim3(NT), 0, st, b, xu, ee_ref, knot_loss,
                     dLdx, dLdref);
  return hipGetLastError();
}

hipError_t launch_pack_neg(const Batch &b, const float *a, float *gz, hipStream_t st) {
  const uint32_t n = b.total_knots * NXU;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_neg, dim3((n + NT - 1) / NT), dim3(NT), 0, st, b, a, gz);
  return hipGetLastError();
}

hipError_t launch_adj_params(const Batch &b, const Params &p, const float *xu, const float *ee_ref,
                             const float *dz, const float *v, float *knot_w, float *gref, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_params, dim3((b.total_knots + NT - 1) / NT), dim3(NT), 0, st, b, p, xu, ee_ref, dz,
                     v, knot_w, gref);
  return hipGetLastError();
}

hipError_t launch_reduce_knots(const Batch &b, const float *vals, uint32_t S, float *out, hipStream_t st) {
  if (b.n_inst == 0) return hipSuccess;
  if (S > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_reduce_knots, dim3(b.n_inst), dim3(64), 0, st, b, vals, S, out);
  return hipGetLastError();
}

hipError_t launch_sub(uint32_t n, const float *a, const float *bb, float *o, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sub, dim3((n + 255) / 256), dim3(256), 0, st, n, a, bb, o);
  return hipGetLastError();
}

hipError_t launch_fill(uint32_t n, float v, float *o, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, st, n, v, o);
  return hipGetLastError();
}

}  // namespace mpc
