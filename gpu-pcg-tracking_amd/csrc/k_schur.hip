// k_schur.hip -- Schur complement, block-Jacobi preconditioner and dz
// recovery for gfx950.
//
// Mapping: one 16-lane group per block row, lane j owns column j of every
// 14x14 block (registers), 4 groups per 64-lane workgroup.  Block inverses
// are Gauss-Jordan eliminations on [M | I] with the pivot column broadcast
// through the group's LDS slice (the regularised blocks are symmetric
// positive definite, so no pivoting is needed).
//
// Math (reference: oldschur.cuh:518-963, signs as there; C stores -A, -B):
//   Q~_k = Q_k + rho I, R~_k = R_k + rho I
//   S(b,b-1) = A Q~_{b-1}^-1          S(b-1,b) = S(b,b-1)^T
//   S(b,b)   = -(A Q~^-1 A^T + B R~^-1 B^T + Q~_b^-1),  S(0,0) = -Q~_0^-1
//   gamma_b  = c_b - (Q~_b^-1 q_b - A Q~^-1 q_{b-1} - B R~^-1 r_{b-1}),
//   gamma_0  = -Q~_0^-1 q_0 (c_0 is not used, oldschur.cuh:639-654)
//   Pinv(b,b) = S(b,b)^-1, Pinv(0,0) = -Q~_0
// G^-1 goes to a separate buffer (the reference overwrites G in place,
// which races between neighbouring block rows).
// dz (oldschur.cuh:966-1105):
//   dx_k = Q~_k^-1 (q_k - lambda_k + A_k^T lambda_{k+1}),  du_k = R~_k^-1 (r_k + B_k^T lambda_{k+1})
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int GROUP = 16;
constexpr int GROUPS_PER_WG = 4;
constexpr int WG = GROUP * GROUPS_PER_WG;

struct SchurLds {
  float A[NXX];    // A_k (column-major)
  float B[98];     // B_k
  float Qi[NXX];   // Q~_k^-1
  float Q1i[NXX];  // Q~_{k+1}^-1
  float Ri[49];    // R~_k^-1
  float AQ[NXX];   // A Q~_k^-1
  float BR[98];    // B R~_k^-1
  float th[NXX];   // theta
  float piv[16];
  float qk[16], q1[16], rk[8], cb[16];
};

// In-place Gauss-Jordan inverse of an n x n column-major LDS matrix
// mat (n <= 14) held as columns by lanes 0..n-1.  The inverse column j is
// returned in x (registers of lane j) and written back to out (may alias mat).
template <int n>
__device__ __forceinline__ void gj_inverse(int lane, const float *mat, float *piv, float *x,
                                           float *out, float add_diag) {
  float a[n];
  const int j = lane < n ? lane : n - 1;
#pragma unroll
  for (int r = 0; r < n; ++r) {
    a[r] = mat[r + n * j] + ((r == j) ? add_diag : 0.f);
    x[r] = (r == j) ? 1.f : 0.f;
  }
#pragma unroll
  for (int p = 0; p < n; ++p) {
    __syncthreads();
    if (lane == p) {
#pragma unroll
      for (int r = 0; r < n; ++r) piv[r] = a[r];
    }
    __syncthreads();
    float pc[n];
#pragma unroll
    for (int r = 0; r < n; ++r) pc[r] = piv[r];
    const float inv = 1.f / pc[p];
    const float ap = a[p] * inv, xp = x[p] * inv;
#pragma unroll
    for (int r = 0; r < n; ++r) {
      if (r == p) continue;
      a[r] -= pc[r] * ap;
      x[r] -= pc[r] * xp;
    }
    a[p] = ap;
    x[p] = xp;
  }
  __syncthreads();
  if (lane < n) {
#pragma unroll
    for (int r = 0; r < n; ++r) out[r + n * lane] = x[r];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(WG) k_form_schur(Batch b, const float *__restrict__ G,
                                                   const float *__restrict__ C,
                                                   const float *__restrict__ g,
                                                   const float *__restrict__ cvec,
                                                   const float *__restrict__ rho_dev, float rho_s,
                                                   const uint32_t *__restrict__ status,
                                                   float *__restrict__ S, float *__restrict__ Pinv,
                                                   float *__restrict__ gamma, float *__restrict__ Ginv,
                                                   int zero_unused) {
  __shared__ SchurLds lds[GROUPS_PER_WG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t gk_raw = blockIdx.x * GROUPS_PER_WG + grp;
  bool valid = gk_raw < b.total_knots;
  const uint32_t gk = valid ? gk_raw : b.total_knots - 1;
  const KnotRef kr = locate(b, gk);
  if (status && !(status[kr.inst] & ST_ACTIVE)) valid = false;
  const float rho = rho_dev ? rho_dev[kr.inst] : rho_s;
  SchurLds &L = lds[grp];
  const uint32_t bk = kr.k;
  const size_t row = (size_t)kr.K0 + bk;  // global block row
  float *Srow = S + (size_t)S_STRIDE * row;
  float *Prow = Pinv + (size_t)S_STRIDE * row;
  float x[NX];
  // Every group runs the same barrier sequence; block row 0 stages its own
  // knot in the 'previous knot' slots and only keeps Q~_0^-1.
  const bool row0 = bk == 0;
  // block row b >= 1: k = b - 1
  const size_t kprev = row0 ? row : row - 1;
  const float *Gk = G + (size_t)G_STRIDE * kprev;
  const float *Gk1 = G + (size_t)G_STRIDE * row;
  const float *Ck = C + (size_t)C_STRIDE * kprev;
  for (int i = lane; i < NXX; i += GROUP) {
    L.A[i] = -Ck[i];
    L.th[i] = Gk[i];    // staging for Q_k
    L.AQ[i] = Gk1[i];   // staging for Q_{k+1}
  }
  for (int i = lane; i < 98; i += GROUP) L.B[i] = -Ck[NXX + i];
  for (int i = lane; i < 49; i += GROUP) L.BR[i] = Gk[NXX + i];  // staging for R_k
  if (lane < NX) {
    L.qk[lane] = g[(size_t)NXU * kprev + lane];
    L.q1[lane] = g[(size_t)NXU * row + lane];
    L.cb[lane] = row0 ? 0.f : cvec[(size_t)NX * row + lane];
  }
  if (lane < NU) L.rk[lane] = g[(size_t)NXU * kprev + NX + lane];
  __syncthreads();
  gj_inverse<NX>(lane, L.th, L.piv, x, L.Qi, rho);
  gj_inverse<NX>(lane, L.AQ, L.piv, x, L.Q1i, rho);
  float xr[NU];
  gj_inverse<NU>(lane, L.BR, L.piv, xr, L.Ri, rho);

  // AQ = A Q~_k^-1 (column j), BR = B R~_k^-1 (column j < 7)
  {
    const int j = lane < NX ? lane : NX - 1;
    float qi[NX];
#pragma unroll
    for (int l = 0; l < NX; ++l) qi[l] = L.Qi[l + NX * j];
    float acc[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) acc[r] = 0.f;
#pragma unroll
    for (int l = 0; l < NX; ++l)
#pragma unroll
      for (int r = 0; r < NX; ++r) acc[r] += L.A[r + NX * l] * qi[l];
    float bacc[NX];
    const int ju = lane < NU ? lane : NU - 1;
    float ri[NU];
#pragma unroll
    for (int l = 0; l < NU; ++l) ri[l] = L.Ri[l + NU * ju];
#pragma unroll
    for (int r = 0; r < NX; ++r) bacc[r] = 0.f;
#pragma unroll
    for (int l = 0; l < NU; ++l)
#pragma unroll
      for (int r = 0; r < NX; ++r) bacc[r] += L.B[r + NX * l] * ri[l];
    __syncthreads();
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.AQ[r + NX * lane] = acc[r];
    }
    if (lane < NU) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.BR[r + NX * lane] = bacc[r];
    }
    __syncthreads();
  }
  // theta(:, j) = sum_l AQ(:, l) A(j, l) + sum_l BR(:, l) B(j, l) + Q~_{k+1}^-1(:, j)
  {
    const int j = lane < NX ? lane : NX - 1;
    float acc[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) acc[r] = L.Q1i[r + NX * j];
#pragma unroll
    for (int l = 0; l < NX; ++l) {
      const float ajl = L.A[j + NX * l];
#pragma unroll
      for (int r = 0; r < NX; ++r) acc[r] += L.AQ[r + NX * l] * ajl;
    }
#pragma unroll
    for (int l = 0; l < NU; ++l) {
      const float bjl = L.B[j + NX * l];
#pragma unroll
      for (int r = 0; r < NX; ++r) acc[r] += L.BR[r + NX * l] * bjl;
    }
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.th[r + NX * lane] = acc[r];
    }
    __syncthreads();
  }
  if (valid && row0) {
    // S(0,0) = -Q~_0^-1, gamma_0 = -Q~_0^-1 q_0, Pinv(0,0) = -Q~_0
    const float *Q0 = G + (size_t)G_STRIDE * row;
    for (int i = lane; i < NXX; i += GROUP) {
      Srow[NXX + i] = -L.Qi[i];
      Prow[NXX + i] = -(Q0[i] + ((i % NX == i / NX) ? rho : 0.f));
    }
    if (lane < NX) {
      float acc = 0.f;
#pragma unroll
      for (int l = 0; l < NX; ++l) acc += L.Qi[lane + NX * l] * L.qk[l];
      gamma[(size_t)NX * row + lane] = -acc;
    }
    if (zero_unused)
      for (int i = lane; i < NXX; i += GROUP) {
        Srow[i] = 0.f;
        Prow[i] = 0.f;
        Prow[2 * NXX + i] = 0.f;
      }
  }
  // stores that do not need theta^-1
  if (valid && !row0) {
    for (int i = lane; i < NXX; i += GROUP) {
      Srow[i] = L.AQ[i];
      Srow[NXX + i] = -L.th[i];
    }
    // right block of the previous row: S(b-1, b) = AQ^T
    float *Sprev = S + (size_t)S_STRIDE * kprev + 2 * NXX;
    for (int i = lane; i < NXX; i += GROUP) {
      const int r = i % NX, col = i / NX;
      Sprev[i] = L.AQ[col + NX * r];
    }
    if (lane < NX) {
      float a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
      for (int l = 0; l < NX; ++l) {
        a1 += L.Q1i[lane + NX * l] * L.q1[l];
        a2 += L.AQ[lane + NX * l] * L.qk[l];
      }
#pragma unroll
      for (int l = 0; l < NU; ++l) a3 += L.BR[lane + NX * l] * L.rk[l];
      gamma[(size_t)NX * row + lane] = L.cb[lane] - (a1 - a2 - a3);
    }
    float *Gi = Ginv + (size_t)G_STRIDE * kprev;
    for (int i = lane; i < NXX; i += GROUP) Gi[i] = L.Qi[i];
    for (int i = lane; i < 49; i += GROUP) Gi[NXX + i] = L.Ri[i];
    if (bk + 1 == kr.N) {
      float *Gl = Ginv + (size_t)G_STRIDE * row;
      for (int i = lane; i < NXX; i += GROUP) Gl[i] = L.Q1i[i];
      if (zero_unused)
        for (int i = lane; i < NXX; i += GROUP) {
          Srow[2 * NXX + i] = 0.f;
          Prow[2 * NXX + i] = 0.f;
        }
    }
    if (zero_unused)
      for (int i = lane; i < NXX; i += GROUP) {
        Prow[i] = 0.f;
        if (bk + 1 < kr.N) Prow[2 * NXX + i] = 0.f;
      }
  }
  // Pinv(b,b) = -theta^-1
  gj_inverse<NX>(lane, L.th, L.piv, x, L.AQ, 0.f);
  if (valid && !row0 && lane < NX) {
#pragma unroll
    for (int r = 0; r < NX; ++r) Prow[NXX + r + NX * lane] = -x[r];
  }
}

struct DzLds {
  float t[16];
  float tu[8];
  float lam1[16];
};

__global__ void __launch_bounds__(WG) k_compute_dz(Batch b, const float *__restrict__ Ginv,
                                                   const float *__restrict__ C,
                                                   const float *__restrict__ g,
                                                   const float *__restrict__ lambda,
                                                   const uint32_t *__restrict__ status,
                                                   float *__restrict__ dz) {
  __shared__ DzLds lds[GROUPS_PER_WG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t gk_raw = blockIdx.x * GROUPS_PER_WG + grp;
  bool valid = gk_raw < b.total_knots;
  const uint32_t gk = valid ? gk_raw : b.total_knots - 1;
  const KnotRef kr = locate(b, gk);
  if (status && !(status[kr.inst] & ST_ACTIVE)) valid = false;
  DzLds &L = lds[grp];
  const size_t knot = (size_t)kr.K0 + kr.k;
  const bool has_next = kr.k + 1 < kr.N;
  const float *Ck = C + (size_t)C_STRIDE * knot;
  if (lane < NX) L.lam1[lane] = has_next ? lambda[(size_t)NX * (knot + 1) + lane] : 0.f;
  __syncthreads();
  if (lane < NX) {
    // t_j = q_j - lambda_k,j + (A^T lambda_{k+1})_j, A = -C_A
    float s = 0.f;
    if (has_next) {
#pragma unroll
      for (int l = 0; l < NX; ++l) s -= Ck[l + NX * lane] * L.lam1[l];
    }
    L.t[lane] = g[(size_t)NXU * knot + lane] - lambda[(size_t)NX * knot + lane] + s;
  }
  if (has_next && lane < NU) {
    float s = 0.f;
#pragma unroll
    for (int l = 0; l < NX; ++l) s -= Ck[NXX + l + NX * lane] * L.lam1[l];
    L.tu[lane] = g[(size_t)NXU * knot + NX + lane] + s;
  }
  __syncthreads();
  if (!valid) return;
  float *dzi = dz + (size_t)NXU * kr.K0 - (size_t)NU * kr.inst + (size_t)NXU * kr.k;
  const float *Gi = Ginv + (size_t)G_STRIDE * knot;
  if (lane < NX) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < NX; ++j) acc += Gi[lane + NX * j] * L.t[j];
    dzi[lane] = acc;
  }
  if (has_next && lane < NU) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < NU; ++j) acc += Gi[NXX + lane + NU * j] * L.tu[j];
    dzi[NX + lane] = acc;
  }
}

}  // namespace

hipError_t launch_form_schur(const Batch &b, const float *G, const float *C, const float *g,
                             const float *c, const float *rho_dev, float rho, const uint32_t *status,
                             float *S, float *Pinv, float *gamma, float *Ginv, int zero_unused,
                             hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_form_schur, dim3((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG),
                     0, st, b, G, C, g, c, rho_dev, rho, status, S, Pinv, gamma, Ginv, zero_unused);
  return hipGetLastError();
}

hipError_t launch_compute_dz(const Batch &b, const float *Ginv, const float *C, const float *g,
                             const float *lambda, const uint32_t *status, float *dz, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_compute_dz, dim3((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG),
                     0, st, b, Ginv, C, g, lambda, status, dz);
  return hipGetLastError();
}

}  // namespace mpc
