// k_direct.hip -- direct solve of the Schur system S lambda = gamma of every
// instance by block LDL^T (block Thomas) of the symmetric positive definite
// A = -S (gfx950).
//
// One 16-lane group per instance, 4 groups per 64-lane workgroup; the
// instances are taken longest first, so the groups of a wave run similar
// trip counts.  Lane j holds column j of the current 14x14 blocks.  Forward
// sweep over the knots (A_b = -S(b,b), B_b = -S(b,b-1), r = -gamma):
//   E_b = B_b D_{b-1}^-1,   D_b = A_b - E_b B_b^T,
//   y_b = r_b - E_b y_{b-1},   z_b = D_b^-1 y_b        (D_0 = A_0, y_0 = r_0)
// with D_b^-1 by Gauss-Jordan in registers (D_b is SPD: no pivoting); E_b
// and D_b^-1 go to the workspace for the backward sweep
//   x_{N-1} = z_{N-1},   x_b = z_b - E_{b+1}^T x_{b+1},   lambda = x
// and for `refine` steps of iterative refinement: the residual
// r = -gamma + S lambda of A lambda = -gamma in fp64 from the fp32 blocks,
// the same two sweeps on it with the stored factors, lambda += x.  The fp32
// elimination alone is off by about cond(S) x eps (1e-4..1e-3 relative on
// these systems); each step multiplies that error by about the same
// factor, down to the fp32 storage of lambda.
// The next knot's blocks are loaded while the current one is processed.
//
// Reference: the QDLDL path of sqpSolve (PCG_SOLVE = 0, toplevel.cuh:160-215,
// :314-335) solves the same system by a sparse LDL^T on the host.
#include "common.h"
#include "gj16.h"

namespace mpc {

namespace {

constexpr int GROUP = 16;
constexpr int GROUPS_PER_WG = 4;
constexpr int WG = GROUP * GROUPS_PER_WG;
constexpr int BPL = (NXX + GROUP - 1) / GROUP;  // B entries per lane (13)

struct DirLds {
  float B[NXX];  // B_b (column-major); refinement: D_b^-1
  float E[NXX];  // E_b (column-major)
  float y[16];   // y_{b-1}, then y_b
  float x[16];   // backward sweep: x_{b+1}
};

// workspace slots of knot row (S layout, 588 floats per knot)
__device__ __forceinline__ float *slot_E(float *work, size_t row) { return work + (size_t)S_STRIDE * row; }
__device__ __forceinline__ float *slot_Dinv(float *work, size_t row) { return work + (size_t)S_STRIDE * row + NXX; }
__device__ __forceinline__ float *slot_z(float *work, size_t row) { return work + (size_t)S_STRIDE * row + 2 * NXX; }

struct KnotLoad {
  float a[NX];   // column j of A_b
  float b[BPL];  // lane's share of B_b
  float r;       // r_b, element j
};

__device__ __forceinline__ void load_knot(KnotLoad &o, const float *__restrict__ S, const float *__restrict__ gamma,
                                          size_t row, bool first, int lane, int j) {
  const float *Sr = S + (size_t)S_STRIDE * row;
#pragma unroll
  for (int r = 0; r < NX; ++r) o.a[r] = -Sr[NXX + r + NX * j];
#pragma unroll
  for (int t = 0; t < BPL; ++t) {
    const int i = lane + GROUP * t;
    o.b[t] = (!first && i < NXX) ? -Sr[i] : 0.f;
  }
  o.r = -gamma[(size_t)NX * row + j];
}

// x_{N-1} = z_{N-1}, x_b = z_b - E_{b+1}^T x_{b+1} (every group runs Nmax
// steps so the barriers line up).  zsrc == nullptr: z_b is in lambda and
// lambda <- x; else z_b is in the workspace's z slot and lambda += x.
__device__ __forceinline__ void backward(DirLds &L, float *__restrict__ work, float *__restrict__ lambda,
                                         const float *zsrc, uint32_t K0, uint32_t N, uint32_t Nmax, bool live,
                                         int lane) {
  if (lane < 16) L.x[lane] = 0.f;
  __syncthreads();
  for (int k = (int)Nmax - 1; k >= 0; --k) {
    const bool on = live && (uint32_t)k < N;
    float xj = 0.f;
    if (on && lane < NX) {
      const size_t row = (size_t)K0 + (uint32_t)k;
      xj = zsrc ? slot_z(work, row)[lane] : lambda[NX * row + lane];
      if ((uint32_t)k + 1 < N) {
        const float *Ec = slot_E(work, row + 1) + NX * lane;  // column j of E_{k+1}
#pragma unroll
        for (int r = 0; r < NX; ++r) xj -= Ec[r] * L.x[r];
      }
    }
    __syncthreads();
    if (on && lane < NX) {
      L.x[lane] = xj;
      float *l = lambda + NX * ((size_t)K0 + (uint32_t)k) + lane;
      *l = zsrc ? *l + xj : xj;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(WG) k_direct_solve(Batch b, const uint32_t *__restrict__ order, uint32_t n_list,
                                                     const float *__restrict__ S, const float *__restrict__ gamma,
                                                     const uint32_t *__restrict__ status, float *__restrict__ work,
                                                     float *__restrict__ lambda, uint32_t refine,
                                                     uint32_t *__restrict__ iters, uint32_t *__restrict__ flag) {
  __shared__ DirLds lds[GROUPS_PER_WG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t slot = blockIdx.x * GROUPS_PER_WG + grp;
  const bool in_list = slot < n_list;
  const uint32_t pick = in_list ? slot : n_list - 1;
  const uint32_t inst = order ? order[pick] : pick;
  const uint32_t K0 = inst_first(b, inst), N = inst_len(b, inst);
  const bool live = in_list && !(status && !(status[inst] & ST_ACTIVE));
  // the workgroup's first group holds its longest instance
  const uint32_t first = blockIdx.x * GROUPS_PER_WG;
  const uint32_t Nmax = inst_len(b, order ? order[first] : first);
  const int j = lane < NX ? lane : NX - 1;
  DirLds &L = lds[grp];
  if (lane < 16) {
    L.y[lane] = 0.f;
    L.x[lane] = 0.f;
  }
  float dinv[NX];  // column j of D_{b-1}^-1
#pragma unroll
  for (int r = 0; r < NX; ++r) dinv[r] = 0.f;
  KnotLoad cur, nxt;
  load_knot(cur, S, gamma, K0, true, lane, j);

  for (uint32_t k = 0; k < Nmax; ++k) {
    const bool on = live && k < N;
    const uint32_t kn = k + 1 < N ? k + 1 : N - 1;  // next knot (clamped)
    load_knot(nxt, S, gamma, (size_t)K0 + kn, kn == 0, lane, j);
#pragma unroll
    for (int t = 0; t < BPL; ++t) {
      const int i = lane + GROUP * t;
      if (i < NXX) L.B[i] = cur.b[t];
    }
    __syncthreads();
    // E_b (column j) = sum_m B(:, m) D_{b-1}^-1(m, j)
    float e[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) e[r] = 0.f;
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      const float d = dinv[m];
#pragma unroll
      for (int r = 0; r < NX; ++r) e[r] += L.B[r + NX * m] * d;
    }
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.E[r + NX * lane] = e[r];
      if (on && k > 0) {
        float *w = slot_E(work, K0 + k) + NX * lane;
#pragma unroll
        for (int r = 0; r < NX; ++r) w[r] = e[r];
      }
    }
    __syncthreads();
    // D_b (column j) = A_b(:, j) - sum_m E(:, m) B(j, m);  y_b(j) = r_j - sum_m E(j, m) y_{b-1}(m)
    float a[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) a[r] = cur.a[r];
    float yj = cur.r;
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      const float bjm = L.B[j + NX * m];
#pragma unroll
      for (int r = 0; r < NX; ++r) a[r] -= L.E[r + NX * m] * bjm;
      yj -= L.E[j + NX * m] * L.y[m];
    }
    pin(a);
    gj_inverse<NX>(lane, a, dinv);
    if (on && lane < NX) {
      float *w = slot_Dinv(work, K0 + k) + NX * lane;
#pragma unroll
      for (int r = 0; r < NX; ++r) w[r] = dinv[r];
    }
    __syncthreads();  // every read of y_{b-1}, B and E is done
    if (lane < NX) L.y[lane] = yj;
    __syncthreads();
    // z_b(j) = sum_m D_b^-1(j, m) y_b(m); D_b^-1 is symmetric, so row j is the lane's column
    float z = 0.f;
#pragma unroll
    for (int m = 0; m < NX; ++m) z += dinv[m] * L.y[m];
    if (on && lane < NX) lambda[(size_t)NX * (K0 + k) + lane] = z;
    cur = nxt;
  }

  // backward sweep: x_b = z_b - E_{b+1}^T x_{b+1}, z_b in lambda
  backward(L, work, lambda, nullptr, K0, N, Nmax, live, lane);

  for (uint32_t pass = 0; pass < refine; ++pass) {
    // forward: r_b = -gamma_b + (S lambda)_b (fp64), y_b = r_b - E_b y_{b-1}, z_b = D_b^-1 y_b
    if (lane < 16) L.y[lane] = 0.f;
    for (uint32_t k = 0; k < Nmax; ++k) {
      const bool on = live && k < N;
      const uint32_t kk = k < N ? k : N - 1;
      const size_t row = (size_t)K0 + kk;
      const float *Sr = S + (size_t)S_STRIDE * row;
      double res = -(double)gamma[(size_t)NX * row + j];
#pragma unroll
      for (int m = 0; m < NX; ++m) res += (double)Sr[NXX + j + NX * m] * (double)lambda[(size_t)NX * row + m];
      if (kk > 0) {
#pragma unroll
        for (int m = 0; m < NX; ++m) res += (double)Sr[j + NX * m] * (double)lambda[(size_t)NX * (row - 1) + m];
      }
      if (kk + 1 < N) {
#pragma unroll
        for (int m = 0; m < NX; ++m)
          res += (double)Sr[2 * NXX + j + NX * m] * (double)lambda[(size_t)NX * (row + 1) + m];
      }
      const float *Eg = slot_E(work, row), *Dg = slot_Dinv(work, row);
      for (int i = lane; i < NXX; i += GROUP) {
        L.E[i] = kk > 0 ? Eg[i] : 0.f;
        L.B[i] = Dg[i];
      }
      __syncthreads();
      float yj = (float)res;
#pragma unroll
      for (int m = 0; m < NX; ++m) yj -= L.E[j + NX * m] * L.y[m];
      __syncthreads();
      if (lane < NX) L.y[lane] = yj;
      __syncthreads();
      float z = 0.f;
#pragma unroll
      for (int m = 0; m < NX; ++m) z += L.B[m + NX * j] * L.y[m];
      if (on && lane < NX) slot_z(work, row)[lane] = z;
      __syncthreads();
    }
    backward(L, work, lambda, work, K0, N, Nmax, live, lane);
  }
  if (live && lane == 0) {
    if (iters) iters[inst] = 0;
    if (flag) flag[inst] = 0;
  }
}

}  // namespace

hipError_t launch_direct_solve(const Batch &b, const uint32_t *order, uint32_t n_list, const float *S,
                               const float *gamma, const uint32_t *status, float *work, float *lambda,
                               uint32_t refine, uint32_t *iters, uint32_t *flag, hipStream_t st) {
  if (n_list == 0) return hipSuccess;
  hipLaunchKernelGGL(k_direct_solve, dim3((n_list + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG), 0, st, b, order,
                     n_list, S, gamma, status, work, lambda, refine, iters, flag);
  return hipGetLastError();
}

}  // namespace mpc
