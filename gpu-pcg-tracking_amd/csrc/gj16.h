// gj16.h -- 14x14 block algebra inside a 16-lane group (gfx950): lane j
// holds column j of a block in registers; a pivot column is broadcast to
// the group by ds_swizzle (Gauss-Jordan) or a DPP row broadcast (the
// whitening factor), so neither needs an LDS buffer or a barrier.  Used by
// the Schur formation (k_schur.hip), the split PCG (k_pcg_split.hip) and
// the direct block-tridiagonal solve (k_direct.hip).
#pragma once
#include "common.h"

namespace mpc {

// value of lane P of the calling lane's 16-lane group (ds_swizzle bitmask
// mode: lane' = (lane & 0x10) | P within each 32-lane half)
template <int P>
__device__ __forceinline__ float bcast16(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x10 | (P << 5)));
}

// The same broadcast by DPP row_newbcast:P (dpp_ctrl 0x150 + P, gfx90a and
// later): lane P of each 16-lane row to the whole row, one VALU instruction
// with no trip through the LDS pipeline.  Lane P must be active.  Used on
// the dependent pivot chain of chol_inv_rows, where the swizzle's LDS
// latency sits on the critical path of every step.
template <int P>
__device__ __forceinline__ float bcast16_dpp(float v) {
  static_assert(P >= 0 && P < 16, "row lane");
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + P, 0xf, 0xf, true));
}

// bcast16 (Dpp: bcast16_dpp) with the lane p given as a value that is a
// constant after loop unrolling (both take the lane as an immediate)
template <bool Dpp = false>
__device__ __forceinline__ float bcast_dyn16(float v, int p) {
#define MPC_BCASE(P) \
  case P: return Dpp ? bcast16_dpp<P>(v) : bcast16<P>(v);
  switch (p) {
    MPC_BCASE(0) MPC_BCASE(1) MPC_BCASE(2) MPC_BCASE(3) MPC_BCASE(4) MPC_BCASE(5) MPC_BCASE(6) MPC_BCASE(7)
    MPC_BCASE(8) MPC_BCASE(9) MPC_BCASE(10) MPC_BCASE(11) MPC_BCASE(12) MPC_BCASE(13) MPC_BCASE(14)
    default: return Dpp ? bcast16_dpp<15>(v) : bcast16<15>(v);
  }
#undef MPC_BCASE
}

// Materialise values at this point of the program: keeps the compiler from
// sinking their computation to a later (conditional) use, which would keep
// all of its inputs alive in registers.
template <int n>
__device__ __forceinline__ void pin(float (&v)[n]) {
#pragma unroll
  for (int r = 0; r < n; ++r) asm volatile("" : "+v"(v[r]));
}

template <int n, int p>
__device__ __forceinline__ void gj_step(float (&a)[NX], float (&x)[NX]) {
  float pc[n];
#pragma unroll
  for (int r = 0; r < n; ++r) pc[r] = bcast16<p>(a[r]);
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
  // the x updates would otherwise sink to their use after the last step and
  // keep every step's pivot column alive
  pin(a);
  pin(x);
}

template <int n, int p = 0>
__device__ __forceinline__ void gj_steps(float (&a)[NX], float (&x)[NX]) {
  if constexpr (p < n) {
    gj_step<n, p>(a, x);
    gj_steps<n, p + 1>(a, x);
  }
}

// Gauss-Jordan inverse of an n x n matrix (n <= 14) held by columns: lane
// j < n holds column j of M in a[0..n) (lanes >= n hold a copy of column
// n-1, which no lane reads); returns column j of M^-1 in x.  Every lane of
// the group must be active.
template <int n>
__device__ __forceinline__ void gj_inverse(int lane, float (&a)[NX], float (&x)[NX]) {
  const int j = lane < n ? lane : n - 1;
#pragma unroll
  for (int r = 0; r < n; ++r) x[r] = (r == j) ? 1.f : 0.f;
  gj_steps<n>(a, x);
}

// Whitening factor of a symmetric positive definite n x n matrix M (n <=
// 14) held by columns (lane j < n: column j in a; lanes >= n a copy of
// column n-1): Gauss elimination of [M | I] without pivoting, each pivot row
// scaled by the pivot's inverse square root once final.  Afterwards lane j
// holds row j of W = U^-1 in x, where M = U^T U (Cholesky, U upper
// triangular; the scaled rows of the left half are the rows of U), so
// W^T M W = I; a is destroyed.  Every lane of the group must be active.
// The pivot's reciprocal and inverse square root are the hardware's
// (v_rcp_f32, v_rsq_f32: 1 ulp), not the correctly rounded sequences: W is
// this build's own factor (no reference operation), and a one-ulp change
// of a multiplier is a rounding-level change of M.
template <int n>
__device__ __forceinline__ void chol_inv_rows(int lane, float (&a)[NX], float (&x)[NX]) {
  const int j = lane < n ? lane : n - 1;
#pragma unroll
  for (int r = 0; r < NX; ++r) x[r] = (r == j) ? 1.f : 0.f;
#pragma unroll
  for (int p = 0; p < n; ++p) {
    float pc[NX];
#pragma unroll
    for (int r = p; r < n; ++r) pc[r] = bcast_dyn16<true>(a[r], p);
    const float inv = __builtin_amdgcn_rcpf(pc[p]);
    const float s = __builtin_amdgcn_rsqf(pc[p]);
    const float ap = a[p], xp = x[p];
#pragma unroll
    for (int r = p + 1; r < n; ++r) {
      const float m = pc[r] * inv;
      a[r] -= m * ap;
      x[r] -= m * xp;
    }
    a[p] = ap * s;
    x[p] = xp * s;
    pin(a);
    pin(x);
  }
}

}  // namespace mpc
