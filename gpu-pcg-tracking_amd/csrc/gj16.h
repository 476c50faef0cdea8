// gj16.h -- 14x14 block algebra inside a 16-lane group (gfx950): lane j
// holds column j of a block in registers; a pivot column is broadcast to
// the group by ds_swizzle, so a Gauss-Jordan inverse needs no LDS round
// trip and no barrier.  Used by the Schur formation (k_schur.hip) and the
// direct block-tridiagonal solve (k_direct.hip).
#pragma once
#include "common.h"

namespace mpc {

// value of lane P of the calling lane's 16-lane group (ds_swizzle bitmask
// mode: lane' = (lane & 0x10) | P within each 32-lane half)
template <int P>
__device__ __forceinline__ float bcast16(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x10 | (P << 5)));
}

// bcast16 with the lane p given as a value that is a constant after loop
// unrolling (ds_swizzle takes its pattern as an immediate)
__device__ __forceinline__ float bcast_dyn16(float v, int p) {
  switch (p) {
    case 0: return bcast16<0>(v);
    case 1: return bcast16<1>(v);
    case 2: return bcast16<2>(v);
    case 3: return bcast16<3>(v);
    case 4: return bcast16<4>(v);
    case 5: return bcast16<5>(v);
    case 6: return bcast16<6>(v);
    case 7: return bcast16<7>(v);
    case 8: return bcast16<8>(v);
    case 9: return bcast16<9>(v);
    case 10: return bcast16<10>(v);
    case 11: return bcast16<11>(v);
    case 12: return bcast16<12>(v);
    case 13: return bcast16<13>(v);
    case 14: return bcast16<14>(v);
    default: return bcast16<15>(v);
  }
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
    for (int r = p; r < n; ++r) pc[r] = bcast_dyn16(a[r], p);
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
