// k_schur.hip -- Schur complement, block-Jacobi preconditioner and dz
// recovery for gfx950.
//
// Two kernels, one 16-lane group per knot / block row, lane j owning
// column j of every 14x14 block (registers), 4 groups per 64-lane workgroup:
//   k_invert_g:   G~^-1 of every knot (Q~_k^-1, R~_k^-1) into the G^-1 buffer;
//   k_form_schur: per block row, the products, theta, its inverse and the
//                 outputs, reading Q~^-1 / R~^-1 from G^-1.
// Block inverses are Gauss-Jordan eliminations on [M | I]; the pivot column
// is broadcast within the group by ds_swizzle (no LDS round trip, no
// barrier).  The regularised blocks are symmetric positive definite, so no
// pivoting is needed.
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
#include <cstddef>
#include <utility>

#include "common.h"
#include "gj16.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int GROUP = 16;
constexpr int GROUPS_PER_WG = 4;
constexpr int WG = GROUP * GROUPS_PER_WG;

// G~_k^-1 = [ (Q_k + rho I)^-1 | (R_k + rho I)^-1 ] of every knot (the
// last knot of an instance has no R block)
__global__ void __launch_bounds__(WG) k_invert_g(Batch b, const float *__restrict__ G,
                                                 const float *__restrict__ rho_dev, float rho_s,
                                                 const uint32_t *__restrict__ status, float *__restrict__ Ginv) {
  const int lane = threadIdx.x % GROUP;
  const uint32_t gk_raw = blockIdx.x * GROUPS_PER_WG + threadIdx.x / GROUP;
  bool valid = gk_raw < b.total_knots;
  const uint32_t gk = valid ? gk_raw : b.total_knots - 1;
  const KnotRef kr = locate(b, gk);
  if (status && !(status[kr.inst] & ST_ACTIVE)) valid = false;
  const float rho = rho_dev ? rho_dev[kr.inst] : rho_s;
  const float *Gk = G + (size_t)G_STRIDE * gk;
  float *Gi = Ginv + (size_t)G_STRIDE * gk;
  float a[NX], x[NX];
  {
    const int j = lane < NX ? lane : NX - 1;
#pragma unroll
    for (int r = 0; r < NX; ++r) a[r] = Gk[r + NX * j] + ((r == j) ? rho : 0.f);
    gj_inverse<NX>(lane, a, x);
    if (valid && lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) Gi[r + NX * lane] = x[r];
    }
  }
  if (kr.k + 1 < kr.N) {  // uniform over the group
    const int j = lane < NU ? lane : NU - 1;
#pragma unroll
    for (int r = 0; r < NU; ++r) a[r] = Gk[NXX + r + NU * j] + ((r == j) ? rho : 0.f);
    gj_inverse<NU>(lane, a, x);
    if (valid && lane < NU) {
#pragma unroll
      for (int r = 0; r < NU; ++r) Gi[NXX + r + NU * lane] = x[r];
    }
  }
}

// 16-byte aligned blocks: AQ and A (reused for the S(b,b) and Pinv(b,b)
// blocks on their way out) are read as float4 rows for coalesced stores
struct alignas(16) SchurLds {
  float AQ[NXX];  // A Q~_k^-1
  float A[NXX];   // A_k (column-major)
  float B[98];    // B_k
  float BR[98];   // B R~_k^-1
  float qk[16], q1[16], rk[8], cb[16];
};
static_assert(NXX % 4 == 0 && sizeof(float) * NXX % 16 == 0, "float4 rows of a 14 x 14 block");

// a 14 x 14 block (element i at src[f(i)]) from the group's LDS to global
// memory as 49 float4 over the 16 lanes; dst is 16-byte aligned whenever
// the S / Pinv base is (rows are 588 floats apart), else plain stores
template <typename F>
__device__ __forceinline__ void store_block(int lane, float *dst, F &&f) {
  if ((reinterpret_cast<uintptr_t>(dst) & 15u) == 0) {
    float4 *d = reinterpret_cast<float4 *>(dst);
    for (int v = lane; v < NXX / 4; v += GROUP) d[v] = make_float4(f(4 * v), f(4 * v + 1), f(4 * v + 2), f(4 * v + 3));
  } else {
    for (int i = lane; i < NXX; i += GROUP) dst[i] = f(i);
  }
}

__global__ void __launch_bounds__(WG) k_form_schur(Batch b, const float *__restrict__ G,
                                                   const float *__restrict__ C,
                                                   const float *__restrict__ g,
                                                   const float *__restrict__ cvec,
                                                   const float *__restrict__ rho_dev, float rho_s,
                                                   const uint32_t *__restrict__ status,
                                                   float *__restrict__ S, float *__restrict__ Pinv,
                                                   float *__restrict__ gamma, const float *__restrict__ Ginv,
                                                   int out_mode) {
  const bool zero_unused = out_mode & SCHUR_ZERO_UNUSED, right = out_mode & SCHUR_RIGHT_BLOCK;
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
  // Every group runs the same sequence; block row 0 stages its own knot in
  // the 'previous knot' slots and only keeps Q~_0^-1.
  const bool row0 = bk == 0;
  const size_t kprev = row0 ? row : row - 1;  // block row b >= 1: k = b - 1
  const float *Ck = C + (size_t)C_STRIDE * kprev;
  const float *Qi = Ginv + (size_t)G_STRIDE * kprev;  // Q~_k^-1
  const float *Ri = Qi + NXX;                         // R~_k^-1
  const float *Q1i = Ginv + (size_t)G_STRIDE * row;   // Q~_b^-1
  for (int i = lane; i < NXX; i += GROUP) L.A[i] = -Ck[i];
  for (int i = lane; i < 98; i += GROUP) L.B[i] = -Ck[NXX + i];
  if (lane < NX) {
    L.qk[lane] = g[(size_t)NXU * kprev + lane];
    L.q1[lane] = g[(size_t)NXU * row + lane];
    L.cb[lane] = (row0 || !cvec) ? 0.f : cvec[(size_t)NX * row + lane];  // no c: c = 0
  }
  if (lane < NU) L.rk[lane] = g[(size_t)NXU * kprev + NX + lane];
  __syncthreads();

  // AQ = A Q~_k^-1 (column j), BR = B R~_k^-1 (column j < 7)
  const int j = lane < NX ? lane : NX - 1;
  {
    float qi[NX];
#pragma unroll
    for (int l = 0; l < NX; ++l) qi[l] = Qi[l + NX * j];
    float acc[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) acc[r] = 0.f;
#pragma unroll
    for (int l = 0; l < NX; ++l)
#pragma unroll
      for (int r = 0; r < NX; ++r) acc[r] += L.A[r + NX * l] * qi[l];
    pin(acc);
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.AQ[r + NX * lane] = acc[r];
    }
  }
  {
    const int ju = lane < NU ? lane : NU - 1;
    float ri[NU];
#pragma unroll
    for (int l = 0; l < NU; ++l) ri[l] = row0 ? 0.f : Ri[l + NU * ju];
    float bacc[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) bacc[r] = 0.f;
#pragma unroll
    for (int l = 0; l < NU; ++l)
#pragma unroll
      for (int r = 0; r < NX; ++r) bacc[r] += L.B[r + NX * l] * ri[l];
    pin(bacc);
    if (lane < NU) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.BR[r + NX * lane] = bacc[r];
    }
  }
  __syncthreads();
  // theta(:, j) = sum_l AQ(:, l) A(j, l) + sum_l BR(:, l) B(j, l) + Q~_b^-1(:, j)
  float th[NX];
#pragma unroll
  for (int r = 0; r < NX; ++r) th[r] = Q1i[r + NX * j];
#pragma unroll
  for (int l = 0; l < NX; ++l) {
    const float ajl = L.A[j + NX * l];
#pragma unroll
    for (int r = 0; r < NX; ++r) th[r] += L.AQ[r + NX * l] * ajl;
  }
#pragma unroll
  for (int l = 0; l < NU; ++l) {
    const float bjl = L.B[j + NX * l];
#pragma unroll
    for (int r = 0; r < NX; ++r) th[r] += L.BR[r + NX * l] * bjl;
  }
  pin(th);
  __syncthreads();  // every lane is done reading A
  // S(b,b) = -theta, transposed into rows through A for coalesced stores
  if (lane < NX) {
#pragma unroll
    for (int r = 0; r < NX; ++r) L.A[r + NX * lane] = -th[r];
  }
  __syncthreads();
  if (valid && row0) {
    // S(0,0) = -Q~_0^-1, gamma_0 = -Q~_0^-1 q_0, Pinv(0,0) = -Q~_0
    const float *Q0 = G + (size_t)G_STRIDE * row;
    for (int i = lane; i < NXX; i += GROUP) {
      Srow[NXX + i] = -Qi[i];
      Prow[NXX + i] = -(Q0[i] + ((i % NX == i / NX) ? rho : 0.f));
    }
    if (lane < NX) {
      float acc = 0.f;
#pragma unroll
      for (int l = 0; l < NX; ++l) acc += Qi[lane + NX * l] * L.qk[l];
      gamma[(size_t)NX * row + lane] = -acc;
    }
    if (zero_unused)
      for (int i = lane; i < NXX; i += GROUP) {
        Srow[i] = 0.f;
        Prow[i] = 0.f;
        Prow[2 * NXX + i] = 0.f;
      }
  }
  if (valid && !row0) {
    store_block(lane, Srow, [&](int i) { return L.AQ[i]; });
    store_block(lane, Srow + NXX, [&](int i) { return L.A[i]; });
    // right block of the previous row: S(b-1, b) = AQ^T (the reference layout;
    // the solvers read it as the transpose of the next row's left block)
    if (right)
      store_block(lane, S + (size_t)S_STRIDE * kprev + 2 * NXX, [&](int i) { return L.AQ[i / NX + NX * (i % NX)]; });
    if (lane < NX) {
      float a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
      for (int l = 0; l < NX; ++l) {
        a1 += Q1i[lane + NX * l] * L.q1[l];
        a2 += L.AQ[lane + NX * l] * L.qk[l];
      }
#pragma unroll
      for (int l = 0; l < NU; ++l) a3 += L.BR[lane + NX * l] * L.rk[l];
      gamma[(size_t)NX * row + lane] = L.cb[lane] - (a1 - a2 - a3);
    }
    if (zero_unused) {
      if (bk + 1 == kr.N && right)
        for (int i = lane; i < NXX; i += GROUP) Srow[2 * NXX + i] = 0.f;
      for (int i = lane; i < NXX; i += GROUP) {
        Prow[i] = 0.f;
        Prow[2 * NXX + i] = 0.f;
      }
    }
  }
  // Pinv(b,b) = -theta^-1
  float x[NX];
  gj_inverse<NX>(lane, th, x);
  __syncthreads();  // every lane is done reading S(b,b) from A
  if (lane < NX) {
#pragma unroll
    for (int r = 0; r < NX; ++r) L.A[r + NX * lane] = -x[r];
  }
  __syncthreads();
  if (valid && !row0) store_block(lane, Prow + NXX, [&](int i) { return L.A[i]; });
}

// ---- the split system for the split-form PCG (k_pcg_split.hip), formed
// with the block rows: per block row b, theta_b as in k_form_schur, W_b =
// chol(theta_b)^-1 (chol_inv_rows), gamma^_b = W_b^T gamma_b and L^_b =
// W_b^T S(b,b-1) W_{b-1}, S(b,b-1) = A Q~_{b-1}^-1.  S, gamma, Pinv and
// G~^-1 are not written.  The inverses are taken in closed form from the
// structure of the cost Hessian (k_form_kkt; iiwa_eepos_plant.cuh:357-375):
//   Q~ = D + v v^T, D = diag(wx + rho), v = (g_q, 0):
//   Q~^-1 = D^-1 - u u^T / s, u = D^-1 v, s = 1 + v^T D^-1 v (Sherman-Morrison),
//   R~^-1 = diag(1 / (wu + rho)),
// so A Q~^-1 = A D^-1 - (A u) u^T / s.  Along v the two terms are of order
// 1 / rho and cancel to 1 / (rho + |v|^2), so D^-1 is taken in fp32 once
// (v_rcp_f32, 1 ulp) and u, s, A u and the cancelling sums in fp64
// (products of fp32 values are exact there): the result is, to fp32
// rounding of its own entries, the inverse of Q~ with its diagonal moved by
// a few ulps.
// The block products (theta, S(b,b-1) W_{b-1} and W_b^T of that) take
// their left factor's columns from the group's lanes by DPP broadcasts
// fused into the FMAs (nb_product), so the factors stay in registers.
// Workgroups of SG groups overlap by one row: group g of workgroup w forms
// row (SG - 1) w + g - 1, and groups g >= 1 take W_{b-1} from group g - 1
// through LDS and write their row (each row from exactly one group).  That
// hand-over is the one workgroup barrier; every other LDS exchange stays
// inside a group, so inside one wave (wave_sync).
constexpr int SG = 16;
constexpr int SWG = GROUP * SG;
constexpr uint32_t SSTEP = SG - 1;

struct alignas(16) SplitRowLds {
  float AQ[NXX];   // A Q~_{b-1}^-1 = S(b, b-1)
  float A[NXX];    // A_{b-1} (column-major); then W_b
  float B[98];     // B_{b-1}; then, with BR, the L^ block on its way out
  float BR[98];    // B R~_{b-1}^-1
  float gm[16];    // gamma_b
};  // 2416 B: four 16-group workgroups fit one CU's LDS
static_assert(offsetof(SplitRowLds, BR) == offsetof(SplitRowLds, B) + 98 * sizeof(float), "B, BR adjacent");

// u = D^-1 v and 1 / s of Q~ = D + v v^T for the g_q of one knot, in fp64
// (every lane the same values, in the same order)
__device__ __forceinline__ void q_lowrank(const float *__restrict__ vq, const float (&dinv)[NQ], bool zero,
                                          double (&u)[NQ], double &inv_s) {
  double s = 1.0;
#pragma unroll
  for (int l = 0; l < NQ; ++l) {
    const double v = zero ? 0.0 : (double)vq[l];
    u[l] = v * (double)dinv[l];
    s = fma(u[l], v, s);
  }
  inv_s = 1.0 / s;
}

__device__ __forceinline__ double bcast_dyn16_f64(double v, int p) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const float lo = bcast_dyn16<true>(__builtin_bit_cast(float, (uint32_t)b), p);
  const float hi = bcast_dyn16<true>(__builtin_bit_cast(float, (uint32_t)(b >> 32)), p);
  return __builtin_bit_cast(double, ((unsigned long long)__builtin_bit_cast(uint32_t, hi) << 32) |
                                        __builtin_bit_cast(uint32_t, lo));
}

// d[r] += s_l[r] * w (r < 14), s_l the s of lane l of the 16-lane group:
// v_fmac_f32_dpp with DPP row_newbcast:l, so a block product whose left
// factor is held by columns (lane l: column l) needs no LDS and no
// separate broadcast instruction.  The s_nop ahead of each seven covers the
// wait states between a VALU write of an operand (or of EXEC, Nop4) and its
// DPP read.  The group's lanes must all be active.
#define MPC_NBF(k) "v_fmac_f32_dpp %" #k ", %" MPC_NBS##k ", %14 row_newbcast:%15 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
#define MPC_NBS0 "7"
#define MPC_NBS1 "8"
#define MPC_NBS2 "9"
#define MPC_NBS3 "10"
#define MPC_NBS4 "11"
#define MPC_NBS5 "12"
#define MPC_NBS6 "13"
#define MPC_NB7(nop)                                                                                          \
  asm(nop MPC_NBF(0) MPC_NBF(1) MPC_NBF(2) MPC_NBF(3) MPC_NBF(4) MPC_NBF(5) MPC_NBF(6)                          \
      : "+v"(d[o]), "+v"(d[o + 1]), "+v"(d[o + 2]), "+v"(d[o + 3]), "+v"(d[o + 4]), "+v"(d[o + 5]), "+v"(d[o + 6]) \
      : "v"(s[o]), "v"(s[o + 1]), "v"(s[o + 2]), "v"(s[o + 3]), "v"(s[o + 4]), "v"(s[o + 5]), "v"(s[o + 6]),       \
        "v"(w), "i"(l))
template <int l, bool Nop4, int o>
__device__ __forceinline__ void nb_fma7(float (&d)[NX], const float (&s)[NX], float w) {
  static_assert(l >= 0 && l < 16 && o + 7 <= NX, "row lane, accumulator range");
  if constexpr (Nop4)
    MPC_NB7("s_nop 4\n");
  else
    MPC_NB7("s_nop 1\n");
}
#undef MPC_NB7
#undef MPC_NBF
#undef MPC_NBS0
#undef MPC_NBS1
#undef MPC_NBS2
#undef MPC_NBS3
#undef MPC_NBS4
#undef MPC_NBS5
#undef MPC_NBS6

// d[r] += sum_{l in Ls, in order} s_l[r] * w[l]: a block product from the
// left factor's columns (one per lane) and the calling lane's row of the
// right factor's transpose
template <bool Nop4, int... Ls>
__device__ __forceinline__ void nb_product(float (&d)[NX], const float (&s)[NX], const float *w,
                                           std::integer_sequence<int, Ls...>) {
  ((nb_fma7<Ls, Nop4 && Ls == 0, 0>(d, s, w[Ls]), nb_fma7<Ls, false, 7>(d, s, w[Ls])), ...);
}

__global__ void __launch_bounds__(SWG) __attribute__((amdgpu_waves_per_eu(4))) k_form_split(Batch b, Params p,
                                                    const float *__restrict__ Cl,
                                                    const float *__restrict__ g, const float *__restrict__ cvec,
                                                    const float *__restrict__ gq,
                                                    const float *__restrict__ rho_dev, float rho_s,
                                                    const uint32_t *__restrict__ status, float *__restrict__ Ps,
                                                    float *__restrict__ ghat, float *__restrict__ rho_used) {
  __shared__ SplitRowLds lds[SG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const int64_t row_raw = (int64_t)blockIdx.x * SSTEP + grp - 1;
  const bool in_range = row_raw >= 0 && row_raw < (int64_t)b.total_knots;
  const uint32_t row = in_range ? (uint32_t)row_raw : (row_raw < 0 ? 0u : b.total_knots - 1);
  // The reads of the block row are issued before the instance lookup: the
  // previous knot's C, g and g_q (for a first knot another instance's last
  // knot, whose C and r are not written: replaced by zeros below, never
  // multiplied) and this knot's g and g_q.
  const size_t kprev = row > 0 ? row - 1 : 0u;
  // the previous knot's lower block rows (compact form), lane l its column
  // l: rows NQ..NX-1 of A (l < 14) and of B (l < 7); the upper rows are the
  // Euler step's (cl_upper_a, and zero in B)
  const float *Clk = Cl + (size_t)CL_STRIDE * kprev;
  const int la = lane < NX ? lane : NX - 1, lb = lane < NU ? lane : NU - 1;
  float ca[NQ], cb[NQ];
#pragma unroll
  for (int r = 0; r < NQ; ++r) {
    ca[r] = Clk[r + NQ * la];
    cb[r] = Clk[CL_A + r + NQ * lb];
  }
  const KnotEnds ke = knot_ends(b, row);
  const uint32_t inst = ke.inst;
  const bool row0 = ke.first;
  const bool owner = grp >= 1 && in_range && !(status && !(status[inst] & ST_ACTIVE));
  SplitRowLds &L = lds[grp];
  // A = -C_A and B = -C_B staged column-major, as negations of the values
  // launch_form_kkt stores (zeros for a first knot)
  if (lane < NX) {
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      L.A[r + NX * lane] = row0 ? 0.f : -cl_upper_a(r, lane, p.dt);
      L.A[NQ + r + NX * lane] = row0 ? 0.f : -ca[r];
    }
  }
  if (lane < NU) {
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      L.B[r + NX * lane] = row0 ? 0.f : -0.f;
      L.B[NQ + r + NX * lane] = row0 ? 0.f : -cb[r];
    }
  }
  const int j = lane < NX ? lane : NX - 1;
  // q_{b-1}, r_{b-1}, q_b: read by every lane of the group (cached)
  const float *qk = g + (size_t)NXU * kprev, *rk = qk + NX, *q1v = g + (size_t)NXU * row;
  const float rho = rho_dev ? rho_dev[inst] : rho_s;
  // D^-1 in fp32 (the same for both knots: one instance, one rho); u and 1/s
  // of Q~_{b-1} (zero for row 0) and of Q~_b in fp64
  float dq[NQ];
#pragma unroll
  for (int l = 0; l < NQ; ++l) dq[l] = __builtin_amdgcn_rcpf(p.wx[l] + rho);
  const float dj = __builtin_amdgcn_rcpf(p.wx[j] + rho);  // D^-1(j, j)
  const float dpj = row0 ? 0.f : dj;       // of Q~_{b-1}^-1 (zero for row 0)
  wave_sync();  // the group's staged A, B (a group is inside one wave)
  // AQ = A Q~_{b-1}^-1 (column j) = A(:, j) D^-1(j, j) - (A u)(u_j / s), the
  // vector A u formed one element per lane and broadcast; BR = B R~^-1
  // (column j < 7)
  float acc[NX];  // column j of AQ (lanes 14, 15: column 13)
  {
    double up[NQ], isp;
    q_lowrank(gq + (size_t)NQ * kprev, dq, row0, up, isp);
    double au = 0.0;
#pragma unroll
    for (int l = 0; l < NQ; ++l) au = fma((double)L.A[j + NX * l], up[l], au);
    const double upj = j < NQ && !row0 ? (double)gq[(size_t)NQ * kprev + j] * (double)dj : 0.0;  // u_{b-1}(j)
    const double cj = upj * isp;
#pragma unroll
    for (int r = 0; r < NX; ++r)
      acc[r] = (float)fma(-bcast_dyn16_f64(au, r), cj, (double)L.A[r + NX * j] * (double)dpj);
    pin(acc);
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.AQ[r + NX * lane] = acc[r];
    }
  }
  float brc[NX];  // column ju of BR
  {
    const int ju = lane < NU ? lane : NU - 1;
    const float wr = row0 ? 0.f : __builtin_amdgcn_rcpf(p.wu[ju] + rho);
#pragma unroll
    for (int r = 0; r < NX; ++r) brc[r] = L.B[r + NX * ju] * wr;
    if (lane < NU) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.BR[r + NX * lane] = brc[r];
    }
  }
  // row j of A and of B (the right factors of theta, transposed)
  float arow[NX], brow[NU];
#pragma unroll
  for (int l = 0; l < NX; ++l) arow[l] = L.A[j + NX * l];
#pragma unroll
  for (int l = 0; l < NU; ++l) brow[l] = L.B[j + NX * l];
  wave_sync();  // the group's AQ and BR (for gamma)
  // theta(:, j) = sum_l AQ(:, l) A(j, l) + sum_l BR(:, l) B(j, l) + Q~_b^-1(:, j)
  float th[NX];
  double isb;
  const double ubj = j < NQ ? (double)gq[(size_t)NQ * row + j] * (double)dj : 0.0;  // u_b(j)
  {
    double ub[NQ];
    q_lowrank(gq + (size_t)NQ * row, dq, false, ub, isb);
    const double cb = ubj * isb;
#pragma unroll
    for (int r = 0; r < NX; ++r) th[r] = (float)((r == j ? (double)dj : 0.0) - (r < NQ ? ub[r] * cb : 0.0));
  }
  // the columns of AQ and BR broadcast from their lanes (DPP), in the
  // order of the sums above
  nb_product<true>(th, acc, arow, std::make_integer_sequence<int, NX>{});
  nb_product<true>(th, brc, brow, std::make_integer_sequence<int, NU>{});
  pin(th);
  // gamma = c - (Q~_b^-1 q_b - A Q~^-1 q_{b-1} - B R~^-1 r_{b-1}) (row 0: -Q~_0^-1 q_0)
  if (lane < NX) {
    float a2 = 0.f, a3 = 0.f;
    double uq = 0.0;
#pragma unroll
    for (int l = 0; l < NQ; ++l) uq = fma((double)gq[(size_t)NQ * row + l] * (double)dq[l], (double)q1v[l], uq);
    const float a1 = (float)fma(-ubj, uq * isb, (double)q1v[lane] * (double)dj);
#pragma unroll
    for (int l = 0; l < NX; ++l) a2 += L.AQ[lane + NX * l] * qk[l];
#pragma unroll
    for (int l = 0; l < NU; ++l) a3 += L.BR[lane + NX * l] * (row0 ? 0.f : rk[l]);
    const float cb = (row0 || !cvec) ? 0.f : cvec[(size_t)NX * row + lane];  // gamma_0 has no c term
    L.gm[lane] = cb - (a1 - a2 - a3);
  }
  float x[NX];
  chol_inv_rows<NX>(lane, th, x);  // row j of W_b
  wave_sync();                     // the group is done reading its A
  if (lane < NX) {
#pragma unroll
    for (int c = 0; c < NX; ++c) L.A[lane + NX * c] = x[c];  // W_b, column-major
  }
  __syncthreads();  // W_b to the next group (the one barrier across waves)
  // gamma^ = W_b^T gamma_b; L^ = W_b^T (S(b,b-1) W_{b-1}) by columns
  float *Pr = Ps + (size_t)S_STRIDE * row;
  float *Lout = L.B;  // B and BR: one 196-float block
  if (owner) {
    if (lane < NX) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < NX; ++k) acc += L.A[k + NX * lane] * L.gm[k];
      ghat[(size_t)NX * row + lane] = acc;
    }
    if (row0 && lane == 0 && rho_used) rho_used[inst] = rho;
    if (!row0) {
      // m = AQ W_{b-1}(:, j) from the AQ columns (acc) of the group's
      // lanes, then L^(:, j) = W_b^T m from the rows of W_b (x), both by
      // DPP broadcasts (the whole group is active here: owner and row0 are
      // the same for all its lanes)
      const float *Wp = lds[grp - 1].A;  // W_{b-1}, column-major
      float wp[NX], m[NX], lh[NX];
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        wp[k] = Wp[k + NX * j];
        m[k] = 0.f;
        lh[k] = 0.f;
      }
      nb_product<true>(m, acc, wp, std::make_integer_sequence<int, NX>{});
      nb_product<true>(lh, x, m, std::make_integer_sequence<int, NX>{});
      if (lane < NX) {
#pragma unroll
        for (int r = 0; r < NX; ++r) Lout[r + NX * lane] = lh[r];
      }
    }
  }
  wave_sync();  // the group's L^ and W_b, stored below
  if (owner) {
    store_block(lane, Pr + NXX, [&](int i) { return L.A[i]; });
    if (!row0) store_block(lane, Pr, [&](int i) { return Lout[i]; });
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
  const KnotEnds ke = knot_ends(b, gk);
  if (status && !(status[ke.inst] & ST_ACTIVE)) valid = false;
  DzLds &L = lds[grp];
  const size_t knot = gk;
  const bool has_next = ke.has_next;
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
  float *dzi = dz + (size_t)NXU * gk - (size_t)NU * ke.inst;
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

// dz of k_compute_dz with Q~_k^-1 and R~_k^-1 in the closed form of
// k_form_split (no G~^-1)
__global__ void __launch_bounds__(WG) k_compute_dz_lr(Batch b, Params p, const float *__restrict__ gq,
                                                      const float *__restrict__ rho_dev,
                                                      const float *__restrict__ Cl,
                                                      const float *__restrict__ g,
                                                      const float *__restrict__ lambda,
                                                      const uint32_t *__restrict__ status,
                                                      float *__restrict__ dz) {
  __shared__ DzLds lds[GROUPS_PER_WG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t gk_raw = blockIdx.x * GROUPS_PER_WG + grp;
  bool valid = gk_raw < b.total_knots;
  const uint32_t gk = valid ? gk_raw : b.total_knots - 1;
  const size_t knot = gk;
  // every read of the knot is issued at entry, before the instance lookup:
  // lambda of the knot and of the next one (a clamped address, selected away
  // at an instance's last knot), the lower block rows of C (compact form:
  // that knot's are never written, so their sums are selected away, not
  // multiplied by zero), g and g_q
  const int la = lane < NX ? lane : NX - 1, lu = lane < NU ? lane : NU - 1;
  const size_t nk = knot + 1 < b.total_knots ? knot + 1 : knot;
  const float lam_next = lambda[(size_t)NX * nk + la], lam_k = lambda[(size_t)NX * knot + la];
  const float *Clk = Cl + (size_t)CL_STRIDE * knot;
  float ca[NQ], cb[NQ], gqv[NQ];
#pragma unroll
  for (int r = 0; r < NQ; ++r) {
    ca[r] = Clk[r + NQ * la];
    cb[r] = Clk[CL_A + r + NQ * lu];
    gqv[r] = gq[(size_t)NQ * knot + r];
  }
  const float gx = g[(size_t)NXU * knot + la], gu = g[(size_t)NXU * knot + NX + lu];
  const float gql = gq[(size_t)NQ * knot + (lane < NQ ? lane : 0)];
  const KnotEnds ke = knot_ends(b, gk);
  if (status && !(status[ke.inst] & ST_ACTIVE)) valid = false;
  const float rho = rho_dev[ke.inst];
  DzLds &L = lds[grp];
  const bool has_next = ke.has_next;
  if (lane < NX) L.lam1[lane] = has_next ? lam_next : 0.f;
  __syncthreads();
  if (lane < NX) {
    // t_j = q_j - lambda_k,j + (A^T lambda_{k+1})_j, A = -C_A; the upper
    // rows of C are the Euler step's (cl_upper_a)
    float s = 0.f;
#pragma unroll
    for (int l = 0; l < NX; ++l) s -= (l < NQ ? cl_upper_a(l, lane, p.dt) : ca[l - NQ]) * L.lam1[l];
    s = has_next ? s : 0.f;
    L.t[lane] = gx - lam_k + s;
  }
  if (has_next && lane < NU) {
    float s = 0.f;
#pragma unroll
    for (int l = 0; l < NX; ++l) s -= (l < NQ ? 0.f : cb[l - NQ]) * L.lam1[l];
    L.tu[lane] = gu + s;
  }
  __syncthreads();
  if (!valid) return;
  float *dzi = dz + (size_t)NXU * gk - (size_t)NU * ke.inst;
  if (lane < NX) {
    // Q~^-1 t = D^-1 t - u (u^T t) / s, the cancelling terms in fp64 (as
    // k_form_split)
    float dq[NQ];
    double u[NQ], inv_s, ut = 0.0;
#pragma unroll
    for (int l = 0; l < NQ; ++l) dq[l] = __builtin_amdgcn_rcpf(p.wx[l] + rho);
    q_lowrank(gqv, dq, false, u, inv_s);
#pragma unroll
    for (int l = 0; l < NQ; ++l) ut = fma(u[l], (double)L.t[l], ut);
    const float dl = __builtin_amdgcn_rcpf(p.wx[lane] + rho);
    const double ul = lane < NQ ? (double)gql * (double)dl : 0.0;
    dzi[lane] = (float)fma(-ul, ut * inv_s, (double)L.t[lane] * (double)dl);
  }
  if (has_next && lane < NU) dzi[NX + lane] = L.tu[lane] * __builtin_amdgcn_rcpf(p.wu[lane] + rho);
}

}  // namespace

hipError_t launch_form_schur(const Batch &b, const float *G, const float *C, const float *g,
                             const float *c, const float *rho_dev, float rho, const uint32_t *status,
                             float *S, float *Pinv, float *gamma, float *Ginv, int out_mode,
                             hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  const dim3 grid((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG);
  hipLaunchKernelGGL(k_invert_g, grid, dim3(WG), 0, st, b, G, rho_dev, rho, status, Ginv);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_form_schur, grid, dim3(WG), 0, st, b, G, C, g, c, rho_dev, rho, status, S, Pinv, gamma,
                     (const float *)Ginv, out_mode);
  return hipGetLastError();
}

hipError_t launch_form_split(const Batch &b, const Params &p, const float *Cl, const float *g, const float *c,
                             const float *gq, const float *rho_dev, float rho, const uint32_t *status, float *Ps,
                             float *ghat, float *rho_used, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  if (!gq || !Cl) return hipErrorInvalidValue;  // g_q and the compact rows of launch_form_kkt
  hipLaunchKernelGGL(k_form_split, dim3((b.total_knots + SSTEP - 1) / SSTEP), dim3(SWG), 0, st, b, p, Cl, g, c, gq,
                     rho_dev, rho, status, Ps, ghat, rho_used);
  return hipGetLastError();
}

hipError_t launch_compute_dz_lr(const Batch &b, const Params &p, const float *gq, const float *rho_dev,
                                const float *Cl, const float *g, const float *lambda, const uint32_t *status,
                                float *dz, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  if (!rho_dev || !gq || !Cl) return hipErrorInvalidValue;  // per-instance rho, g_q and the compact rows
  hipLaunchKernelGGL(k_compute_dz_lr, dim3((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG), 0, st, b,
                     p, gq, rho_dev, Cl, g, lambda, status, dz);
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
