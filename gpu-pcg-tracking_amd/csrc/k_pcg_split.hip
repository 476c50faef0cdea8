// k_pcg_split.hip -- block-Jacobi PCG in split-preconditioned form for
// gfx950: one 512-lane workgroup per bin of instances (N <= 256 each).
//
// Block-Jacobi PCG on S lambda = gamma (preconditioner Pinv = S(b,b)^-1) is,
// in exact arithmetic, plain CG on the transformed system
//     S^ lambda^ = gamma^,  S^ = W^T S W,  gamma^ = W^T gamma,  lambda = W lambda^,
// with W = diag(W_b), W_b = U_b^-1 and -S(b,b) = theta_b = U_b^T U_b
// (Cholesky).  S^ has -I on its diagonal and L^_b = W_b^T S(b,b-1) W_{b-1}
// below it; the preconditioner becomes -I, and the residuals satisfy
// r^ = W^T r, so r^T z^ = r^T Pinv r: the exit test |r^T z| < tol and every
// alpha, beta are those of the reference's PCG (GPU-PCG pcg<T,14,N>, call site
// toplevel.cuh:143-157), and the iterates map to its iterates.  Only L^ is
// kept on chip (one 14 x 14 block per knot instead of S(b,b), Pinv(b,b) and
// S(b,b-1)), and an iteration takes two block products instead of four.
//
// k_split_prep forms L^_b, W_b and gamma^_b from S and gamma (one 16-lane
// group per knot, W_b by Gauss elimination of theta_b without pivoting:
// theta is symmetric positive definite); k_pcg_split runs the iteration.
//
// Layout of the split system (the Pinv buffer's rows): L^_b at offset 0 of
// row b, W_b at offset NXX (both 14 x 14, column-major), gamma^ in its own
// buffer.
//
// Workgroup: 32 groups of 16 lanes (8 waves), group g serving knots
// 8g .. 8g+7 of its instance, lane i element i; up to 256 knots, bins of
// up to 8 instances, each on whole waves (as in pcg_res_body.h).  The
// blocks are held pre-rotated for the DPP products of pcg_onchip.h, in
// registers and LDS slots of the group's own (below).
#include "common.h"
#include "gj16.h"
#include "pcg_onchip.h"

namespace mpc {

namespace {

using namespace onchip;

constexpr int GROUP = 16;

// ---------------------------------------------------------------- prep ----
// Workgroups of PG groups overlap by one knot: group g of workgroup w serves
// row kg = (PG - 1) w + g - 1.  Every group forms W of its row; groups
// g >= 1 then take W of the row before from group g - 1 through LDS.
constexpr int PG = 16;
constexpr int PWG = PG * GROUP;

struct alignas(16) PrepLds {
  float W[NXX];  // W_kg, column-major
  float L[NXX];  // S(kg, kg-1), column-major
  float gm[16];
};

__global__ void __launch_bounds__(PWG) k_split_prep(Batch b, const float *__restrict__ S,
                                                    const float *__restrict__ gamma,
                                                    const uint32_t *__restrict__ status, float *__restrict__ Ps,
                                                    float *__restrict__ ghat) {
  __shared__ PrepLds lds[PG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const int64_t kg_raw = (int64_t)blockIdx.x * (PG - 1) + grp - 1;
  const bool in_range = kg_raw >= 0 && kg_raw < (int64_t)b.total_knots;
  const uint32_t kg = in_range ? (uint32_t)kg_raw : 0u;
  const KnotRef kr = locate(b, kg);
  const bool active = in_range && !(status && !(status[kr.inst] & ST_ACTIVE));
  const bool owner = grp > 0 && active;
  PrepLds &L = lds[grp];
  const int j = lane < NX ? lane : NX - 1;
  const float *Sr = S + (size_t)S_STRIDE * kg;
  // 1. theta = -S(kg,kg) by columns; W_kg = chol(theta)^-1 by rows
  float a[NX], x[NX];
#pragma unroll
  for (int r = 0; r < NX; ++r) a[r] = -Sr[NXX + r + NX * j];
  chol_inv_rows<NX>(lane, a, x);
  // x[c] = W^T(c, j) = W(j, c): lane j holds row j of W
  if (lane < NX) {
#pragma unroll
    for (int c = 0; c < NX; ++c) L.W[lane + NX * c] = x[c];
  }
  const bool row0 = kr.k == 0;
  for (int i = lane; i < NXX; i += GROUP) L.L[i] = row0 ? 0.f : Sr[i];
  if (lane < NX) L.gm[lane] = gamma[(size_t)NX * kg + lane];
  __syncthreads();
  // 2. row kg: M = S(kg,kg-1) W_{kg-1} (column j per lane), L^ = W_kg^T M,
  //    gamma^ = W_kg^T gamma_kg
  float *Pr = Ps + (size_t)S_STRIDE * kg;
  if (owner) {
    const PrepLds &Pv = lds[grp - 1];
    float m[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) m[r] = 0.f;
    if (!row0) {
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        const float wkj = Pv.W[k + NX * j];
#pragma unroll
        for (int r = 0; r < NX; ++r) m[r] += L.L[r + NX * k] * wkj;
      }
    }
    float lh[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < NX; ++k) acc += L.W[k + NX * r] * m[k];
      lh[r] = acc;
    }
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) Pr[r + NX * lane] = lh[r];
      float gacc = 0.f;
#pragma unroll
      for (int k = 0; k < NX; ++k) gacc += L.W[k + NX * lane] * L.gm[k];
      ghat[(size_t)NX * kg + lane] = gacc;
    }
    for (int i = lane; i < NXX; i += GROUP) Pr[NXX + i] = L.W[i];
  }
}

// --------------------------------------------------------------- solve ----
// One 512-lane workgroup (8 waves, two per SIMD, up to 256 VGPRs each) per
// bin: 32 groups of 16 lanes, group g serving knots 8g .. 8g+7 of its
// instance.  A block L^_b enters two products per iteration, L^_b x_{b-1}
// and L^_b^T x_b.  Held as rotated rows (Mr[n] = L(i, src(i, n))) the first
// is 16 multiply-adds with the vector rotated by DPP; held as rotated rows of
// its transpose (Kr[n] = L(src(i, n), i)) so is the second.  From rows
// alone the transposed product needs the products rotated instead (Horner
// form: 16 products and 15 rotated additions).  Per group, of the nine
// blocks L^_{8g} .. L^_{8g+8} (the last is the next group's first):
//   registers: rows of L^_1 .. L^_7, transposed rows of L^_3 and L^_4;
//   LDS slots: rows of L^_0, transposed rows of L^_1, L^_2 and L^_8;
// so L^_5 .. L^_7 take the Horner form.  Two 16-register buffers take the
// slot reads, each issued one product group ahead of its use, and the
// products run as four or five interleaved chains: with two waves per SIMD
// a wave's own independent instructions hide the latency of the chains.
constexpr int NT = 512;
constexpr int NGRP = NT / 16;     // 32 groups
constexpr int RK = 8;             // knots per group
constexpr int KNOTS = NGRP * RK;  // 256
constexpr int MAXI = 8;

// A block slot: rotated row element n of lane i at ((n >> 1) * 16 + i) * 2 +
// (n & 1), read as 8 x 64-bit words (each 16-lane read covers the 32 banks
// once); lanes 14, 15 read zeros (never written).
constexpr int SLOT = 256;

struct SplitShared {
  float R0[NGRP * SLOT];     // rows of L^_{8g}
  float K1[NGRP * SLOT];     // rows of L^_{8g+1}^T
  float K2[NGRP * SLOT];     // rows of L^_{8g+2}^T
  float K8[NGRP * SLOT];     // rows of L^_{8g+8}^T
  float zb[NGRP * 2 * 16];   // staged vector element of each group's first and last knot
  float red[2][MAXI][NGRP];  // dot-product partials: one per group, by instance and group
  uint32_t nlive;            // instances of the bin still iterating
};
constexpr uint32_t SPLIT_CLEAR4 = (uint32_t)(offsetof(SplitShared, nlive) / 16);
static_assert(offsetof(SplitShared, nlive) % 16 == 0, "slots, zb and red are cleared as one float4 range");

__device__ __forceinline__ void write_slot(float *Sl, uint32_t i, const float (&M)[16]) {
  float2 *s2 = reinterpret_cast<float2 *>(Sl) + i;
#pragma unroll
  for (int c = 0; c < 8; ++c) s2[16 * c] = make_float2(M[2 * c], M[2 * c + 1]);
}

__device__ __forceinline__ void read_slot(float (&M)[16], const float *Sl, uint32_t i) {
  const float2 *s2 = reinterpret_cast<const float2 *>(Sl) + i;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float2 v = s2[16 * c];
    M[2 * c] = v.x;
    M[2 * c + 1] = v.y;
  }
}

// total of the 16 lanes of each row, in every lane of the row (fixed order;
// each step adds two values in either order, so the lanes agree bitwise)
__device__ __forceinline__ float row_total(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror
  return v;
}

__device__ __forceinline__ float ror1(float v) { return dpp_f<0x121>(v); }

// (M_k^T x_k)_i for three blocks with rotated rows M_k: sum_n ror_n(m_n),
// m_n = M[(16 - n) & 15] x, in Horner form m_0 + ror_1(m_1 + ror_1(m_2 + ...
// + ror_1(m_15))): 16 products and 15 row_ror:1 additions per block, the
// three chains interleaved
__device__ __forceinline__ void rot_mtv3(const float (&M0)[16], float x0, const float (&M1)[16], float x1,
                                         const float (&M2)[16], float x2, float &c0, float &c1, float &c2) {
  // separate products: a fused multiply-add would take the rotated sum as
  // its addend, which DPP cannot supply (it rotates the first source only)
#pragma clang fp contract(off)
  float a0 = M0[1] * x0, a1 = M1[1] * x1, a2 = M2[1] * x2;
#pragma unroll
  for (int n = 14; n >= 0; --n) {
    const int k = (16 - n) & 15;
    a0 = ror1(a0) + M0[k] * x0;
    a1 = ror1(a1) + M1[k] * x1;
    a2 = ror1(a2) + M2[k] * x2;
  }
  c0 = a0;
  c1 = a1;
  c2 = a2;
}

// Four / five interleaved accumulations a_k += sum_{n=1..15} M_k[n] * ror_n(x_k),
// three terms per asm statement; one s_nop 4 ahead of the first (the wait
// states between a VALU write of x_k and its DPP read)
#define MPC_TERM(n, acc, x, m) \
  "v_fmac_f32_dpp %" #acc ", %" #x ", %" #m " row_ror:" #n " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
#define MPC_T4(n, m0, m1, m2, m3) MPC_TERM(n, 0, 4, m0) MPC_TERM(n, 1, 5, m1) MPC_TERM(n, 2, 6, m2) MPC_TERM(n, 3, 7, m3)
#define MPC_B4(pre, n0, n1, n2)                                                                                 \
  asm(pre MPC_T4(n0, 8, 9, 10, 11) MPC_T4(n1, 12, 13, 14, 15) MPC_T4(n2, 16, 17, 18, 19)                        \
      : "+v"(a), "+v"(b), "+v"(c), "+v"(e)                                                                     \
      : "v"(xa), "v"(xb), "v"(xc), "v"(xe), "v"(Ma[n0]), "v"(Mb[n0]), "v"(Mc[n0]), "v"(Me[n0]), "v"(Ma[n1]),     \
        "v"(Mb[n1]), "v"(Mc[n1]), "v"(Me[n1]), "v"(Ma[n2]), "v"(Mb[n2]), "v"(Mc[n2]), "v"(Me[n2]))
__device__ __forceinline__ void rot_fma15x4(float &a, float xa, const float (&Ma)[16], float &b, float xb,
                                            const float (&Mb)[16], float &c, float xc, const float (&Mc)[16],
                                            float &e, float xe, const float (&Me)[16]) {
  MPC_B4("s_nop 4\n", 1, 2, 3);
  MPC_B4("", 4, 5, 6);
  MPC_B4("", 7, 8, 9);
  MPC_B4("", 10, 11, 12);
  MPC_B4("", 13, 14, 15);
}
#undef MPC_B4
#undef MPC_T4
#define MPC_T5(n, m0, m1, m2, m3, m4) \
  MPC_TERM(n, 0, 5, m0) MPC_TERM(n, 1, 6, m1) MPC_TERM(n, 2, 7, m2) MPC_TERM(n, 3, 8, m3) MPC_TERM(n, 4, 9, m4)
#define MPC_B5(pre, n0, n1, n2)                                                                                  \
  asm(pre MPC_T5(n0, 10, 11, 12, 13, 14) MPC_T5(n1, 15, 16, 17, 18, 19) MPC_T5(n2, 20, 21, 22, 23, 24)          \
      : "+v"(a), "+v"(b), "+v"(c), "+v"(e), "+v"(f)                                                             \
      : "v"(xa), "v"(xb), "v"(xc), "v"(xe), "v"(xf), "v"(Ma[n0]), "v"(Mb[n0]), "v"(Mc[n0]), "v"(Me[n0]),        \
        "v"(Mf[n0]), "v"(Ma[n1]), "v"(Mb[n1]), "v"(Mc[n1]), "v"(Me[n1]), "v"(Mf[n1]), "v"(Ma[n2]), "v"(Mb[n2]), \
        "v"(Mc[n2]), "v"(Me[n2]), "v"(Mf[n2]))
__device__ __forceinline__ void rot_fma15x5(float &a, float xa, const float (&Ma)[16], float &b, float xb,
                                            const float (&Mb)[16], float &c, float xc, const float (&Mc)[16],
                                            float &e, float xe, const float (&Me)[16], float &f, float xf,
                                            const float (&Mf)[16]) {
  MPC_B5("s_nop 4\n", 1, 2, 3);
  MPC_B5("", 4, 5, 6);
  MPC_B5("", 7, 8, 9);
  MPC_B5("", 10, 11, 12);
  MPC_B5("", 13, 14, 15);
}
#undef MPC_B5
#undef MPC_T5
#undef MPC_TERM

// rotated rows of the column-major 14 x 14 block Lh (T = true: of its
// transpose); zeros where has is false and in lanes 14, 15
template <bool T = false>
__device__ __forceinline__ void load_lrows(float (&Lr)[16], bool has, uint32_t i, uint32_t d,
                                           const float *__restrict__ Lh) {
  const bool row = has && i < (uint32_t)NX;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const uint32_t j = (i + (uint32_t)n * d) & 15u;
    Lr[n] = row && j < (uint32_t)NX ? Lh[T ? j + NX * i : i + NX * j] : 0.f;
  }
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_pcg_split(Batch b, const uint4 *__restrict__ items, const uint32_t *__restrict__ bin_off,
            const float *__restrict__ Ps, const float *__restrict__ ghat, float *__restrict__ lambda,
            uint32_t max_iter, float tol, const uint32_t *__restrict__ status, uint32_t *__restrict__ iters_out,
            uint32_t *__restrict__ flag_out, float *__restrict__ eta_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  SplitShared &sh = *reinterpret_cast<SplitShared *>(smem_raw);
  const uint32_t tid = threadIdx.x, bin = blockIdx.x;
  const uint32_t g = tid >> 4, i = tid & 15u;
  const uint32_t d = ror_step(tid);
  {
    float4 *z4 = reinterpret_cast<float4 *>(smem_raw);
    for (uint32_t idx = tid; idx < SPLIT_CLEAR4; idx += NT) z4[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid == 0) sh.nlive = 0u;
  }
  // my instance in this bin (items: {inst, first group, groups, -})
  const uint32_t i0 = bin_off[bin], ni = bin_off[bin + 1] - i0;
  int slot = -1;
  uint32_t inst = 0, g0 = 0;
  for (uint32_t jx = 0; jx < ni; ++jx) {
    const uint4 it = items[i0 + jx];
    if (g >= it.y && g < it.y + it.z) {
      slot = (int)jx;
      inst = it.x;
      g0 = it.y;
    }
  }
  const bool mine = slot >= 0;
  const uint32_t N = mine ? inst_len(b, inst) : 0u;
  const uint32_t K0 = mine ? inst_first(b, inst) : 0u;
  const bool live = mine && !(status && !(status[inst] & ST_ACTIVE));
  const uint32_t cap = pcg_cap(max_iter, N);
  const uint32_t lg = g - g0;
  const bool leader = lg == 0 && i == 0 && mine;  // the instance's first lane
  const bool lane_ok = i < (uint32_t)NX;
  __syncthreads();  // LDS cleared

  // ---- load: lambda^ = W^-1 lambda by back substitution (W upper
  // triangular; each solved element broadcast to the group by DPP, every
  // lane active) and gamma^, then the blocks (fewer values live at once)
  const uint32_t b0 = lg * RK;
  float lam[RK], r[RK], p[RK], y[RK], gh[RK];  // z = -r is not stored
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    const uint32_t bk = b0 + (uint32_t)t;
    const bool has = live && bk < N && lane_ok;
    const float *Wr = Ps + (size_t)S_STRIDE * (K0 + (has ? bk : 0u)) + NXX;
    float wrow[NX];
#pragma unroll
    for (int c = 0; c < NX; ++c) wrow[c] = has ? Wr[i + NX * c] : 0.f;
    float acc = has ? lambda[(size_t)NX * (K0 + bk) + i] : 0.f;
    float lh = 0.f;
#pragma unroll
    for (int k = NX - 1; k >= 0; --k) {
      const float xk = bcast_dyn16<true>(acc / wrow[k], k);
      lh = (int)i == k ? xk : lh;
      asm volatile("" : "+v"(lh));  // select now, not after the last step
      acc -= wrow[k] * xk;
    }
    lam[t] = has ? lh : 0.f;
    gh[t] = has ? ghat[(size_t)NX * (K0 + bk) + i] : 0.f;
  }
  float *R0g = sh.R0 + g * SLOT, *K1g = sh.K1 + g * SLOT, *K2g = sh.K2 + g * SLOT, *K8g = sh.K8 + g * SLOT;
  auto blk = [&](uint32_t t, bool &has) {
    const uint32_t bk = b0 + t;
    has = live && bk < N && bk > 0;  // L^_0 = 0
    return Ps + (size_t)S_STRIDE * (K0 + (has ? bk : 0u));
  };
  {
    const uint32_t ts[4] = {0u, 1u, 2u, (uint32_t)RK};
    float *dst[4] = {R0g, K1g, K2g, K8g};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      bool has;
      const float *Pr = blk(ts[u], has);
      float Ls[16];
      if (u == 0)
        load_lrows(Ls, has, i, d, Pr);
      else
        load_lrows<true>(Ls, has, i, d, Pr);
      if (lane_ok) write_slot(dst[u], i, Ls);
      asm volatile("" ::: "memory");  // store now: Ls need not stay live across the other loads
    }
  }
  float Lr[RK - 1][16];  // Lr[t - 1]: rows of L^_t, t = 1 .. 7
  float Kr[2][16];       // Kr[t - 3]: rows of L^_t^T, t = 3, 4
#pragma unroll
  for (int t = 1; t < RK; ++t) {
    bool has;
    const float *Pr = blk((uint32_t)t, has);
    load_lrows(Lr[t - 1], has, i, d, Pr);
  }
#pragma unroll
  for (int t = 3; t < 5; ++t) {
    bool has;
    const float *Pr = blk((uint32_t)t, has);
    load_lrows<true>(Kr[t - 3], has, i, d, Pr);
  }
  const bool in_m = lg > 0, in_p = b0 + RK < N;
  // zb: [group][first, last knot][lane]
  float *zbg = sh.zb + g * 32 + i;
  auto stage = [&](const float (&v)[RK]) {
    zbg[0] = v[0];
    zbg[16] = v[RK - 1];
  };
  auto stage_neg = [&](const float (&v)[RK]) {
    zbg[0] = -v[0];
    zbg[16] = -v[RK - 1];
  };
  auto neighbours = [&](float &vm, float &vp) {
    vm = in_m ? zbg[-16] : 0.f;
    vp = in_p ? zbg[32] : 0.f;
  };
  // y = S^ x: y_t = -x_t + L^_t x_{t-1} + L^_{t+1}^T x_{t+1}; xm / xp: the
  // elements of knots 8lg - 1 and 8lg + 8.  Lanes 14, 15 of every block row
  // are zero, so y (and r, p, lambda^) stay zero there.
  // two slot buffers: each slot read is issued one product group ahead of
  // its use; A holds the rows of L^_0 between calls
  float A[16], B[16];
  read_slot(A, R0g, i);
  auto smv = [&](const float (&x)[RK], float xm, float xp) {
    read_slot(B, K1g, i);
    // first chain of every knot: rows of L^_t times x_{t-1}
    y[0] = fmaf(A[0], xm, -x[0]);
#pragma unroll
    for (int t = 1; t < 5; ++t) y[t] = fmaf(Lr[t - 1][0], x[t - 1], -x[t]);
    rot_fma15x5(y[0], xm, A, y[1], x[0], Lr[0], y[2], x[1], Lr[1], y[3], x[2], Lr[2], y[4], x[3], Lr[3]);
    read_slot(A, K2g, i);
#pragma unroll
    for (int t = 5; t < 8; ++t) y[t] = fmaf(Lr[t - 1][0], x[t - 1], -x[t]);
    y[0] = fmaf(B[0], x[1], y[0]);
    rot_fma15x4(y[5], x[4], Lr[4], y[6], x[5], Lr[5], y[7], x[6], Lr[6], y[0], x[1], B);
    read_slot(B, K8g, i);
    float c4, c5, c6;
    rot_mtv3(Lr[4], x[5], Lr[5], x[6], Lr[6], x[7], c4, c5, c6);
    // second chains: rows of L^_{t+1}^T times x_{t+1}
    y[1] = fmaf(A[0], x[2], y[1]);
    y[2] = fmaf(Kr[0][0], x[3], y[2]);
    y[3] = fmaf(Kr[1][0], x[4], y[3]);
    y[7] = fmaf(B[0], xp, y[7]);
    rot_fma15x4(y[1], x[2], A, y[2], x[3], Kr[0], y[3], x[4], Kr[1], y[7], xp, B);
    read_slot(A, R0g, i);
    y[4] += c4;
    y[5] += c5;
    y[6] += c6;
  };
  // dot products: a row total per group to LDS (red[k][instance][group]),
  // then, after the barrier, the instance's 32 entries (zero past its
  // groups) summed by every row: two per lane, then the row total
  float *redg = &sh.red[0][mine ? slot : 0][lg & (NGRP - 1)];
  auto post = [&](int k, float part) {
    const float v = row_total(part);
    if (i == 0) redg[k * MAXI * NGRP] = v;
  };
  const float2 *red2 = reinterpret_cast<const float2 *>(&sh.red[0][mine ? slot : 0][0]) + i;
  auto total = [&](int k) {
    const float2 v = red2[k * MAXI * NGRP / 2];
    return row_total(v.x + v.y);
  };
  stage(lam);
  __syncthreads();
  float pm, pp;
  neighbours(pm, pp);

  // ---- prologue: r = gamma^ - S^ lambda^, z = -r, p = z, eta = r.z
  smv(lam, pm, pp);
  float part = 0.f;
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    r[t] = lane_ok ? gh[t] - y[t] : 0.f;
    part -= r[t] * r[t];
  }
  __syncthreads();  // every neighbour lambda read before zb is reused
  post(1, part);
  stage_neg(r);
  __syncthreads();
  float eta = total(1);
  neighbours(pm, pp);
#pragma unroll
  for (int t = 0; t < RK; ++t) p[t] = -r[t];
  uint32_t it = 0;
  // tol < 0: exit relative to the first residual, |eta| < -tol |eta_0|
  const float tol_i = tol < 0.f ? fmaxf(-tol * fabsf(eta), FLT_MIN) : tol;
  bool conv = !live || fabsf(eta) < tol_i;
  float eta_last = eta;
  // the bin's count of iterating instances: each instance's first lane adds
  // itself here and leaves when it stops (between the second barrier of its
  // last iteration and the first of the next), so after the first barrier of
  // an iteration every wave sees the same count
  if (leader && !conv && cap > 0) atomicAdd(&sh.nlive, 1u);

  for (;;) {
    const bool run = live && !conv && it < cap;  // uniform over the instance's waves
    if (run) {
      smv(p, pm, pp);
      float pyp = 0.f;
#pragma unroll
      for (int t = 0; t < RK; ++t) pyp += p[t] * y[t];
      post(0, pyp);
    }
    __syncthreads();
    if (sh.nlive == 0u) break;
    if (run) {
      const float alpha = eta / total(0);
      float rz = 0.f;
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        lam[t] += alpha * p[t];
        r[t] -= alpha * y[t];
        rz -= r[t] * r[t];
      }
      stage_neg(r);
      post(1, rz);
    }
    __syncthreads();
    if (run) {
      float zm, zp;
      neighbours(zm, zp);
      const float eta_new = total(1);
      eta_last = eta_new;
      ++it;
      conv = fabsf(eta_new) < tol_i;
      if (leader && (conv || it >= cap)) atomicSub(&sh.nlive, 1u);
      if (!conv) {
        const float beta = eta_new / eta;
        eta = eta_new;
#pragma unroll
        for (int t = 0; t < RK; ++t) p[t] = fmaf(beta, p[t], -r[t]);
        pm = fmaf(beta, pm, zm);
        pp = fmaf(beta, pp, zp);
      }
    }
  }

  if (!live) return;
  // lambda = W lambda^, two knots at a time
#pragma unroll
  for (int t = 0; t < RK; t += 2) {
    float Wr[2][16];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const uint32_t bk = b0 + (uint32_t)(t + u);
      const bool has = bk < N && lane_ok;
      const float *Wb = Ps + (size_t)S_STRIDE * (K0 + (bk < N ? bk : 0u)) + NXX;
#pragma unroll
      for (int n = 0; n < 16; ++n) {
        const uint32_t j = (i + (uint32_t)n * d) & 15u;
        Wr[u][n] = has && j < (uint32_t)NX ? Wb[i + NX * j] : 0.f;
      }
    }
    float va, vb;
    rot_mv2(Wr[0], lam[t], va, Wr[1], lam[t + 1], vb);
    if (b0 + t < N && lane_ok) lambda[(size_t)NX * (K0 + b0 + t) + i] = va;
    if (b0 + t + 1 < N && lane_ok) lambda[(size_t)NX * (K0 + b0 + t + 1) + i] = vb;
  }
  if (lg == 0 && i == 0) {
    if (iters_out) iters_out[inst] = it;
    if (flag_out) flag_out[inst] = conv ? 0u : 1u;
    if (eta_out) eta_out[inst] = fabsf(eta_last);
  }
}

}  // namespace

uint32_t pcg_split_max_knots() { return (uint32_t)KNOTS; }
uint32_t pcg_split_bin_groups() { return (uint32_t)NGRP; }
uint32_t pcg_split_max_items() { return (uint32_t)MAXI; }
uint32_t pcg_split_groups(uint32_t N) {
  const uint32_t groups = (N + RK - 1) / RK;
  return (groups + 3u) & ~3u;  // whole waves
}

hipError_t launch_split_prep(const Batch &b, const float *S, const float *gamma, const uint32_t *status, float *Ps,
                             float *ghat, hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  const dim3 grid((b.total_knots + PG - 2) / (PG - 1));
  hipLaunchKernelGGL(k_split_prep, grid, dim3(PWG), 0, st, b, S, gamma, status, Ps, ghat);
  return hipGetLastError();
}

hipError_t launch_pcg_split(const Batch &b, const uint4 *items, const uint32_t *bin_off, uint32_t n_bins,
                            const float *Ps, const float *ghat, float *lambda, uint32_t max_iter, float tol,
                            const uint32_t *status, uint32_t *iters, uint32_t *flag, float *eta, hipStream_t st) {
  if (n_bins == 0) return hipSuccess;
  static bool attr_set = false;
  const size_t smem = sizeof(SplitShared);
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void *)k_pcg_split, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    (void)hipGetLastError();
    attr_set = true;
  }
  hipLaunchKernelGGL(k_pcg_split, dim3(n_bins), dim3(NT), smem, st, b, items, bin_off, Ps, ghat, lambda, max_iter,
                     tol, status, iters, flag, eta);
  return hipGetLastError();
}

}  // namespace mpc
