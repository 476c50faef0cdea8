// pcg_onchip.h -- block-row arithmetic shared by the on-chip PCG kernels
// (k_pcg_res.hip, k_pcg_team.hip) for gfx950.
//
// A knot b of the Schur system is served by a 16-lane group that matches a
// DPP row of the wave: lane i < 14 owns row i of block row b; lanes 14 and
// 15 carry zeros.  A vector (p, r, z, lambda) lives in registers, element i
// of knot b in lane i.  Each group serves 4 consecutive knots, so the
// neighbours of its inner knots are in its own registers and only the
// first and last knot's outer neighbours pass through LDS.  A 14x14 block
// times a vector is computed without
// moving the vector through LDS: under the DPP rotation row_ror:n lane i
// reads the lane src(i, n) = (i + n * d) mod 16 of its row (d = +-1, found
// once per kernel by rotating the lane id), so
//     (M x)_i = sum_{n=0..15} M(i, src(i, n)) * ror_n(x)_i
// with each lane holding its row of M "rotated": Mr[n] = M(i, src(i, n)),
// zero where src >= 14 or i >= 14.  One v_fmac_f32_dpp per term.
//
// Blocks of block row b: D = S(b,b) and P = Pinv(b,b) in registers;
// L_b = S(b,b-1) in LDS, pre-rotated per lane (Lr[n] at lane offset
// i * L_LANE + n, read as 8 x 64-bit words; L_LANE = 18 spreads the 14 lanes
// over all 32 banks).  S(b,b+1) = L_{b+1}^T is applied from the rows of
// L_{b+1} in "contribution" form: lane j forms L_{b+1}(j, i) x_j for the lane
// i that reads it under row_ror:n, i.e.
//     (L^T x)_i = sum_n ror_n(Lr[(16 - n) mod 16] * x)_i,
// so no transposed copy is stored.
#pragma once
#include "common.h"

namespace mpc {
namespace onchip {

constexpr int L_LANE = 18;              // floats per lane row of a stored L block
constexpr int L_SLOT = NX * L_LANE;     // 252 floats per knot slot (lanes 14, 15 read into the next slot)
constexpr int L_GUARD = 2 * L_LANE;     // finite tail read by lanes 14, 15 past the last slot

// direction of row_ror on this device: src(i, 1) - i (mod 16), 1 or 15.
// Must run with every lane of the row active; tid = the lane's thread index.
__device__ __forceinline__ uint32_t ror_step(uint32_t tid) {
  const int i = (int)(tid & 15u);
  const int s = __builtin_amdgcn_mov_dpp(i, 0x121, 0xF, 0xF, true);
  return (uint32_t)(s - i) & 15u;
}

// Two interleaved accumulations (independent dependency chains):
//   a += sum_{n=1..15} Ma[n] * ror_n(xa),  b += sum_{n=1..15} Mb[n] * ror_n(xb).
// A leading s_nop 4 covers the wait states required between a VALU (or
// EXEC) write and a DPP read of xa / xb.
__device__ __forceinline__ void rot_fma15x2(float &a, float xa, const float (&Ma)[16], float &b, float xb,
                                            const float (&Mb)[16]) {
#define MPC_RF2(n, oa, ob)                                                                        \
  "v_fmac_f32_dpp %0, %2, %" #oa " row_ror:" #n " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
  "v_fmac_f32_dpp %1, %3, %" #ob " row_ror:" #n " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
  asm("s_nop 4\n" MPC_RF2(1, 4, 11) MPC_RF2(2, 5, 12) MPC_RF2(3, 6, 13) MPC_RF2(4, 7, 14) MPC_RF2(5, 8, 15)
          MPC_RF2(6, 9, 16) MPC_RF2(7, 10, 17)
      : "+v"(a), "+v"(b)
      : "v"(xa), "v"(xb), "v"(Ma[1]), "v"(Ma[2]), "v"(Ma[3]), "v"(Ma[4]), "v"(Ma[5]), "v"(Ma[6]), "v"(Ma[7]),
        "v"(Mb[1]), "v"(Mb[2]), "v"(Mb[3]), "v"(Mb[4]), "v"(Mb[5]), "v"(Mb[6]), "v"(Mb[7]));
  asm("s_nop 4\n" MPC_RF2(8, 4, 12) MPC_RF2(9, 5, 13) MPC_RF2(10, 6, 14) MPC_RF2(11, 7, 15) MPC_RF2(12, 8, 16)
          MPC_RF2(13, 9, 17) MPC_RF2(14, 10, 18) MPC_RF2(15, 11, 19)
      : "+v"(a), "+v"(b)
      : "v"(xa), "v"(xb), "v"(Ma[8]), "v"(Ma[9]), "v"(Ma[10]), "v"(Ma[11]), "v"(Ma[12]), "v"(Ma[13]),
        "v"(Ma[14]), "v"(Ma[15]), "v"(Mb[8]), "v"(Mb[9]), "v"(Mb[10]), "v"(Mb[11]), "v"(Mb[12]), "v"(Mb[13]),
        "v"(Mb[14]), "v"(Mb[15]));
#undef MPC_RF2
}

// (Ma xa)_i and (Mb xb)_i for rotated rows Ma, Mb of full 16-lane rows
__device__ __forceinline__ void rot_mv2(const float (&Ma)[16], float xa, float &ya, const float (&Mb)[16], float xb,
                                        float &yb) {
  ya = Ma[0] * xa;
  yb = Mb[0] * xb;
  rot_fma15x2(ya, xa, Ma, yb, xb, Mb);
}

// wave64 sum, identical in every lane, in a fixed order: quads, eights and
// rows by DPP, then (row0 + row1) + (row2 + row3)
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
  return (r0 + r1) + (r2 + r3);
}

// Registers of one block row (knot k of the instance starting at global
// knot K0) for lane i; zeros for absent rows and lanes >= 14.  Also writes
// L_k = S(k,k-1) pre-rotated into the LDS slot Ls (lanes i < 14; zeros for
// k = 0).
struct RowRegs {
  float D[16], P[16];
};

__device__ __forceinline__ void load_row(RowRegs &R, float *Ls, bool has, uint32_t i, uint32_t d,
                                         const float *__restrict__ S, const float *__restrict__ Pinv,
                                         uint32_t K0, uint32_t k) {
  const bool row = has && i < (uint32_t)NX;
  const float *Sb = S + (size_t)S_STRIDE * (K0 + k);
  const float *Pb = Pinv + (size_t)S_STRIDE * (K0 + k);
  const bool left = row && k > 0;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const uint32_t j = (i + (uint32_t)n * d) & 15u;
    const bool col = j < (uint32_t)NX;
    R.D[n] = row && col ? Sb[NXX + i + NX * j] : 0.f;
    R.P[n] = row && col ? Pb[NXX + i + NX * j] : 0.f;
    if (row) Ls[i * L_LANE + n] = left && col ? Sb[i + NX * j] : 0.f;
  }
}

__device__ __forceinline__ void read_lrow(float (&Lr)[16], const float *Ls, uint32_t i) {
  const float2 *l2 = reinterpret_cast<const float2 *>(Ls + i * L_LANE);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float2 v = l2[c];
    Lr[2 * c] = v.x;
    Lr[2 * c + 1] = v.y;
  }
}

// (L^T x)_i from the rotated rows Lr of L held by the lanes of the row.
// The products are formed first so that no DPP read directly follows the
// VALU write of its source; two partial sums (odd / even n) shorten the
// dependency chain.
__device__ __forceinline__ float rot_mtv(const float (&Lr)[16], float x) {
  float m[16];
#pragma unroll
  for (int n = 0; n < 16; ++n) m[n] = Lr[(16 - n) & 15] * x;
  float a = m[0], b = dpp_f<0x121>(m[1]);
  a += dpp_f<0x122>(m[2]);
  b += dpp_f<0x123>(m[3]);
  a += dpp_f<0x124>(m[4]);
  b += dpp_f<0x125>(m[5]);
  a += dpp_f<0x126>(m[6]);
  b += dpp_f<0x127>(m[7]);
  a += dpp_f<0x128>(m[8]);
  b += dpp_f<0x129>(m[9]);
  a += dpp_f<0x12A>(m[10]);
  b += dpp_f<0x12B>(m[11]);
  a += dpp_f<0x12C>(m[12]);
  b += dpp_f<0x12D>(m[13]);
  a += dpp_f<0x12E>(m[14]);
  b += dpp_f<0x12F>(m[15]);
  return a + b;
}

// y = S x for the RK consecutive knots of a group: Ls0 is the LDS slot of
// its first knot (slot t + 1 holds L of the next knot, finite and multiplied
// by xp = 0 after the last knot); xm / xp are the elements of the knots just
// before / after the group.  Each slot's rows are read once.
template <int RK>
__device__ __forceinline__ void group_smv(const RowRegs (&R)[RK], const float *Ls0, uint32_t i, const float (&x)[RK],
                                          float xm, float xp, float (&y)[RK]) {
  float La[16], Lb[16];
  read_lrow(La, Ls0, i);
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    float(&Lc)[16] = (t & 1) ? Lb : La;  // L_b rows of knot t
    float(&Ln)[16] = (t & 1) ? La : Lb;  // rows of knot t + 1
    read_lrow(Ln, Ls0 + (t + 1) * L_SLOT, i);
    const float xl = t > 0 ? x[t - 1] : xm, xr = t < RK - 1 ? x[t + 1] : xp;
    float a, bm;
    rot_mv2(R[t].D, x[t], a, Lc, xl, bm);
    const float bp = rot_mtv(Ln, xr);
    y[t] = (a + bm) + bp;
  }
}

// z = Pinv r for the RK knots of a group, two interleaved chains at a time
template <int RK>
__device__ __forceinline__ void group_precond(const RowRegs (&R)[RK], const float (&r)[RK], float (&z)[RK]) {
  static_assert(RK % 2 == 0, "knots are taken in pairs");
#pragma unroll
  for (int t = 0; t < RK; t += 2) rot_mv2(R[t].P, r[t], z[t], R[t + 1].P, r[t + 1], z[t + 1]);
}

}  // namespace onchip
}  // namespace mpc
