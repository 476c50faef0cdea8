// pcg_res_body.h -- the workgroup body of the resident on-chip PCG
// (gfx950), shared by its own kernel (k_pcg_res.hip) and the mixed
// cooperative launch that runs bins beside teams (k_pcg_team.hip).
//
// Same algorithm and exit rule as k_pcg.hip (GPU-PCG pcg<T,14,N>, call
// site toplevel.cuh:143-157), but every instance's matrix is read from HBM
// once per solve, and vectors never leave registers except for one
// neighbour exchange per iteration (block-row arithmetic: pcg_onchip.h).
//
// A 512-lane workgroup holds up to 128 knots: 32 groups of 16 lanes, each
// serving RK = 4 consecutive knots (4 lg + t of its instance, t = 0..3); a
// 256-lane one holds 64 knots, and two of them share a CU.
// Several instances share a workgroup ("bin"): each owns whole waves, so
// its dot products are a wave64 reduction plus a fixed-order sum over its
// own waves, and its arithmetic does not depend on which bin it is packed
// into.  Instances of a bin iterate in lockstep; a converged instance
// skips its arithmetic until the bin is done.
//
// With the stair preconditioner (STAIR), z = D^-1 (2 r - S D^-1 r): one more
// block product, one more staged vector and one more barrier per iteration.
//
// Per iteration, two workgroup barriers (block Jacobi):
//   y = S p (registers + L from LDS); p.y -> B1 -> alpha; lambda, r;
//   z = Pinv r; stage z of the group's first and last knot; r.z -> B2 ->
//   beta; p = z + beta p for the own knots and, from the staged z, for the
//   copies of the neighbouring groups' elements p_{4lg-1}[i], p_{4lg+4}[i]
//   (the same fused multiply-add as their owners, so every copy stays
//   bitwise equal to its owner).
#pragma once
#include "pcg_onchip.h"

namespace mpc {
namespace resident {

using namespace onchip;

constexpr int RK = 4;    // knots per group
constexpr int MAXI = 8;  // instances per bin

// Workgroup shapes: NT = 512 lanes (32 groups, 8 waves, 128 knots; one per
// CU) or 256 (16 groups, 4 waves, 64 knots; two per CU share its LDS).
template <int NT>
struct ResShape {
  static_assert(NT == 256 || NT == 512, "workgroup of 4 or 8 waves");
  static constexpr int NGRP = NT / 16, NWAV = NT / 64, KNOTS = NGRP * RK;
};

template <int NT>
struct ResShared {
  static constexpr int KNOTS = ResShape<NT>::KNOTS, NWAV = ResShape<NT>::NWAV;
  float L[(KNOTS + 1) * L_SLOT + L_GUARD];  // S(b,b-1) per knot slot (+ one zero slot), pre-rotated rows
  float zb[(KNOTS + 1) * 16];         // staged vector element per knot slot and lane
  float wb[(KNOTS + 1) * 16];         // stair preconditioner: staged D^-1 r
  float red[2][NWAV];
  uint32_t act[NWAV];
  static_assert(((KNOTS + 1) * L_SLOT + L_GUARD) % 4 == 0, "L and zb are cleared as one float4 range");
};

__device__ __forceinline__ float seg_sum(const float *red, uint32_t w0, uint32_t nw) {
  float s = 0.f;
  for (uint32_t v = 0; v < nw; ++v) s += red[w0 + v];
  return s;
}

// The PCG of bin `bin` (its instances: items[bin_off[bin] .. bin_off[bin+1]))
// on an NT-lane workgroup with the ResShared<NT> block at smem; tid is
// threadIdx.x (passed in so that a caller's loop can keep everything
// derived from it inside the loop).
template <bool STAIR, int NT>
__device__ __forceinline__ void body(uint32_t tid, uint32_t bin, unsigned char *smem, Batch b,
                                     const uint4 *__restrict__ items,
                                     const uint32_t *__restrict__ bin_off, const float *__restrict__ S,
                                     const float *__restrict__ Pinv, const float *__restrict__ gamma,
                                     float *__restrict__ lambda, uint32_t max_iter, float tol,
                                     const uint32_t *__restrict__ status, uint32_t *__restrict__ iters_out,
                                     uint32_t *__restrict__ flag_out, float *__restrict__ eta_out) {
  constexpr int NGRP = ResShape<NT>::NGRP, NWAV = ResShape<NT>::NWAV, KNOTS = ResShape<NT>::KNOTS;
  ResShared<NT> &sh = *reinterpret_cast<ResShared<NT> *>(smem);
  const uint32_t g = tid >> 4, i = tid & 15u, w = tid >> 6;
  const uint32_t d = ror_step(tid);
  {
    float4 *z4 = reinterpret_cast<float4 *>(sh.L);
    constexpr uint32_t n4 = ((KNOTS + 1) * L_SLOT + L_GUARD + (KNOTS + 1) * 16) / 4;
    for (uint32_t idx = tid; idx < n4; idx += NT) z4[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
  }

  // my instance in this bin (items: {inst, first group, groups, -}); a null
  // item table means one instance (0) starting at group 0
  uint32_t i0 = 0, ni = 1;
  if (items) {
    i0 = bin_off[bin];
    ni = bin_off[bin + 1] - i0;
  }
  int slot = -1;
  uint32_t inst = 0, g0 = 0, ng = NGRP;
  const uint32_t ng1 = ((inst_len(b, 0) + RK - 1) / RK + 3u) & ~3u;  // lone instance: as in a bin
  for (uint32_t j = 0; j < ni; ++j) {
    const uint4 it = items ? items[i0 + j] : make_uint4(0u, 0u, ng1, 0u);
    if (g >= it.y && g < it.y + it.z) {
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
  const uint32_t cap = pcg_cap(max_iter, N);
  const uint32_t lg = g - g0, w0 = g0 >> 2, nw = ng >> 2, lbase = g0 * RK;
  const bool lane_ok = i < (uint32_t)NX;
  __syncthreads();  // LDS cleared

  // ---- load: D, Pinv rows into registers, S(b,b-1) into LDS
  const uint32_t b0 = lg * RK, s0 = lbase + b0;  // first knot of the group and its slot
  RowRegs R[RK];
  float lam[RK], r[RK], p[RK], y[RK], z[RK];
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    const uint32_t bk = b0 + (uint32_t)t;
    const bool has = live && bk < N;
    load_row(R[t], sh.L + (s0 + t) * L_SLOT, has, i, d, S, Pinv, K0, bk);
    lam[t] = has && lane_ok ? lambda[(size_t)NX * (K0 + bk) + i] : 0.f;
  }
  // first / last element of the group to LDS; neighbouring groups' copies back
  const bool in_m = lg > 0, in_p = b0 + RK < N;
  float *zb0 = sh.zb + s0 * 16 + i;
  auto stage = [&](const float (&v)[RK]) {
    zb0[0] = v[0];
    zb0[(RK - 1) * 16] = v[RK - 1];
  };
  auto neighbours = [&](float &vm, float &vp) {
    vm = in_m ? zb0[-16] : 0.f;
    vp = in_p ? zb0[RK * 16] : 0.f;
  };
  // stair preconditioner: the first / last element of w = D^-1 r, the same way
  float *wb0 = sh.wb + s0 * 16 + i;
  auto stage_w = [&](const float (&v)[RK]) {
    wb0[0] = v[0];
    wb0[(RK - 1) * 16] = v[RK - 1];
  };
  auto neighbours_w = [&](float &vm, float &vp) {
    vm = in_m ? wb0[-16] : 0.f;
    vp = in_p ? wb0[RK * 16] : 0.f;
  };
  // y = S x for the group's knots; xm / xp: elements of knots 4lg-1, 4lg+4
  auto smv = [&](const float (&x)[RK], float xm, float xp) { group_smv<RK>(R, sh.L + s0 * L_SLOT, i, x, xm, xp, y); };
  stage(lam);
  __syncthreads();
  float pm, pp;
  neighbours(pm, pp);

  // ---- prologue: r = gamma - S lambda, z = Pinv r, p = z, eta = r.z
  smv(lam, pm, pp);
  float part = 0.f;
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    const uint32_t bk = b0 + (uint32_t)t;
    r[t] = live && bk < N && lane_ok ? gamma[(size_t)NX * (K0 + bk) + i] - y[t] : 0.f;
  }
  if constexpr (STAIR) {
    // z = D^-1 (2 r - S w), w = D^-1 r
    float wv[RK], v[RK], wm, wp;
    group_precond<RK>(R, r, wv);
    stage_w(wv);
    __syncthreads();
    neighbours_w(wm, wp);
    smv(wv, wm, wp);
#pragma unroll
    for (int t = 0; t < RK; ++t) v[t] = lane_ok ? 2.f * r[t] - y[t] : 0.f;
    group_precond<RK>(R, v, z);
  } else {
    group_precond<RK>(R, r, z);
  }
#pragma unroll
  for (int t = 0; t < RK; ++t) part += r[t] * z[t];
  part = wave_sum(part);
  __syncthreads();  // every neighbour lambda read before zb is reused
  if ((tid & 63u) == 0) sh.red[1][w] = part;
  stage(z);
  __syncthreads();
  float eta = seg_sum(sh.red[1], w0, nw);
  neighbours(pm, pp);
#pragma unroll
  for (int t = 0; t < RK; ++t) p[t] = z[t];
  uint32_t it = 0;
  // tol < 0: exit relative to the first residual, |eta| < -tol |eta_0|
  const float tol_i = tol < 0.f ? fmaxf(-tol * fabsf(eta), FLT_MIN) : tol;
  bool conv = !live || fabsf(eta) < tol_i;
  float eta_last = eta;  // of the last exit test

  for (;;) {
    const bool run = live && !conv && it < cap;  // uniform over the instance's waves
    float pyp = 0.f;
    if (run) {
      smv(p, pm, pp);
#pragma unroll
      for (int t = 0; t < RK; ++t) pyp += p[t] * y[t];
      pyp = wave_sum(pyp);
    }
    if ((tid & 63u) == 0) {
      sh.red[0][w] = pyp;
      sh.act[w] = run ? 1u : 0u;
    }
    __syncthreads();
    uint32_t any = 0;
#pragma unroll
    for (int v = 0; v < NWAV; ++v) any |= sh.act[v];
    if (!any) break;
    float rz = 0.f;
    float wv[RK];
    if (run) {
      const float alpha = eta / seg_sum(sh.red[0], w0, nw);
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        lam[t] += alpha * p[t];
        r[t] = lane_ok ? r[t] - alpha * y[t] : 0.f;
      }
      if constexpr (STAIR) {
        group_precond<RK>(R, r, wv);
        stage_w(wv);
      } else {
        group_precond<RK>(R, r, z);
#pragma unroll
        for (int t = 0; t < RK; ++t) rz += r[t] * z[t];
        stage(z);
        rz = wave_sum(rz);
      }
    }
    if constexpr (STAIR) {
      __syncthreads();  // staged w visible
      if (run) {
        float v[RK], wm, wp;
        neighbours_w(wm, wp);
        smv(wv, wm, wp);
#pragma unroll
        for (int t = 0; t < RK; ++t) v[t] = lane_ok ? 2.f * r[t] - y[t] : 0.f;
        group_precond<RK>(R, v, z);
#pragma unroll
        for (int t = 0; t < RK; ++t) rz += r[t] * z[t];
        stage(z);
        rz = wave_sum(rz);
      }
    }
    if ((tid & 63u) == 0) sh.red[1][w] = rz;
    __syncthreads();
    if (run) {
      float zm, zp;
      neighbours(zm, zp);
      const float eta_new = seg_sum(sh.red[1], w0, nw);
      eta_last = eta_new;
      ++it;
      if (fabsf(eta_new) < tol_i) {
        conv = true;
      } else {
        const float beta = eta_new / eta;
        eta = eta_new;
#pragma unroll
        for (int t = 0; t < RK; ++t) p[t] = fmaf(beta, p[t], z[t]);
        pm = fmaf(beta, pm, zm);
        pp = fmaf(beta, pp, zp);
      }
    }
  }

  if (!live) return;
#pragma unroll
  for (int t = 0; t < RK; ++t)
    if (b0 + t < N && lane_ok) lambda[(size_t)NX * (K0 + b0 + t) + i] = lam[t];
  if (lg == 0 && i == 0) {
    if (iters_out) iters_out[inst] = it;
    if (flag_out) flag_out[inst] = conv ? 0u : 1u;
    if (eta_out) eta_out[inst] = fabsf(eta_last);
  }
}

}  // namespace resident
}  // namespace mpc
