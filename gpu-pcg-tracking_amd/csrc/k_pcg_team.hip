// k_pcg_team.hip -- PCG for horizons 128 < N <= 256: one instance split
// over a team of two workgroups (two CUs), each holding its half of the
// Schur system on chip with the block-row arithmetic of pcg_onchip.h
// (gfx950).
//
// Half 0 owns knots [0, N0), half 1 owns [N0, N), N0 = ceil(N/2) rounded up
// to a multiple of 4; each workgroup's 32 groups of 16 lanes serve 4
// consecutive knots each (local knot 4 g + t).  Half 0 also keeps
// L_{N0} = S(N0, N0-1) in the LDS slot after its own, for its last row's
// S(N0-1, N0) p_{N0} term.  The knots next to the split (the last of half
// 0's group N0/4 - 1, the first of half 1's group 0) take their outer
// neighbour's vector element from the partner.  Per
// iteration the halves exchange, through an L2 mailbox: their p.Sp
// partials, then their r.z partials together with z at the boundary row,
// from which each side forms the neighbour's new p with the same fused
// multiply-add the owner uses.  Totals are summed half 0 + half 1 on both
// sides, so both halves take identical decisions.  Same algorithm and exit
// rule as k_pcg.hip (GPU-PCG pcg<T,14,N>, toplevel.cuh:143-157).
//
// Each launch is cooperative (all workgroups co-resident) and solves one
// team per workgroup pair; more teams than fit at once are solved by
// consecutive launches.  The two halves of a team sit 8 workgroup ids apart
// so they land on the same XCD and share its L2.  Mailbox words carry their epoch, so no memory fences are needed.
// Every wait is bounded: after ~2^22 polls a half gives up, sets the error
// word and leaves the kernel; its partner then times out as well.
#include "common.h"
#include "pcg_onchip.h"

namespace mpc {

namespace {

using namespace onchip;

constexpr int NT = 512;
constexpr int NGRP = NT / 16;  // 32 groups
constexpr int NWAV = NT / 64;  // 8 waves
constexpr int RK = 4;
constexpr int KNOTS = NGRP * RK;  // knots per half
constexpr uint32_t SPIN_LIMIT = 1u << 22;

struct TeamShared {
  float L[(KNOTS + 1) * L_SLOT + L_GUARD];
  float zb[(KNOTS + 1) * 16];
  float red[2][NWAV];
  float send[16], recv[16];
  uint32_t ok;
};
static_assert(((KNOTS + 1) * L_SLOT + L_GUARD) % 4 == 0, "L and zb are cleared as one float4 range");

__device__ __forceinline__ float sum8(const float *red) {
  float s = 0.f;
#pragma unroll
  for (int v = 0; v < NWAV; ++v) s += red[v];
  return s;
}

// Wave 0 only: post word0 (lane 0) and sh.send[1..n) to the partner and
// wait for its n words, stored into sh.recv; sh.ok = 0 on a time-out.
// Every mailbox word carries its epoch, so relaxed (fence-free) atomics
// suffice: a word is taken only once its tag shows the current epoch.
__device__ void mailbox_swap(TeamMailbox *mb, int half, uint32_t epoch, TeamShared &sh, float word0, int n) {
  const int lane = (int)threadIdx.x;
  unsigned long long *mine = mb->w[epoch & 1u][half];
  const unsigned long long *theirs = mb->w[epoch & 1u][1 - half];
  const unsigned long long tag = (unsigned long long)epoch << 32;
  if (lane < n) {
    const float v = lane == 0 ? word0 : sh.send[lane];
    __hip_atomic_store(mine + lane, tag | __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  bool got = lane >= n;
  unsigned long long v = 0;
  uint32_t spins = 0;
  while (!__all(got)) {
    if (!got) {
      v = __hip_atomic_load(theirs + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      got = (uint32_t)(v >> 32) == epoch;
    }
    if (++spins > SPIN_LIMIT) break;
    if (!__all(got)) __builtin_amdgcn_s_sleep(1);
  }
  const bool all_got = __all(got);
  if (lane < n && got) sh.recv[lane] = __uint_as_float((uint32_t)v);
  if (lane == 0) sh.ok = all_got ? 1u : 0u;
}

__global__ void __launch_bounds__(NT) k_pcg_team(Batch b, const uint32_t *__restrict__ team_inst,
                                                 uint32_t n_teams, uint32_t team0, TeamMailbox *mboxes,
                                                 const float *__restrict__ S,
                                                 const float *__restrict__ Pinv,
                                                 const float *__restrict__ gamma,
                                                 float *__restrict__ lambda, uint32_t max_iter,
                                                 float tol, const uint32_t *__restrict__ status,
                                                 uint32_t *__restrict__ iters_out,
                                                 uint32_t *__restrict__ flag_out,
                                                 uint32_t *__restrict__ err) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TeamShared &sh = *reinterpret_cast<TeamShared *>(smem_raw);
  const uint32_t tid = threadIdx.x, g = tid >> 4, i = tid & 15u, w = tid >> 6;
  const uint32_t d = ror_step();
  const int half = (int)((blockIdx.x >> 3) & 1u);
  const uint32_t pair = (blockIdx.x >> 4) * 8u + (blockIdx.x & 7u);
  TeamMailbox *mb = mboxes + pair;
  const bool lane_ok = i < (uint32_t)NX;
  uint32_t epoch = 0;

  {
    const uint32_t team = team0 + pair;
    if (team >= n_teams) return;
    const uint32_t inst = team_inst[team];
    if (status && !(status[inst] & ST_ACTIVE)) return;  // same decision on both halves
    const uint32_t N = inst_len(b, inst), K0 = inst_first(b, inst);
    const uint32_t N0 = (((N + 1) / 2) + 3u) & ~3u;  // <= 128 for N <= 256
    const uint32_t kb = half ? N0 : 0u, ke = half ? N : N0;  // own knots [kb, ke)
    const uint32_t n_own = ke - kb;

    {
      float4 *z4 = reinterpret_cast<float4 *>(sh.L);
      constexpr uint32_t n4 = ((KNOTS + 1) * L_SLOT + L_GUARD + (KNOTS + 1) * 16) / 4;
      for (uint32_t idx = tid; idx < n4; idx += NT) z4[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    // half 0: L_{N0} into the slot after its own knots
    if (half == 0 && N0 < N && g == 0 && lane_ok) {
      const float *Sn = S + (size_t)S_STRIDE * (K0 + N0);
      float *Ls = sh.L + n_own * L_SLOT;
#pragma unroll
      for (int n = 0; n < 16; ++n) {
        const uint32_t j = (i + (uint32_t)n * d) & 15u;
        Ls[i * L_LANE + n] = j < (uint32_t)NX ? Sn[i + NX * j] : 0.f;
      }
    }

    const uint32_t b0 = g * RK;  // first local knot of the group (and its slot)
    RowRegs R[RK];
    float lam[RK], r[RK], p[RK], y[RK], z[RK];
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      const uint32_t lk = b0 + (uint32_t)t;
      const bool has = lk < n_own;
      load_row(R[t], sh.L + lk * L_SLOT, has, i, d, S, Pinv, K0, kb + lk);
      lam[t] = has && lane_ok ? lambda[(size_t)NX * (K0 + kb + lk) + i] : 0.f;
    }
    // the group's outer neighbours: in the own zb, from the partner (recv), or none
    const bool own_m = g > 0 && b0 < n_own, own_p = b0 + RK < n_own;
    const bool far_m = half == 1 && g == 0, far_p = half == 0 && b0 + RK == n_own && ke < N;
    const bool bnd = lane_ok && (half ? g == 0 : b0 + RK == n_own);  // rows next to the split
    float *zb0 = sh.zb + b0 * 16 + i;
    auto stage = [&](const float (&v)[RK]) {
      zb0[0] = v[0];
      zb0[(RK - 1) * 16] = v[RK - 1];
      if (bnd) sh.send[1 + i] = half ? v[0] : v[RK - 1];
    };
    auto neighbours = [&](float &vm, float &vp) {
      vm = own_m ? zb0[-16] : (far_m && lane_ok ? sh.recv[1 + i] : 0.f);
      vp = own_p ? zb0[RK * 16] : (far_p && lane_ok ? sh.recv[1 + i] : 0.f);
    };
    auto smv = [&](const float (&x)[RK], float xm, float xp) {
#pragma unroll
      for (int t = 0; t < RK; ++t)
        y[t] = row_smv(R[t], sh.L + (b0 + t) * L_SLOT, i, x[t], t > 0 ? x[t - 1] : xm,
                       t < RK - 1 ? x[t + 1] : xp);
    };
    stage(lam);
    __syncthreads();
    // boundary lambda to the partner
    if (tid < 64) mailbox_swap(mb, half, ++epoch, sh, 0.f, 1 + NX);
    __syncthreads();
    if (!sh.ok) {
      if (tid == 0) atomicOr(err, 1u);
      return;
    }
    float pm, pp;
    neighbours(pm, pp);

    // prologue: r = gamma - S lambda, z = Pinv r, p = z, eta = r.z
    smv(lam, pm, pp);
    float part = 0.f;
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      const uint32_t lk = b0 + (uint32_t)t;
      r[t] = lk < n_own && lane_ok ? gamma[(size_t)NX * (K0 + kb + lk) + i] - y[t] : 0.f;
      z[t] = rot_mv(R[t].P, r[t]);
      part += r[t] * z[t];
    }
    part = wave_sum(part);
    __syncthreads();  // neighbour and halo reads done before zb / send are reused
    if ((tid & 63u) == 0) sh.red[1][w] = part;
    stage(z);
    __syncthreads();
    if (tid < 64) mailbox_swap(mb, half, ++epoch, sh, sum8(sh.red[1]), 1 + NX);
    __syncthreads();
    if (!sh.ok) {
      if (tid == 0) atomicOr(err, 1u);
      return;
    }
    float own = sum8(sh.red[1]);
    float eta = half ? sh.recv[0] + own : own + sh.recv[0];
    neighbours(pm, pp);
#pragma unroll
    for (int t = 0; t < RK; ++t) p[t] = z[t];
    uint32_t it = 0;
    bool conv = fabsf(eta) < tol;

    while (!conv && it < max_iter) {
      smv(p, pm, pp);
      float pyp = 0.f;
#pragma unroll
      for (int t = 0; t < RK; ++t) pyp += p[t] * y[t];
      pyp = wave_sum(pyp);
      if ((tid & 63u) == 0) sh.red[0][w] = pyp;
      __syncthreads();
      if (tid < 64) mailbox_swap(mb, half, ++epoch, sh, sum8(sh.red[0]), 1);
      __syncthreads();
      if (!sh.ok) {
        if (tid == 0) atomicOr(err, 1u);
        return;
      }
      own = sum8(sh.red[0]);
      const float pAp = half ? sh.recv[0] + own : own + sh.recv[0];
      const float alpha = eta / pAp;
      float rz = 0.f;
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        lam[t] += alpha * p[t];
        r[t] = lane_ok ? r[t] - alpha * y[t] : 0.f;
        z[t] = rot_mv(R[t].P, r[t]);
        rz += r[t] * z[t];
      }
      stage(z);
      rz = wave_sum(rz);
      if ((tid & 63u) == 0) sh.red[1][w] = rz;
      __syncthreads();
      if (tid < 64) mailbox_swap(mb, half, ++epoch, sh, sum8(sh.red[1]), 1 + NX);
      __syncthreads();
      if (!sh.ok) {
        if (tid == 0) atomicOr(err, 1u);
        return;
      }
      own = sum8(sh.red[1]);
      const float eta_new = half ? sh.recv[0] + own : own + sh.recv[0];
      ++it;
      if (fabsf(eta_new) < tol) {
        conv = true;
        break;
      }
      const float beta = eta_new / eta;
      eta = eta_new;
      float zm, zp;
      neighbours(zm, zp);
#pragma unroll
      for (int t = 0; t < RK; ++t) p[t] = fmaf(beta, p[t], z[t]);
      pm = fmaf(beta, pm, zm);
      pp = fmaf(beta, pp, zp);
    }
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      const uint32_t lk = b0 + (uint32_t)t;
      if (lk < n_own && lane_ok) lambda[(size_t)NX * (K0 + kb + lk) + i] = lam[t];
    }
    if (half == 0 && tid == 0) {
      if (iters_out) iters_out[inst] = it;
      if (flag_out) flag_out[inst] = conv ? 0u : 1u;
    }
  }
}

__global__ void k_reset_mailboxes(TeamMailbox *mb, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * 64u) (&mb[0].w[0][0][0])[i] = 0ull;
}

}  // namespace

uint32_t pcg_team_max_knots() { return 2u * KNOTS; }

// Largest cooperative grid of k_pcg_team on this device (0: not supported).
uint32_t pcg_team_max_grid() {
  static int cached = -1;
  if (cached >= 0) return (uint32_t)cached;
  int dev = 0, coop = 0, cus = 0, per_cu = 0;
  const size_t smem = sizeof(TeamShared);
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    cached = 0;
    return 0;
  }
  (void)hipFuncSetAttribute((const void *)k_pcg_team, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  (void)hipGetLastError();
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_team, NT, smem) != hipSuccess) per_cu = 0;
  (void)hipGetLastError();
  cached = (per_cu * cus) & ~15;  // whole groups of 16 (8 pairs, one per XCD slot)
  return (uint32_t)cached;
}

hipError_t launch_pcg_team(const Batch &b, const uint32_t *team_inst, uint32_t n_teams, TeamMailbox *mb,
                           uint32_t n_mb, const float *S, const float *Pinv, const float *gamma,
                           float *lambda, uint32_t max_iter, float tol, const uint32_t *status,
                           uint32_t *iters, uint32_t *flag, uint32_t *err, hipStream_t st) {
  if (n_teams == 0) return hipSuccess;
  const uint32_t cap = pcg_team_max_grid();
  if (cap < 16) return hipErrorNotSupported;
  const uint32_t per_launch = cap / 2;  // teams per launch
  for (uint32_t team0 = 0; team0 < n_teams; team0 += per_launch) {
    const uint32_t n = n_teams - team0 < per_launch ? n_teams - team0 : per_launch;
    const uint32_t grid = 16u * ((n + 7u) / 8u);  // pairs in whole groups of 8
    if (grid / 2 > n_mb) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reset_mailboxes, dim3((grid / 2 * 64 + 255) / 256), dim3(256), 0, st, mb, grid / 2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    Batch bb = b;
    uint32_t t0 = team0;
    void *args[] = {&bb, (void *)&team_inst, &n_teams, &t0, &mb, (void *)&S, (void *)&Pinv, (void *)&gamma,
                    &lambda, &max_iter, &tol, (void *)&status, &iters, &flag, &err};
    e = hipLaunchCooperativeKernel((const void *)k_pcg_team, dim3(grid), dim3(NT), args,
                                   (unsigned)sizeof(TeamShared), st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mpc
