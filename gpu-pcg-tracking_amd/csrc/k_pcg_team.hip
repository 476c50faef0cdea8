// k_pcg_team.hip -- PCG for horizons 128 < N <= 256: one instance split
// over a team of two workgroups (two CUs), each holding its half of the
// Schur system on chip like k_pcg_res.hip (gfx950).
//
// Half 0 owns knots [0, N0), half 1 owns [N0, N), N0 = ceil(N/2).  Half 0
// also keeps S(N0, N0-1) for its last row's S(N0-1, N0) p_{N0} term; each
// half keeps the neighbour's boundary p in a halo slot.  Per iteration the
// halves exchange, through an L2 mailbox: their p.Sp partials, then their
// r.z partials together with z at the boundary row, from which each side
// forms the neighbour's new boundary p with the same fused multiply-add the
// owner uses.  Totals are summed half 0 + half 1 on both sides, so both
// halves take identical decisions.  Same algorithm and exit rule as
// k_pcg.hip (GPU-PCG pcg<T,14,N>, toplevel.cuh:143-157).
//
// The launch is cooperative (all workgroups co-resident); the two halves of
// a team sit 8 workgroup ids apart so they land on the same XCD and share
// its L2.  Mailbox words carry their epoch, so no memory fences are needed.
// Every wait is bounded: after ~2^22 polls a half gives up, sets the error
// word and leaves its instance unsolved.
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int NT = 1024;
constexpr int NGRP = NT / 16;
constexpr int NWAV = NT / 64;
constexpr int RK = 2;
constexpr int SLOTS = NGRP * RK + 1;
constexpr uint32_t SPIN_LIMIT = 1u << 22;

struct TeamShared {
  float L[SLOTS * NXX];
  float pv[SLOTS * NX];
  float rv[SLOTS * NX];
  float red[2][NWAV];
  float send[16], recv[16];
  uint32_t ok;
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wg_sum(float v, float *red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float s = 0.f;
#pragma unroll
  for (int w = 0; w < NWAV; ++w) s += red[w];
  return s;
}

// Send sh.send[0..n) to the partner, receive its payload into sh.recv.
// Every mailbox word carries its epoch, so relaxed (fence-free) atomics
// suffice: a word is taken only once its tag shows the current epoch.
// Returns false (uniformly) if the partner did not answer in time.
__device__ bool team_exchange(TeamMailbox *mb, int half, uint32_t epoch, TeamShared &sh, int n,
                              uint32_t *err) {
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    unsigned long long *mine = mb->w[epoch & 1u][half];
    const unsigned long long *theirs = mb->w[epoch & 1u][1 - half];
    const unsigned long long tag = (unsigned long long)epoch << 32;
    if (lane < n)
      __hip_atomic_store(mine + lane, tag | __float_as_uint(sh.send[lane]), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
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
  __syncthreads();
  const bool ok = sh.ok != 0;
  if (!ok && threadIdx.x == 0) atomicOr(err, 1u);
  return ok;
}

__global__ void __launch_bounds__(NT) k_pcg_team(Batch b, const uint32_t *__restrict__ team_inst,
                                                 uint32_t n_teams, TeamMailbox *mboxes,
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
  const int tid = threadIdx.x, g = tid >> 4, i = tid & 15;
  const int half = (int)((blockIdx.x >> 3) & 1u);
  const uint32_t pair = (blockIdx.x >> 4) * 8u + (blockIdx.x & 7u);
  const uint32_t n_pairs = gridDim.x >> 1;
  TeamMailbox *mb = mboxes + pair;
  uint32_t epoch = 0;

  for (uint32_t team = pair; team < n_teams; team += n_pairs) {
    const uint32_t inst = team_inst[team];
    if (status && !(status[inst] & ST_ACTIVE)) continue;  // same decision on both halves
    const uint32_t N = inst_len(b, inst), K0 = inst_first(b, inst);
    const uint32_t N0 = (N + 1) / 2;
    const uint32_t kb = half ? N0 : 0u, ke = half ? N : N0;  // own knots [kb, ke)
    const uint32_t off = half ? 1u : 0u;                     // slot(k) = k - kb + off
    const uint32_t n_own = ke - kb;
    const uint32_t sh_slot = half ? 0u : n_own;  // slot of the halo knot (N0 - 1 or N0)
    const bool lane_ok = i < NX;

    // L blocks of own knots (+ S(N0, N0-1) on half 0) into LDS
    const uint32_t l_first = half ? kb : 1u, l_last = half ? ke : N0 + 1;  // knots [l_first, l_last)
    for (uint32_t idx = tid; idx < (l_last - l_first) * NXX; idx += NT) {
      const uint32_t k = l_first + idx / NXX, e = idx % NXX;
      sh.L[(k - kb + off) * NXX + e] = S[(size_t)S_STRIDE * (K0 + k) + e];
    }
    float D[RK][NX], P[RK][NX];
    float lam[RK], r[RK], p[RK], y[RK], z[RK];
    bool has[RK];
    uint32_t kn[RK];
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      kn[t] = kb + (uint32_t)g + NGRP * t;
      has[t] = lane_ok && kn[t] < ke;
      const float *Sd = S + (size_t)S_STRIDE * (K0 + kn[t]) + NXX;
      const float *Pd = Pinv + (size_t)S_STRIDE * (K0 + kn[t]) + NXX;
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        D[t][j] = has[t] ? Sd[i + NX * j] : 0.f;
        P[t][j] = has[t] ? Pd[i + NX * j] : 0.f;
      }
      lam[t] = has[t] ? lambda[(size_t)NX * (K0 + kn[t]) + i] : 0.f;
      if (has[t]) sh.pv[(kn[t] - kb + off) * NX + i] = lam[t];
    }
    // boundary row of this half (its value goes to the partner)
    const uint32_t kbound = half ? kb : ke - 1;
    const int bg = (int)((kbound - kb) % NGRP), bt = (int)((kbound - kb) / NGRP);

    // halo lambda
    __syncthreads();
    if (tid < NX) sh.send[tid] = sh.pv[(kbound - kb + off) * NX + tid];
    bool alive = team_exchange(mb, half, ++epoch, sh, NX, err);
    if (alive && tid < NX) sh.pv[sh_slot * NX + tid] = sh.recv[tid];
    __syncthreads();

    auto smv = [&]() {
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        float acc = 0.f;
        if (has[t]) {
          const uint32_t k = kn[t], ks = k - kb + off;
          const float *x0 = sh.pv + ks * NX;
#pragma unroll
          for (int j = 0; j < NX; ++j) acc += D[t][j] * x0[j];
          if (k > 0) {
            const float *Lb = sh.L + ks * NXX;
            const float *xm = x0 - NX;
#pragma unroll
            for (int j = 0; j < NX; ++j) acc += Lb[i + NX * j] * xm[j];
          }
          if (k + 1 < N) {
            const float *Ln = sh.L + (ks + 1) * NXX;
            const float *xp = x0 + NX;
#pragma unroll
            for (int j = 0; j < NX; ++j) acc += Ln[NX * i + j] * xp[j];
          }
        }
        y[t] = acc;
      }
    };
    auto precond = [&]() {
      float part = 0.f;
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        float acc = 0.f;
        if (has[t]) {
          const float *rr = sh.rv + (kn[t] - kb + off) * NX;
#pragma unroll
          for (int j = 0; j < NX; ++j) acc += P[t][j] * rr[j];
          part += r[t] * acc;
        }
        z[t] = acc;
      }
      return part;
    };
    auto stage_r = [&]() {
#pragma unroll
      for (int t = 0; t < RK; ++t)
        if (has[t]) sh.rv[(kn[t] - kb + off) * NX + i] = r[t];
      wave_sync();
    };
    // z of the boundary row into sh.send[1..14]
    auto send_zb = [&]() {
#pragma unroll
      for (int t = 0; t < RK; ++t)
        if (t == bt && g == bg && lane_ok) sh.send[1 + i] = z[t];
    };

    // prologue
    smv();
#pragma unroll
    for (int t = 0; t < RK; ++t) r[t] = has[t] ? gamma[(size_t)NX * (K0 + kn[t]) + i] - y[t] : 0.f;
    stage_r();
    float part = wg_sum(precond(), sh.red[0]);
    if (tid == 0) sh.send[0] = part;
    send_zb();
    alive = alive && team_exchange(mb, half, ++epoch, sh, 1 + NX, err);
    float eta = half ? sh.recv[0] + part : part + sh.recv[0];
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      p[t] = z[t];
      if (has[t]) sh.pv[(kn[t] - kb + off) * NX + i] = p[t];
    }
    if (tid < NX) sh.pv[sh_slot * NX + tid] = sh.recv[1 + tid];
    uint32_t it = 0;
    bool conv = fabsf(eta) < tol;
    while (alive && !conv && it < max_iter) {
      __syncthreads();  // p and the halo visible
      smv();
      float pp = 0.f;
#pragma unroll
      for (int t = 0; t < RK; ++t) pp += p[t] * y[t];
      pp = wg_sum(pp, sh.red[1]);
      if (tid == 0) sh.send[0] = pp;
      if (!(alive = team_exchange(mb, half, ++epoch, sh, 1, err))) break;
      const float pAp = half ? sh.recv[0] + pp : pp + sh.recv[0];
      const float alpha = eta / pAp;
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        lam[t] += alpha * p[t];
        r[t] -= alpha * y[t];
      }
      stage_r();
      const float rz = wg_sum(precond(), sh.red[0]);
      if (tid == 0) sh.send[0] = rz;
      send_zb();
      if (!(alive = team_exchange(mb, half, ++epoch, sh, 1 + NX, err))) break;
      const float eta_new = half ? sh.recv[0] + rz : rz + sh.recv[0];
      ++it;
      if (fabsf(eta_new) < tol) {
        conv = true;
        break;
      }
      const float beta = eta_new / eta;
      eta = eta_new;
#pragma unroll
      for (int t = 0; t < RK; ++t) {
        p[t] = fmaf(beta, p[t], z[t]);
        if (has[t]) sh.pv[(kn[t] - kb + off) * NX + i] = p[t];
      }
      if (tid < NX) {
        float *ph = sh.pv + sh_slot * NX;
        ph[tid] = fmaf(beta, ph[tid], sh.recv[1 + tid]);
      }
    }
    __syncthreads();
    if (alive) {
#pragma unroll
      for (int t = 0; t < RK; ++t)
        if (has[t]) lambda[(size_t)NX * (K0 + kn[t]) + i] = lam[t];
      if (half == 0 && tid == 0) {
        if (iters_out) iters_out[inst] = it;
        if (flag_out) flag_out[inst] = conv ? 0u : 1u;
      }
    }
    __syncthreads();  // LDS reuse by the next team
  }
}

__global__ void k_reset_mailboxes(TeamMailbox *mb, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * 64u) (&mb[0].w[0][0][0])[i] = 0ull;
}

}  // namespace

uint32_t pcg_team_max_knots() { return 2u * NGRP * RK; }

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
  uint32_t grid = 16u * ((n_teams + 7u) / 8u);
  if (grid > cap) grid = cap;
  if (grid / 2 > n_mb) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_reset_mailboxes, dim3((grid / 2 * 64 + 255) / 256), dim3(256), 0, st, mb, grid / 2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  Batch bb = b;
  void *args[] = {&bb, (void *)&team_inst, &n_teams, &mb, (void *)&S, (void *)&Pinv, (void *)&gamma, &lambda,
                  &max_iter, &tol, (void *)&status, &iters, &flag, &err};
  return hipLaunchCooperativeKernel((const void *)k_pcg_team, dim3(grid), dim3(NT), args,
                                    (unsigned)sizeof(TeamShared), st);
}

}  // namespace mpc
