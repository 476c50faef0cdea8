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
// neighbour's vector element from the partner.
//
// With the stair preconditioner (STAIR) a third exchange per iteration
// carries the row of w = D^-1 r at the split for the extra block product.
//
// Exchange through an L2 mailbox, twice per iteration (p.Sp, then r.z with
// z of the row next to the split): before the workgroup barrier that
// publishes the wave partials in LDS, every wave also posts its partial to
// the partner, and the boundary group posts its row; after the barrier
// every wave polls the partner's 8 partials itself (the boundary wave also
// its row), so no second barrier is needed.  Totals are
// (sum of half 0's waves) + (sum of half 1's waves), each in wave order, on
// both sides, so both halves take identical decisions.  Same algorithm and
// exit rule as k_pcg.hip (GPU-PCG pcg<T,14,N>, toplevel.cuh:143-157).
//
// Each launch is cooperative (all workgroups co-resident) and solves one
// team per workgroup pair; more teams than fit at once are solved by
// consecutive launches.  The two halves of a team sit 8 workgroup ids apart
// so they land on the same XCD and share its L2.  Mailbox words carry their
// epoch, so relaxed (fence-free) atomics suffice.  Every wait is bounded:
// a wave that gives up sets the error word and the workgroup's fail flag,
// and the workgroup leaves at its next barrier with the instance's PCG flag
// set to PCG_FLAG_FAULT; its partner then times out as well.  Under stream
// capture (HIP graphs) the launches are ordinary kernel launches of the same
// grid, which fits the device at once (at most one workgroup per CU); the
// bounded waits turn a partner that is not co-resident into that flag
// instead of a hang.
#include "common.h"
#include "pcg_onchip.h"
#include "pcg_res_body.h"

namespace mpc {

namespace {

using namespace onchip;

constexpr int NT = 512;
constexpr int NGRP = NT / 16;  // 32 groups
constexpr int NWAV = NT / 64;  // 8 waves
constexpr int RK = 4;
constexpr int KNOTS = NGRP * RK;  // knots per half
constexpr uint32_t SPIN_LIMIT = 1u << 16;
constexpr uint32_t MB_ROW = 8;  // first mailbox word of the boundary row
static_assert(MB_ROW == NWAV && MB_ROW + NX <= 32, "mailbox layout");

struct TeamShared {
  float L[(KNOTS + 1) * L_SLOT + L_GUARD];
  float zb[(KNOTS + 1) * 16];
  float wb[(KNOTS + 1) * 16];  // stair preconditioner: staged D^-1 r
  float red[2][NWAV];
  float recv[16];
  uint32_t fail;
};
static_assert(((KNOTS + 1) * L_SLOT + L_GUARD) % 4 == 0, "L and zb are cleared as one float4 range");

__device__ __forceinline__ float sum8(const float *red) {
  float s = 0.f;
#pragma unroll
  for (int v = 0; v < NWAV; ++v) s += red[v];
  return s;
}

// x of lanes 0..7, summed in lane order (the order of sum8)
__device__ __forceinline__ float sum_lanes8(float x) {
  float s = 0.f;
#pragma unroll
  for (int v = 0; v < NWAV; ++v) s += __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), v));
  return s;
}

__device__ __forceinline__ void mb_post(TeamMailbox *mb, int half, uint32_t e, uint32_t word, float v) {
  __hip_atomic_store(&mb->w[e & 1u][half][word], ((unsigned long long)e << 32) | __float_as_uint(v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lanes lo <= lane < hi of the calling wave wait for the partner's mailbox
// word (lane) of epoch e and return its value (0 in other lanes); ok is
// cleared when a word did not arrive in time.  Every mailbox word carries
// its epoch, so a word is taken only once its tag shows the current epoch.
__device__ __forceinline__ float mb_wait(const TeamMailbox *mb, int half, uint32_t e, uint32_t lane, uint32_t lo,
                                         uint32_t hi, bool &ok) {
  const unsigned long long *theirs = mb->w[e & 1u][1 - half];
  const bool want = lane >= lo && lane < hi;
  bool got = !want;
  unsigned long long v = 0;
  uint32_t spins = 0;
  while (!__all(got)) {
    if (!got) {
      v = __hip_atomic_load(theirs + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      got = (uint32_t)(v >> 32) == e;
    }
    if (++spins > SPIN_LIMIT) break;
    if (!__all(got)) __builtin_amdgcn_s_sleep(1);
  }
  if (!__all(got)) ok = false;
  return want ? __uint_as_float((uint32_t)v) : 0.f;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Half `(blk >> 3) & 1` of team team0 + pair(blk) on a 512-lane workgroup
// with the TeamShared block at smem; blk is the workgroup's id among the
// team workgroups of the launch.
template <bool STAIR>
__device__ __forceinline__ void team_body(uint32_t blk, unsigned char *smem, Batch b,
                                          const uint32_t *__restrict__ team_inst, uint32_t n_teams, uint32_t team0,
                                          TeamMailbox *mboxes, const float *__restrict__ S,
                                          const float *__restrict__ Pinv, const float *__restrict__ gamma,
                                          float *__restrict__ lambda, uint32_t max_iter, float tol,
                                          const uint32_t *__restrict__ status, uint32_t *__restrict__ iters_out,
                                          uint32_t *__restrict__ flag_out, uint32_t *__restrict__ err) {
  TeamShared &sh = *reinterpret_cast<TeamShared *>(smem);
  const uint32_t tid = threadIdx.x, g = tid >> 4, i = tid & 15u, w = tid >> 6, lane64 = tid & 63u;
  const uint32_t d = ror_step();
  const int half = (int)((blk >> 3) & 1u);
  const uint32_t pair = (blk >> 4) * 8u + (blk & 7u);
  const uint32_t team = team0 + pair;
  if (team >= n_teams) return;
  const uint32_t inst = team_inst[team];
  if (status && !(status[inst] & ST_ACTIVE)) return;  // same decision on both halves
  TeamMailbox *mb = mboxes + pair;
  const bool lane_ok = i < (uint32_t)NX;
  const uint32_t N = inst_len(b, inst), K0 = inst_first(b, inst);
  const uint32_t N0 = (((N + 1) / 2) + 3u) & ~3u;  // <= 128 for N <= 256
  const uint32_t kb = half ? N0 : 0u, ke = half ? N : N0;  // own knots [kb, ke)
  const uint32_t n_own = ke - kb;

  {
    float4 *z4 = reinterpret_cast<float4 *>(sh.L);
    constexpr uint32_t n4 = ((KNOTS + 1) * L_SLOT + L_GUARD + (KNOTS + 1) * 16) / 4;
    for (uint32_t idx = tid; idx < n4; idx += NT) z4[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid == 0) sh.fail = 0u;
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
  // the group's outer neighbours: in the own zb, from the partner, or none
  const bool own_m = g > 0 && b0 < n_own, own_p = b0 + RK < n_own;
  const bool far_m = half == 1 && g == 0, far_p = half == 0 && b0 + RK == n_own && ke < N;
  const uint32_t gb = half ? 0u : n_own / RK - 1u, wb = gb >> 2;  // group / wave at the split
  const bool row_lane = g == gb && lane_ok;
  float *zb0 = sh.zb + b0 * 16 + i;
  uint32_t e = 0;  // mailbox epoch
  bool ok = true;
  // a wait timed out somewhere in the workgroup: the solve is abandoned and
  // the instance flagged (both halves write the same value)
  auto fault = [&]() {
    if (tid == 0 && flag_out) flag_out[inst] = PCG_FLAG_FAULT;
  };
  // own first / last element to LDS, the row at the split to the partner
  auto stage = [&](const float (&v)[RK]) {
    zb0[0] = v[0];
    zb0[(RK - 1) * 16] = v[RK - 1];
    if (row_lane) mb_post(mb, half, e, MB_ROW + i, half ? v[0] : v[RK - 1]);
  };
  // after the barrier: the partner's partials (every wave) and, when the
  // exchange carries it, its row (wave wb, into sh.recv)
  auto collect = [&](bool partials, bool row) {
    const bool rw = row && w == wb;
    float x = 0.f;
    if (partials || rw) {
      x = mb_wait(mb, half, e, lane64, partials ? 0u : MB_ROW, rw ? MB_ROW + NX : MB_ROW, ok);
      if (rw) {
        if (lane64 >= MB_ROW && lane64 < MB_ROW + NX) sh.recv[lane64 - MB_ROW] = x;
        wave_sync();
      }
    }
    if (!ok) {
      sh.fail = 1u;
      if (lane64 == 0) atomicOr(err, 1u);
    }
    return partials ? sum_lanes8(x) : 0.f;
  };
  auto neighbours = [&](float &vm, float &vp) {
    vm = own_m ? zb0[-16] : (far_m && lane_ok ? sh.recv[i] : 0.f);
    vp = own_p ? zb0[RK * 16] : (far_p && lane_ok ? sh.recv[i] : 0.f);
  };
  // stair preconditioner: w = D^-1 r staged like z, in its own buffer, and
  // its row at the split exchanged on an epoch of its own (row only)
  float *wb0 = sh.wb + b0 * 16 + i;
  auto stage_w = [&](const float (&v)[RK]) {
    wb0[0] = v[0];
    wb0[(RK - 1) * 16] = v[RK - 1];
    if (row_lane) mb_post(mb, half, e, MB_ROW + i, half ? v[0] : v[RK - 1]);
  };
  auto neighbours_w = [&](float &vm, float &vp) {
    vm = own_m ? wb0[-16] : (far_m && lane_ok ? sh.recv[i] : 0.f);
    vp = own_p ? wb0[RK * 16] : (far_p && lane_ok ? sh.recv[i] : 0.f);
  };
  auto smv = [&](const float (&x)[RK], float xm, float xp) { group_smv<RK>(R, sh.L + b0 * L_SLOT, i, x, xm, xp, y); };
  auto total = [&](float mine, float theirs) { return half ? theirs + mine : mine + theirs; };

  // boundary lambda to the partner
  ++e;
  stage(lam);
  __syncthreads();
  collect(false, true);
  float pm, pp;
  neighbours(pm, pp);

  // prologue: r = gamma - S lambda, z = Pinv r, p = z, eta = r.z
  smv(lam, pm, pp);
  float part = 0.f;
#pragma unroll
  for (int t = 0; t < RK; ++t) {
    const uint32_t lk = b0 + (uint32_t)t;
    r[t] = lk < n_own && lane_ok ? gamma[(size_t)NX * (K0 + kb + lk) + i] - y[t] : 0.f;
  }
  if constexpr (STAIR) {
    float wv[RK], v[RK], wm, wp;
    group_precond<RK>(R, r, wv);
    __syncthreads();  // every read of the lambda row in recv is done
    ++e;
    stage_w(wv);
    __syncthreads();
    if (sh.fail) return fault();
    collect(false, true);
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
  __syncthreads();  // neighbour and row reads done before zb / recv are reused
  ++e;
  if (lane64 == 0) {
    sh.red[1][w] = part;
    mb_post(mb, half, e, w, part);
  }
  stage(z);
  __syncthreads();
  if (sh.fail) return fault();
  float eta = total(sum8(sh.red[1]), collect(true, true));
  neighbours(pm, pp);
#pragma unroll
  for (int t = 0; t < RK; ++t) p[t] = z[t];
  uint32_t it = 0;
  // tol < 0: exit relative to the first residual, |eta| < -tol |eta_0|
  const float tol_i = tol < 0.f ? fmaxf(-tol * fabsf(eta), FLT_MIN) : tol;
  bool conv = fabsf(eta) < tol_i;

  while (!conv && it < max_iter) {
    smv(p, pm, pp);
    float pyp = 0.f;
#pragma unroll
    for (int t = 0; t < RK; ++t) pyp += p[t] * y[t];
    pyp = wave_sum(pyp);
    ++e;
    if (lane64 == 0) {
      sh.red[0][w] = pyp;
      mb_post(mb, half, e, w, pyp);
    }
    __syncthreads();
    if (sh.fail) return fault();
    const float alpha = eta / total(sum8(sh.red[0]), collect(true, false));
    float rz = 0.f;
#pragma unroll
    for (int t = 0; t < RK; ++t) {
      lam[t] += alpha * p[t];
      r[t] = lane_ok ? r[t] - alpha * y[t] : 0.f;
    }
    if constexpr (STAIR) {
      // z = D^-1 (2 r - S w), w = D^-1 r; the previous z row in recv was
      // consumed before this iteration's first barrier
      float wv[RK], wm, wp;
      group_precond<RK>(R, r, wv);
      ++e;
      stage_w(wv);
      __syncthreads();
      if (sh.fail) return fault();
      collect(false, true);
      neighbours_w(wm, wp);
      smv(wv, wm, wp);
#pragma unroll
      for (int t = 0; t < RK; ++t) wv[t] = lane_ok ? 2.f * r[t] - y[t] : 0.f;  // w is no longer needed
      group_precond<RK>(R, wv, z);
    } else {
      group_precond<RK>(R, r, z);
    }
#pragma unroll
    for (int t = 0; t < RK; ++t) rz += r[t] * z[t];
    rz = wave_sum(rz);
    ++e;
    if (lane64 == 0) {
      sh.red[1][w] = rz;
      mb_post(mb, half, e, w, rz);
    }
    stage(z);
    __syncthreads();
    if (sh.fail) return fault();
    float zm = own_m ? zb0[-16] : 0.f, zp = own_p ? zb0[RK * 16] : 0.f;  // issued before the wait
    const float eta_new = total(sum8(sh.red[1]), collect(true, true));
    if (far_m && lane_ok) zm = sh.recv[i];
    if (far_p && lane_ok) zp = sh.recv[i];
    ++it;
    if (fabsf(eta_new) < tol_i) {
      conv = true;
      break;
    }
    const float beta = eta_new / eta;
    eta = eta_new;
#pragma unroll
    for (int t = 0; t < RK; ++t) p[t] = fmaf(beta, p[t], z[t]);
    pm = fmaf(beta, pm, zm);
    pp = fmaf(beta, pp, zp);
  }
  __syncthreads();
  if (sh.fail) return fault();
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

template <bool STAIR>
__global__ void __launch_bounds__(NT) k_pcg_team(Batch b, const uint32_t *__restrict__ team_inst,
                                                 uint32_t n_teams, uint32_t team0, TeamMailbox *mboxes,
                                                 const float *__restrict__ S, const float *__restrict__ Pinv,
                                                 const float *__restrict__ gamma, float *__restrict__ lambda,
                                                 uint32_t max_iter, float tol, const uint32_t *__restrict__ status,
                                                 uint32_t *__restrict__ iters_out, uint32_t *__restrict__ flag_out,
                                                 uint32_t *__restrict__ err) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  team_body<STAIR>(blockIdx.x, smem_raw, b, team_inst, n_teams, team0, mboxes, S, Pinv, gamma, lambda, max_iter, tol,
                   status, iters_out, flag_out, err);
}

// The last, partly filled team launch with resident bins in its spare
// workgroups: ids below team_grid run the team body, the others bins
// bin0 + (id - team_grid).  Bins need no co-residency; they only fill
// workgroup slots that would idle until the launch ends.
template <bool STAIR>
__global__ void __launch_bounds__(NT) k_pcg_mixed(Batch b, const uint32_t *__restrict__ team_inst,
                                                  uint32_t n_teams, uint32_t team0, TeamMailbox *mboxes,
                                                  uint32_t team_grid, const uint4 *__restrict__ items,
                                                  const uint32_t *__restrict__ bin_off, uint32_t bin0,
                                                  const float *__restrict__ S, const float *__restrict__ Pinv,
                                                  const float *__restrict__ gamma, float *__restrict__ lambda,
                                                  uint32_t max_iter, float tol, const uint32_t *__restrict__ status,
                                                  uint32_t *__restrict__ iters_out, uint32_t *__restrict__ flag_out,
                                                  uint32_t *__restrict__ err) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  if (blockIdx.x < team_grid)
    team_body<STAIR>(blockIdx.x, smem_raw, b, team_inst, n_teams, team0, mboxes, S, Pinv, gamma, lambda, max_iter,
                     tol, status, iters_out, flag_out, err);
  else
    resident::body<STAIR, 512>(bin0 + (blockIdx.x - team_grid), smem_raw, b, items, bin_off, S, Pinv, gamma, lambda,
                          max_iter, tol, status, iters_out, flag_out);
}

__global__ void k_reset_mailboxes(TeamMailbox *mb, uint32_t n) {
  constexpr uint32_t W = sizeof(TeamMailbox) / sizeof(unsigned long long);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * W) (&mb[0].w[0][0][0])[i] = 0ull;
}

}  // namespace

uint32_t pcg_team_max_knots() { return 2u * KNOTS; }

constexpr size_t MIXED_SMEM =
    sizeof(TeamShared) > sizeof(resident::ResShared<512>) ? sizeof(TeamShared) : sizeof(resident::ResShared<512>);

// Largest cooperative grid of k_pcg_team (and k_pcg_mixed) on this device
// (0: not supported).
uint32_t pcg_team_max_grid() {
  static int cached = -1;
  if (cached >= 0) return (uint32_t)cached;
  int dev = 0, coop = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    cached = 0;
    return 0;
  }
  // every variant (block Jacobi, stair; teams only, mixed) must fit; the
  // launches use the smallest
  const void *fns[4] = {(const void *)k_pcg_team<false>, (const void *)k_pcg_team<true>,
                        (const void *)k_pcg_mixed<false>, (const void *)k_pcg_mixed<true>};
  const size_t smem[4] = {sizeof(TeamShared), sizeof(TeamShared), MIXED_SMEM, MIXED_SMEM};
  int per_cu = 1 << 30;
  for (int k = 0; k < 4; ++k) {
    int n = 0;
    (void)hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem[k]);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fns[k], NT, smem[k]) != hipSuccess) n = 0;
    per_cu = n < per_cu ? n : per_cu;
  }
  (void)hipGetLastError();
  cached = (per_cu * cus) & ~15;  // whole groups of 16 (8 pairs, one per XCD slot)
  return (uint32_t)cached;
}

// Launch plan of n_teams teams: the number of cooperative launches and how
// many of n_bins resident bins the last one solves in its spare workgroups.
void pcg_team_plan(uint32_t n_teams, uint32_t n_bins, uint32_t *launches, uint32_t *bins_taken) {
  *launches = 0;
  *bins_taken = 0;
  const uint32_t cap = pcg_team_max_grid();
  if (n_teams == 0 || cap < 16) return;
  const uint32_t per_launch = cap / 2;
  *launches = (n_teams + per_launch - 1) / per_launch;
  const uint32_t n_last = n_teams - (*launches - 1) * per_launch;
  const uint32_t spare = cap - 16u * ((n_last + 7u) / 8u);
  *bins_taken = spare < n_bins ? spare : n_bins;
}

hipError_t launch_pcg_team(const Batch &b, const uint32_t *team_inst, uint32_t n_teams, TeamMailbox *mb,
                           uint32_t n_mb, const uint4 *items, const uint32_t *bin_off, uint32_t n_bins,
                           uint32_t *bins_taken, const float *S, const float *Pinv, const float *gamma,
                           float *lambda, uint32_t max_iter, float tol, const uint32_t *status,
                           uint32_t *iters, uint32_t *flag, uint32_t *err, bool stair, hipStream_t st) {
  if (bins_taken) *bins_taken = 0;
  if (n_teams == 0) return hipSuccess;
  const uint32_t cap = pcg_team_max_grid();
  if (cap < 16) return hipErrorNotSupported;
  const uint32_t per_launch = cap / 2;  // teams per launch
  constexpr uint32_t W = sizeof(TeamMailbox) / sizeof(unsigned long long);
  // under stream capture: ordinary launches of the same grids (see the
  // header comment), since a graph node carries no co-residency guarantee
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  {
    const hipError_t ec = hipStreamIsCapturing(st, &cs);
    if (ec != hipSuccess) return ec;
  }
  const bool capturing = cs == hipStreamCaptureStatusActive;
  auto launch = [&](const void *fn, uint32_t grid, void **args, size_t smem) {
    return capturing ? hipLaunchKernel(fn, dim3(grid), dim3(NT), args, smem, st)
                     : hipLaunchCooperativeKernel(fn, dim3(grid), dim3(NT), args, (unsigned)smem, st);
  };
  for (uint32_t team0 = 0; team0 < n_teams; team0 += per_launch) {
    const uint32_t n = n_teams - team0 < per_launch ? n_teams - team0 : per_launch;
    const uint32_t grid = 16u * ((n + 7u) / 8u);  // pairs in whole groups of 8
    if (grid / 2 > n_mb) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reset_mailboxes, dim3((grid / 2 * W + 255) / 256), dim3(256), 0, st, mb, grid / 2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    Batch bb = b;
    uint32_t t0 = team0;
    // the last launch takes the first bins into its spare workgroups
    uint32_t nb = 0;
    if (team0 + n == n_teams && items && bins_taken) {
      uint32_t nl = 0;
      pcg_team_plan(n_teams, n_bins, &nl, &nb);
    }
    if (nb == 0) {
      void *args[] = {&bb, (void *)&team_inst, &n_teams, &t0, &mb, (void *)&S, (void *)&Pinv, (void *)&gamma,
                      &lambda, &max_iter, &tol, (void *)&status, &iters, &flag, &err};
      const void *fn = stair ? (const void *)k_pcg_team<true> : (const void *)k_pcg_team<false>;
      e = launch(fn, grid, args, sizeof(TeamShared));
    } else {
      uint32_t tg = grid, bin0 = 0;
      void *args[] = {&bb, (void *)&team_inst, &n_teams, &t0, &mb, &tg, (void *)&items, (void *)&bin_off, &bin0,
                      (void *)&S, (void *)&Pinv, (void *)&gamma, &lambda, &max_iter, &tol, (void *)&status, &iters,
                      &flag, &err};
      const void *fn = stair ? (const void *)k_pcg_mixed<true> : (const void *)k_pcg_mixed<false>;
      e = launch(fn, grid + nb, args, MIXED_SMEM);
      if (e == hipSuccess) *bins_taken = nb;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mpc
