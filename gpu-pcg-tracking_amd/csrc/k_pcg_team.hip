// k_pcg_team.hip -- PCG for horizons 128 < N <= 512: one instance split
// over a team of Q = 2 (N <= 256) or Q = 4 (N <= 512) workgroups (Q CUs),
// each holding its part of the Schur system on chip with the block-row
// arithmetic of pcg_onchip.h (gfx950).
//
// Part p owns knots [p C, min(N, (p+1) C)), C = ceil(N/Q) rounded up to a
// multiple of 4 (at most 128); each workgroup's 32 groups of 16 lanes serve 4
// consecutive knots each (local knot 4 g + t).  Every part but the last also
// keeps L_ke = S(ke, ke-1) of the next part's first knot in the LDS slot
// after its own, for its last row's S(ke-1, ke) p_ke term.  The knots next
// to a split (the last of a part's last group, the first of the next part's
// group 0) take their outer neighbour's vector element from that part.
//
// With the stair preconditioner (STAIR) a third exchange per iteration
// carries the rows of w = D^-1 r at the splits for the extra block product.
//
// Exchange through an L2 mailbox, twice per iteration (p.Sp, then r.z with
// z of the rows next to the splits): before the workgroup barrier that
// publishes the wave partials in LDS, every wave also posts its partial to
// the mailbox, and the boundary groups post their rows; after the barrier
// every wave polls the other parts' partials itself (the boundary waves also
// the rows they need), so no second barrier is needed.  Totals are
// ((part 0 + part 1) + part 2) + part 3, each part's the sum of its waves in
// wave order, on every part, so all parts take identical decisions.  Same
// algorithm and exit rule as k_pcg.hip (GPU-PCG pcg<T,14,N>,
// toplevel.cuh:143-157).
//
// Two launch forms, both cooperative (all workgroups co-resident):
// - k_pcg_team: one team per Q workgroups; more teams than fit at once are
//   solved by consecutive launches (four-workgroup teams, single instances);
// - k_pcg_pool: one persistent launch of one workgroup per CU for a PCG
//   phase's two-workgroup teams and 512-lane resident bins.  Workgroups are
//   paired into team slots; the slot's part 0 takes the next team from a
//   queue (a device counter) and passes its index to part 1 through the
//   slot's mailbox; when the team queue is empty every workgroup takes
//   bins from a second queue (resident::body).  A CU whose team converged
//   early takes the next item instead of idling until the slowest team of a
//   static launch is done.  Teams are queued before bins and the bins
//   largest first.
// The parts of a team sit 8 workgroup ids apart so they land on the same
// XCD and share its L2.  Mailbox words carry their epoch, so relaxed
// (fence-free) atomics suffice; four epoch buffers keep a word until every
// reader is done with it, and a pool slot alternates between two sets of
// them from one team to the next (TeamMailbox).  Every wait is bounded: a
// wave that gives up sets the error word and the workgroup's fail flag, and
// the workgroup leaves the solve at its next barrier; its partners then
// time out as well.  A part counts itself in the instance's `done` word
// only after a solve without a fault, and the finalising kernel after the
// launch sets PCG_FLAG_FAULT for every team instance whose count is short of
// Q (an abandoned solve, or a team the pool never reached), so a fault wins
// over the result part 0 writes, whatever the order.
// `poison` (host test hook; 0 = off) makes part poison - 1 of every team
// post its words under a wrong epoch tag: its partners' waits then run out
// and every team ends with PCG_FLAG_FAULT, every workgroup exiting
// normally.  Under stream
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
constexpr int KNOTS = NGRP * RK;  // knots per part
constexpr uint32_t SPIN_LIMIT = 1u << 16;
// mailbox words of a part: 0..7 wave partials, MB_FIRST.. its first row,
// MB_LAST.. its last row; lanes of a polling wave: 0 .. 8(Q-1)-1 the other
// parts' partials, LANE_M.. the row before the part, LANE_P.. the row after
constexpr uint32_t MB_FIRST = 8, MB_LAST = 24, LANE_M = 32, LANE_P = 48;
static_assert(MB_FIRST == NWAV && MB_LAST >= MB_FIRST + NX && MB_LAST + NX <= 40, "mailbox layout");
static_assert(8 * 3 <= LANE_M && LANE_M + NX <= LANE_P && LANE_P + NX <= 64, "polling lanes");

struct TeamShared {
  float L[(KNOTS + 1) * L_SLOT + L_GUARD];
  float zb[(KNOTS + 1) * 16];
  float wb[(KNOTS + 1) * 16];  // stair preconditioner: staged D^-1 r
  float red[2][NWAV];
  float recv[2][16];  // rows from the part before / after
  uint32_t fail;
};
static_assert(((KNOTS + 1) * L_SLOT + L_GUARD) % 4 == 0, "L and zb are cleared as one float4 range");

__device__ __forceinline__ float sum8(const float *red) {
  float s = 0.f;
#pragma unroll
  for (int v = 0; v < NWAV; ++v) s += red[v];
  return s;
}

// x of lanes l0 .. l0+7 (l0 uniform), summed in lane order (the order of sum8)
__device__ __forceinline__ float sum_lanes8(float x, uint32_t l0) {
  float s = 0.f;
#pragma unroll
  for (uint32_t v = 0; v < (uint32_t)NWAV; ++v)
    s += __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), (int)(l0 + v)));
  return s;
}

// tag_xor: 0, or a bit pattern that makes the tag wrong (the poison hook)
// buffer of epoch e in the set `set` (0 or 1) of the mailbox
__device__ __forceinline__ uint32_t mb_buf(uint32_t set, uint32_t e) { return (set << 2) | (e & 3u); }

__device__ __forceinline__ void mb_post(TeamMailbox *mb, uint32_t set, uint32_t part, uint32_t e, uint32_t word,
                                        float v, uint32_t tag_xor) {
  __hip_atomic_store(&mb->w[mb_buf(set, e)][part][word],
                     ((unsigned long long)(e ^ tag_xor) << 32) | __float_as_uint(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Lanes with `want` wait for the mailbox word at src to carry epoch e and
// return its value (0 in other lanes, whose src is not read); ok is cleared
// when a word did not arrive in time.  Every mailbox word carries its epoch,
// so a word is taken only once its tag shows the current epoch.
__device__ __forceinline__ float mb_wait(const unsigned long long *src, bool want, uint32_t e, bool &ok) {
  bool got = !want;
  unsigned long long v = 0;
  uint32_t spins = 0;
  while (!__all(got)) {
    if (!got) {
      v = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      got = (uint32_t)(v >> 32) == e;
    }
    if (++spins > SPIN_LIMIT) break;
    if (!__all(got)) __builtin_amdgcn_s_sleep(1);
  }
  if (!__all(got)) ok = false;
  return want ? __uint_as_float((uint32_t)v) : 0.f;
}

// Part `part` of the PCG of team `team` (instance team_inst[team]) on a
// 512-lane workgroup with the TeamShared block at smem, exchanging through
// mailbox mb, buffer set `set`, epochs e0 + 1, e0 + 2, ...  Returns false
// when a wait ran out (the solve was abandoned; `done` not counted).
template <bool STAIR, int Q>
__device__ __forceinline__ bool team_solve(uint32_t tid, uint32_t part, uint32_t team, uint32_t set, uint32_t e0,
                                           TeamMailbox *mb,
                                           unsigned char *smem, Batch b, const uint32_t *__restrict__ team_inst,
                                           const float *__restrict__ S, const float *__restrict__ Pinv,
                                           const float *__restrict__ gamma, float *__restrict__ lambda,
                                           uint32_t max_iter, float tol, const uint32_t *__restrict__ status,
                                           uint32_t *__restrict__ iters_out, uint32_t *__restrict__ flag_out,
                                           float *__restrict__ eta_out, uint32_t *__restrict__ done,
                                           uint32_t *__restrict__ err, uint32_t poison) {
  TeamShared &sh = *reinterpret_cast<TeamShared *>(smem);
  const uint32_t g = tid >> 4, i = tid & 15u, w = tid >> 6, lane64 = tid & 63u;
  const uint32_t d = ror_step(tid);
  static_assert(Q == 2 || Q == 4, "teams of 2 or 4 workgroups");
  const uint32_t inst = team_inst[team];
  if (status && !(status[inst] & ST_ACTIVE)) return true;  // same decision on every part
  const uint32_t tag_xor = poison == part + 1u ? 0x80000000u : 0u;
  const bool lane_ok = i < (uint32_t)NX;
  const uint32_t N = inst_len(b, inst), K0 = inst_first(b, inst);
  const uint32_t chunk = (((N + Q - 1) / Q) + 3u) & ~3u;  // <= 128 for N <= 128 Q
  const uint32_t kb = part * chunk, ke = kb + chunk < N ? kb + chunk : N;  // own knots [kb, ke)
  const uint32_t n_own = ke - kb;  // > 0: the host gives Q = 4 teams N > 256
  const uint32_t cap = pcg_cap(max_iter, N);
  const bool has_prev = part > 0, has_next = ke < N;

  {
    float4 *z4 = reinterpret_cast<float4 *>(sh.L);
    constexpr uint32_t n4 = ((KNOTS + 1) * L_SLOT + L_GUARD + (KNOTS + 1) * 16) / 4;
    for (uint32_t idx = tid; idx < n4; idx += NT) z4[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid == 0) sh.fail = 0u;
  }
  __syncthreads();
  // every part but the last: L_ke into the slot after its own knots
  if (has_next && g == 0 && lane_ok) {
    const float *Sn = S + (size_t)S_STRIDE * (K0 + ke);
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
  // the group's outer neighbours: in the own zb, from the next / previous
  // part, or none
  const bool own_m = g > 0 && b0 < n_own, own_p = b0 + RK < n_own;
  const bool far_m = has_prev && g == 0, far_p = has_next && b0 + RK == n_own;
  const uint32_t gl = n_own / RK - 1u, wl = gl >> 2;  // last group / its wave (n_own % RK == 0 with has_next)
  const bool first_lane = has_prev && g == 0 && lane_ok, last_lane = has_next && g == gl && lane_ok;
  float *zb0 = sh.zb + b0 * 16 + i;
  uint32_t e = e0;  // mailbox epoch
  bool ok = true;
  // a wait timed out somewhere in the workgroup: the solve is abandoned
  // (this part is not counted in `done`, so the instance gets PCG_FLAG_FAULT)
  auto fault = [&]() { return false; };
  // own first / last element to LDS, the rows at the splits to the mailbox
  auto post_rows = [&](const float (&v)[RK]) {
    if (first_lane) mb_post(mb, set, part, e, MB_FIRST + i, v[0], tag_xor);
    if (last_lane) mb_post(mb, set, part, e, MB_LAST + i, v[RK - 1], tag_xor);
  };
  auto stage = [&](const float (&v)[RK]) {
    zb0[0] = v[0];
    zb0[(RK - 1) * 16] = v[RK - 1];
    post_rows(v);
  };
  // after the barrier: the other parts' partials (every wave) and, when the
  // exchange carries them, the rows next to the part (the waves of group 0
  // and of the last group, into sh.recv); returns the partial lanes
  // each polling lane's role and mailbox word (within one epoch's parity):
  // 1 = a partial of another part, 2 = the row before the part (wave 0),
  // 3 = the row after it (wave wl)
  constexpr uint32_t PART_WORDS = sizeof(TeamMailbox{}.w[0][0]) / sizeof(unsigned long long);  // 40
  uint32_t role = 0, word = 0;
  if (lane64 < 8u * (Q - 1)) {
    const uint32_t q = lane64 >> 3;
    role = 1u;
    word = (q < part ? q : q + 1u) * PART_WORDS + (lane64 & 7u);
  } else if (has_prev && w == 0 && lane64 >= LANE_M && lane64 < LANE_M + NX) {
    role = 2u;
    word = (part - 1u) * PART_WORDS + MB_LAST + lane64 - LANE_M;
  } else if (has_next && w == wl && lane64 >= LANE_P && lane64 < LANE_P + NX) {
    role = 3u;
    word = (part + 1u) * PART_WORDS + MB_FIRST + lane64 - LANE_P;
  }
  auto collect = [&](bool partials, bool rows) {
    const bool rm = rows && has_prev && w == 0, rp = rows && has_next && w == wl;
    const bool want = (role == 1u && partials) || (role == 2u && rows) || (role == 3u && rows);
    float x = 0.f;
    if (partials || rm || rp) {
      x = mb_wait(&mb->w[mb_buf(set, e)][0][0] + word, want, e, ok);
      if (want && role == 2u) sh.recv[0][lane64 - LANE_M] = x;
      if (want && role == 3u) sh.recv[1][lane64 - LANE_P] = x;
      if (rm || rp) wave_sync();
    }
    if (!ok) {
      sh.fail = 1u;
      if (lane64 == 0) atomicOr(err, 1u);
    }
    return x;
  };
  auto neighbours = [&](float &vm, float &vp) {
    vm = own_m ? zb0[-16] : (far_m && lane_ok ? sh.recv[0][i] : 0.f);
    vp = own_p ? zb0[RK * 16] : (far_p && lane_ok ? sh.recv[1][i] : 0.f);
  };
  // stair preconditioner: w = D^-1 r staged like z, in its own buffer, and
  // its row at the split exchanged on an epoch of its own (row only)
  float *wb0 = sh.wb + b0 * 16 + i;
  auto stage_w = [&](const float (&v)[RK]) {
    wb0[0] = v[0];
    wb0[(RK - 1) * 16] = v[RK - 1];
    post_rows(v);
  };
  auto neighbours_w = [&](float &vm, float &vp) {
    vm = own_m ? wb0[-16] : (far_m && lane_ok ? sh.recv[0][i] : 0.f);
    vp = own_p ? wb0[RK * 16] : (far_p && lane_ok ? sh.recv[1][i] : 0.f);
  };
  auto smv = [&](const float (&x)[RK], float xm, float xp) { group_smv<RK>(R, sh.L + b0 * L_SLOT, i, x, xm, xp, y); };
  // ((part 0 + part 1) + part 2) + part 3: the own part from LDS, the others
  // from the partial lanes x of collect
  auto total = [&](float mine, float x) {
    float t = 0.f;
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)Q; ++q) t += q == part ? mine : sum_lanes8(x, 8u * (q < part ? q : q - 1u));
    return t;
  };

  // boundary lambda to the partner
  ++e;
  stage(lam);
  __syncthreads();
  collect(false, true);
  float pm, pp;
  neighbours(pm, pp);

  // prologue: r = gamma - S lambda, z = Pinv r, p = z, eta = r.z
  smv(lam, pm, pp);
  float part_sum = 0.f;
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
  for (int t = 0; t < RK; ++t) part_sum += r[t] * z[t];
  part_sum = wave_sum(part_sum);
  __syncthreads();  // neighbour and row reads done before zb / recv are reused
  ++e;
  if (lane64 == 0) {
    sh.red[1][w] = part_sum;
    mb_post(mb, set, part, e, w, part_sum, tag_xor);
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
  float eta_last = eta;  // of the last exit test

  while (!conv && it < cap) {
    smv(p, pm, pp);
    float pyp = 0.f;
#pragma unroll
    for (int t = 0; t < RK; ++t) pyp += p[t] * y[t];
    pyp = wave_sum(pyp);
    ++e;
    if (lane64 == 0) {
      sh.red[0][w] = pyp;
      mb_post(mb, set, part, e, w, pyp, tag_xor);
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
      mb_post(mb, set, part, e, w, rz, tag_xor);
    }
    stage(z);
    __syncthreads();
    if (sh.fail) return fault();
    float zm = own_m ? zb0[-16] : 0.f, zp = own_p ? zb0[RK * 16] : 0.f;  // issued before the wait
    const float eta_new = total(sum8(sh.red[1]), collect(true, true));
    eta_last = eta_new;
    if (far_m && lane_ok) zm = sh.recv[0][i];
    if (far_p && lane_ok) zp = sh.recv[1][i];
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
  if (tid == 0) {
    if (part == 0) {
      if (iters_out) iters_out[inst] = it;
      if (eta_out) eta_out[inst] = fabsf(eta_last);
      if (flag_out) flag_out[inst] = conv ? 0u : 1u;
    }
    if (done) atomicAdd(done + inst, 1u);
  }
  return true;
}

template <bool STAIR, int Q>
__global__ void __launch_bounds__(NT) k_pcg_team(Batch b, const uint32_t *__restrict__ team_inst,
                                                 uint32_t n_teams, uint32_t team0, TeamMailbox *mboxes,
                                                 const float *__restrict__ S, const float *__restrict__ Pinv,
                                                 const float *__restrict__ gamma, float *__restrict__ lambda,
                                                 uint32_t max_iter, float tol, const uint32_t *__restrict__ status,
                                                 uint32_t *__restrict__ iters_out, uint32_t *__restrict__ flag_out,
                                                 float *__restrict__ eta_out, uint32_t *__restrict__ done,
                                                 uint32_t *__restrict__ err, uint32_t poison) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr uint32_t QS = Q == 4 ? 2u : 1u;  // log2 Q
  const uint32_t blk = blockIdx.x;
  const uint32_t part = (blk >> 3) & (uint32_t)(Q - 1);
  const uint32_t slot = (blk >> (3u + QS)) * 8u + (blk & 7u);
  if (team0 + slot >= n_teams) return;
  team_solve<STAIR, Q>(threadIdx.x, part, team0 + slot, 0u, 0u, mboxes + slot, smem_raw, b, team_inst, S, Pinv, gamma,
                       lambda, max_iter, tol, status, iters_out, flag_out, eta_out, done, err, poison);
}

// Largest pool round: its epochs start at round << POOL_ROUND_SHIFT, so a
// team solve may use 2^POOL_ROUND_SHIFT epochs (about 2^19 PCG iterations)
// and a slot 2^(32 - POOL_ROUND_SHIFT) - 1 teams (the host checks both).
constexpr uint32_t POOL_ROUND_SHIFT = 20;

// The persistent PCG phase: team slot (blk >> 4) * 8 + (blk & 7) with part
// (blk >> 3) & 1 takes two-workgroup teams from ctr->team_next while there
// are any, then every workgroup takes bins from ctr->bin_next.
template <bool STAIR>
__global__ void __launch_bounds__(NT) k_pcg_pool(Batch b, const uint32_t *__restrict__ team_inst, uint32_t n_teams,
                                                 TeamMailbox *mboxes, PoolCounters *ctr,
                                                 const uint4 *__restrict__ items, const uint32_t *__restrict__ bin_off,
                                                 uint32_t n_bins, const float *__restrict__ S,
                                                 const float *__restrict__ Pinv, const float *__restrict__ gamma,
                                                 float *__restrict__ lambda, uint32_t max_iter, float tol,
                                                 const uint32_t *__restrict__ status, uint32_t *__restrict__ iters_out,
                                                 uint32_t *__restrict__ flag_out, float *__restrict__ eta_out,
                                                 uint32_t *__restrict__ done, uint32_t *__restrict__ err,
                                                 uint32_t poison) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ uint32_t s_item, s_round, s_team;
  const uint32_t blk = blockIdx.x;
  const uint32_t part = (blk >> 3) & 1u, slot = (blk >> 4) * 8u + (blk & 7u);
  TeamMailbox *mb = mboxes + slot;
  // One loop, one item per pass, a team or a bin in exclusive branches (as
  // two loops one after the other the kernel needs more registers than it
  // has and spills in the team iteration).
  bool teams = n_teams > 0;  // uniform: this workgroup still takes teams
  for (uint32_t r = 1;; ++r) {
    if (threadIdx.x == 0) {
      uint32_t t = 0xFFFFFFFFu, rr = r;
      if (teams) {
        // part 0 takes the next team and posts its index under the round's
        // tag; part 1 waits for that tag or a later one (part 0 runs ahead
        // past teams that are no longer active, which need no partner) and
        // adopts its round; no tag in time means no team
        if (part == 0) {
          t = atomicAdd(&ctr->team_next, 1u);
          __hip_atomic_store(&mb->next, ((unsigned long long)r << 32) | t, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        } else {
          for (uint32_t spins = 0; spins <= SPIN_LIMIT; ++spins) {
            const unsigned long long v = __hip_atomic_load(&mb->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(v >> 32) >= r) {
              t = (uint32_t)v;
              rr = (uint32_t)(v >> 32);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
      }
      const bool is_team = teams && t < n_teams;
      if (!is_team) t = atomicAdd(&ctr->bin_next, 1u);  // the team queue is done for this workgroup
      s_item = t;
      s_round = rr;
      s_team = is_team ? 1u : 0u;
    }
    __syncthreads();
    // read through readfirstlane: the item is uniform, and the compiler keeps
    // what it derives from it (instance, offsets, addresses) in scalar
    // registers, as for the static team launch's block-index-derived team
    const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_item);
    r = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_round);
    const bool is_team = __builtin_amdgcn_readfirstlane((int)s_team) != 0;
    __syncthreads();  // every lane has read the item before it is written again
    // the thread index made opaque per pass: what the bodies derive from it
    // is recomputed inside the loop instead of held in registers across it
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (is_team) {
      // consecutive teams of a slot use the two buffer sets in turn, so a late
      // reader of one team's last words is not overwritten by the next team
      if (!team_solve<STAIR, 2>(tid, part, t, r & 1u, r << POOL_ROUND_SHIFT, mb, smem_raw, b, team_inst, S, Pinv,
                                gamma, lambda, max_iter, tol, status, iters_out, flag_out, eta_out, done, err,
                                poison))
        teams = false;  // the partner is out of step: this workgroup takes no further team
    } else {
      teams = false;
      if (t >= n_bins) break;
      resident::body<STAIR, 512>(tid, t, smem_raw, b, items, bin_off, S, Pinv, gamma, lambda, max_iter, tol, status,
                                 iters_out, flag_out, eta_out);
    }
    __syncthreads();  // LDS of this item read by every wave before the next one clears it
  }
}

// Before a team launch: its n_mb mailboxes cleared, `done` of every active
// instance among teams team_inst[team0 .. team0 + n) set to 0, and the pool
// counters (ctr, may be null) reset.
__global__ void k_team_reset(TeamMailbox *mb, uint32_t n_mb, const uint32_t *__restrict__ team_inst, uint32_t team0,
                             uint32_t n, const uint32_t *__restrict__ status, uint32_t *__restrict__ done,
                             PoolCounters *ctr) {
  constexpr uint32_t W = sizeof(TeamMailbox) / sizeof(unsigned long long);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_mb * W) (&mb[0].w[0][0][0])[i] = 0ull;
  if (i < n) {
    const uint32_t inst = team_inst[team0 + i];
    if (!status || (status[inst] & ST_ACTIVE)) done[inst] = 0u;
  }
  if (ctr && i == 0) {
    ctr->team_next = 0u;
    ctr->bin_next = 0u;
  }
}

// After it: PCG_FLAG_FAULT for every active team instance that fewer than q
// parts finished (an abandoned solve, or one never started)
__global__ void k_team_finalize(const uint32_t *__restrict__ team_inst, uint32_t team0, uint32_t n, uint32_t q,
                                const uint32_t *__restrict__ status, const uint32_t *__restrict__ done,
                                uint32_t *__restrict__ flag_out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t inst = team_inst[team0 + i];
  if (status && !(status[inst] & ST_ACTIVE)) return;
  if (done[inst] != q && flag_out) flag_out[inst] = PCG_FLAG_FAULT;
}

}  // namespace

uint32_t pcg_team_max_knots(uint32_t q) { return q * (uint32_t)KNOTS; }

constexpr size_t POOL_SMEM =
    sizeof(TeamShared) > sizeof(resident::ResShared<512>) ? sizeof(TeamShared) : sizeof(resident::ResShared<512>);

// Largest cooperative grid of k_pcg_team and k_pcg_pool on this device (0:
// not supported).
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
  // every variant (block Jacobi, stair; two or four parts; pool) must fit;
  // the launches use the smallest
  const void *fns[6] = {(const void *)k_pcg_team<false, 2>, (const void *)k_pcg_team<true, 2>,
                        (const void *)k_pcg_team<false, 4>, (const void *)k_pcg_team<true, 4>,
                        (const void *)k_pcg_pool<false>,    (const void *)k_pcg_pool<true>};
  const size_t smem[6] = {sizeof(TeamShared), sizeof(TeamShared), sizeof(TeamShared),
                          sizeof(TeamShared), POOL_SMEM,          POOL_SMEM};
  int per_cu = 1 << 30;
  for (int k = 0; k < 6; ++k) {
    int n = 0;
    (void)hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem[k]);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fns[k], NT, smem[k]) != hipSuccess) n = 0;
    per_cu = n < per_cu ? n : per_cu;
  }
  (void)hipGetLastError();
  cached = (per_cu * cus) & ~15;  // whole groups of 16 (8 two-part teams, one per XCD slot)
  return (uint32_t)cached;
}

// Teams of q workgroups per cooperative launch: whole groups of 8 teams (one
// per XCD slot, each team's parts on one XCD) within the grid limit.
uint32_t pcg_team_per_launch(uint32_t q) { return pcg_team_max_grid() / (8u * q) * 8u; }

// Cooperative team launches of n_teams teams of q workgroups.
uint32_t pcg_team_launches(uint32_t n_teams, uint32_t q) {
  const uint32_t per_launch = pcg_team_per_launch(q);
  return n_teams == 0 || per_launch == 0 ? 0u : (n_teams + per_launch - 1) / per_launch;
}

// Whether the pool can take n_teams teams with max_iter iterations each (its
// epoch and round ranges) on this device.
bool pcg_pool_ok(uint32_t n_teams, uint32_t max_iter) {
  return pcg_team_max_grid() >= 16u && n_teams < (1u << (32u - POOL_ROUND_SHIFT)) - 1u &&
         (max_iter == PCG_CAP_BY_N || max_iter < (1u << (POOL_ROUND_SHIFT - 2u)));
}

namespace {

// launch fn cooperatively, or as an ordinary launch of the same grid under
// stream capture (a graph node carries no co-residency guarantee; the grid
// fits the device at once, and the bounded waits turn a partner that is
// not resident into PCG_FLAG_FAULT instead of a hang)
hipError_t launch_coop(const void *fn, uint32_t grid, void **args, size_t smem, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const hipError_t ec = hipStreamIsCapturing(st, &cs);
  if (ec != hipSuccess) return ec;
  return cs == hipStreamCaptureStatusActive
             ? hipLaunchKernel(fn, dim3(grid), dim3(NT), args, smem, st)
             : hipLaunchCooperativeKernel(fn, dim3(grid), dim3(NT), args, (unsigned)smem, st);
}

hipError_t team_reset(TeamMailbox *mb, uint32_t n_mb, const uint32_t *team_inst, uint32_t team0, uint32_t n,
                      const uint32_t *status, uint32_t *done, PoolCounters *ctr, hipStream_t st) {
  constexpr uint32_t W = sizeof(TeamMailbox) / sizeof(unsigned long long);
  const uint32_t m = n_mb * W > n ? n_mb * W : n;
  hipLaunchKernelGGL(k_team_reset, dim3((m + 255) / 256), dim3(256), 0, st, mb, n_mb, team_inst, team0, n, status,
                     done, ctr);
  return hipGetLastError();
}

hipError_t team_finalize(const uint32_t *team_inst, uint32_t team0, uint32_t n, uint32_t q, const uint32_t *status,
                         const uint32_t *done, uint32_t *flag, hipStream_t st) {
  if (n == 0 || !flag) return hipSuccess;
  hipLaunchKernelGGL(k_team_finalize, dim3((n + 255) / 256), dim3(256), 0, st, team_inst, team0, n, q, status, done,
                     flag);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_pcg_team(const Batch &b, uint32_t q, const uint32_t *team_inst, uint32_t n_teams,
                           TeamMailbox *mb, uint32_t n_mb, const float *S, const float *Pinv, const float *gamma,
                           float *lambda, uint32_t max_iter, float tol, const uint32_t *status, uint32_t *iters,
                           uint32_t *flag, float *eta, uint32_t *done, uint32_t *err, bool stair, uint32_t poison,
                           hipStream_t st) {
  if (n_teams == 0) return hipSuccess;
  if (q != 2 && q != 4) return hipErrorInvalidValue;
  const uint32_t per_launch = pcg_team_per_launch(q);  // teams per launch
  if (per_launch == 0) return hipErrorNotSupported;
  for (uint32_t team0 = 0; team0 < n_teams; team0 += per_launch) {
    const uint32_t n = n_teams - team0 < per_launch ? n_teams - team0 : per_launch;
    const uint32_t grid = 8u * q * ((n + 7u) / 8u);  // teams in whole groups of 8
    if (grid / q > n_mb) return hipErrorInvalidValue;
    hipError_t e = team_reset(mb, grid / q, team_inst, team0, n, status, done, nullptr, st);
    if (e != hipSuccess) return e;
    Batch bb = b;
    uint32_t t0 = team0;
    void *args[] = {&bb, (void *)&team_inst, &n_teams, &t0, &mb, (void *)&S, (void *)&Pinv, (void *)&gamma,
                    &lambda, &max_iter, &tol, (void *)&status, &iters, &flag, &eta, &done, &err, &poison};
    const void *fn = q == 4 ? (stair ? (const void *)k_pcg_team<true, 4> : (const void *)k_pcg_team<false, 4>)
                            : (stair ? (const void *)k_pcg_team<true, 2> : (const void *)k_pcg_team<false, 2>);
    e = launch_coop(fn, grid, args, sizeof(TeamShared), st);
    if (e == hipSuccess) e = team_finalize(team_inst, team0, n, q, status, done, flag, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_pcg_pool(const Batch &b, const uint32_t *team_inst, uint32_t n_teams, TeamMailbox *mb,
                           uint32_t n_mb, PoolCounters *ctr, const uint4 *items, const uint32_t *bin_off,
                           uint32_t n_bins, const float *S, const float *Pinv, const float *gamma, float *lambda,
                           uint32_t max_iter, float tol, const uint32_t *status, uint32_t *iters, uint32_t *flag,
                           float *eta, uint32_t *done, uint32_t *err, bool stair, uint32_t poison, hipStream_t st) {
  if (n_teams == 0 && n_bins == 0) return hipSuccess;
  const uint32_t grid = pcg_team_max_grid();
  if (!pcg_pool_ok(n_teams, max_iter) || grid / 2 > n_mb) return hipErrorInvalidValue;
  hipError_t e = team_reset(mb, grid / 2, team_inst, 0, n_teams, status, done, ctr, st);
  if (e != hipSuccess) return e;
  Batch bb = b;
  void *args[] = {&bb,          (void *)&team_inst, &n_teams, &mb,     &ctr,    (void *)&items, (void *)&bin_off,
                  &n_bins,      (void *)&S,         (void *)&Pinv, (void *)&gamma, &lambda, &max_iter, &tol,
                  (void *)&status, &iters,          &flag,    &eta,    &done,   &err,           &poison};
  e = launch_coop(stair ? (const void *)k_pcg_pool<true> : (const void *)k_pcg_pool<false>, grid, args, POOL_SMEM,
                  st);
  if (e == hipSuccess) e = team_finalize(team_inst, 0, n_teams, 2, status, done, flag, st);
  return e;
}

}  // namespace mpc
