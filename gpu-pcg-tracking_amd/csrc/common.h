// common.h -- batch layout shared by the kernels and the host driver.
//
// A batch is a list of instances with horizons N_i packed back to back.
// K_i = sum_{j<i} N_j is the first global knot of instance i.
//   caller-facing:  xu / dz at 21*K_i - 7*i, lambda at 14*K_i, ee_ref at 6*K_i
//   workspace:      per-knot strides (G 245, C 294, g 21, c 14, S / Pinv 588)
//                   at (stride * K_i); a single instance at K = 0 is exactly
//                   the reference layout of schur.cuh / toplevel.cuh.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace mpc {

constexpr int NQ = 7;
constexpr int NX = 14;
constexpr int NU = 7;
constexpr int NXU = 21;
constexpr int NXX = 196;
constexpr int G_STRIDE = 245;  // per-knot stride of G / G^-1: Q (196) + R (49)
constexpr int C_STRIDE = 294;  // per-knot stride of C: -A (196) + -B (98)
constexpr int S_STRIDE = 588;  // per-block-row stride of S / Pinv

struct Batch {
  const uint32_t *knot_inst;  // [total_knots] instance of each global knot (nullptr: one instance)
  const uint32_t *inst_k0;    // [n_inst + 1] first global knot of each instance
  uint32_t n_inst;
  uint32_t total_knots;
};

struct KnotRef {
  uint32_t inst, k, N, K0;
};

__device__ __forceinline__ KnotRef locate(const Batch &b, uint32_t gk) {
  KnotRef r;
  if (b.knot_inst == nullptr) {
    r.inst = 0;
    r.K0 = 0;
    r.N = b.total_knots;
  } else {
    r.inst = b.knot_inst[gk];
    r.K0 = b.inst_k0[r.inst];
    r.N = b.inst_k0[r.inst + 1] - r.K0;
  }
  r.k = gk - r.K0;
  return r;
}

// Instance of global knot gk and its place in the instance, from knot_inst
// alone: three independent reads that share cache lines, so what depends on
// them waits one round trip instead of locate's two (knot_inst, then
// inst_k0).  Instances lie back to back, each at least two knots long: a
// knot starts its instance when its predecessor belongs to another one and
// has a successor when the next knot belongs to the same one.  Its state
// x_k sits at NXU * gk - NU * inst of the batch's xu (every earlier
// instance's last knot has no u).
struct KnotEnds {
  uint32_t inst;
  bool first, has_next;
};

__device__ __forceinline__ KnotEnds knot_ends(const Batch &b, uint32_t gk) {
  KnotEnds e;
  const bool last = gk + 1 >= b.total_knots;
  if (b.knot_inst == nullptr) {
    e.inst = 0;
    e.first = gk == 0;
    e.has_next = !last;
  } else {
    e.inst = b.knot_inst[gk];
    const uint32_t ip = b.knot_inst[gk > 0 ? gk - 1 : gk], in = b.knot_inst[last ? gk : gk + 1];
    e.first = gk == 0 || ip != e.inst;
    e.has_next = !last && in == e.inst;
  }
  return e;
}

// LDS ordering among the lanes of one wave: a wave's LDS accesses execute
// in order, so only compiler reordering must be prevented; unlike
// __syncthreads() this neither waits for the other waves nor for
// outstanding global loads
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t inst_first(const Batch &b, uint32_t inst) {
  return b.knot_inst == nullptr ? 0u : b.inst_k0[inst];
}
__device__ __forceinline__ uint32_t inst_len(const Batch &b, uint32_t inst) {
  return b.knot_inst == nullptr ? b.total_knots : b.inst_k0[inst + 1] - b.inst_k0[inst];
}

// Numeric parameters passed by value to the kernels.
struct Params {
  float dt, q_ee, mu;
  float wx[14];  // per-dimension state weights (reference: 0 x 7, QD x 7)
  float wu[7];   // per-dimension control weights (reference: R x 7)
  float rho_factor, rho_max, rho_min, rho_reset;
  float pcg_tol;
  uint32_t pcg_max_iter;
  uint32_t flags;
  // SQP wall-clock budget in device clock ticks (wall_clock64), 0 = none
  unsigned long long budget_ticks;
};

// mailbox of a PCG team of up to 4 workgroups: per buffer and part, 40
// self-validating words {epoch (high 32 bits), float bits (low 32 bits)}:
// words 0..7 the partial sums of the part's 8 waves, 8..21 the vector
// elements of its first row (for the part before it), 24..37 of its last
// row (for the part after it).  Epoch e of a solve uses buffer
// 4 set + (e mod 4).  Four buffers per set: a part overwrites the words of
// epoch e only when it posts epoch e + 4, and between the two it has
// waited for the partials of every wave of every other part at an
// all-to-all epoch, which each wave posts only after its reads of epoch e
// (no two consecutive epochs are rows-only exchanges).  Two sets: the
// persistent pool alternates them between the consecutive teams of a slot.
// `next` carries the pool's team index from part 0 to part 1 {round, index}.
struct TeamMailbox {
  unsigned long long w[8][4][40];
  unsigned long long next;
};

// work queues of the persistent PCG pool (k_pcg_pool), one cache line each
struct PoolCounters {
  uint32_t team_next;
  uint32_t pad0[31];
  uint32_t bin_next;
  uint32_t pad1[31];
};

// per-instance device state of the batched SQP (struct of arrays)
struct InstState {
  float *rho;        // [n] regularisation (in/out)
  float *drho;       // [n] rho multiplier (toplevel.cuh:89)
  float *merit0;     // [n] merit of the current iterate
  uint32_t *status;  // [n] MPC_ST_* bits
  uint32_t *sqp_done;  // [n] SQP iterations taken
  float *xs;         // [14 n] initial state of each solve
  unsigned long long *t0;  // [1] device clock at the start of the solve
};

// one control step of the closed-loop harness (k_closed_loop.hip)
struct ClStep {
  uint32_t N, substeps, sqp_iters;
  float sim_dt, sigma;
  uint64_t seed;
};

constexpr uint32_t ST_ACTIVE = 1u, ST_PCG_MAXITER = 2u, ST_RHO_ABORT = 4u, ST_NONFINITE = 8u, ST_PCG_FAULT = 16u,
                   ST_TIME_BUDGET = 32u;
// per-instance PCG flag: 0 converged, 1 max_iter reached, PCG_FLAG_FAULT the
// solve was abandoned (a team exchange timed out)
constexpr uint32_t PCG_FLAG_FAULT = 2u;
constexpr uint32_t FLAG_LAST_KNOT_OWN_STATE = 1u, FLAG_PCG_STREAM = 2u, FLAG_DIRECT_SOLVE = 4u,
                   FLAG_PCG_STAIR = 8u, FLAG_PCG_CAP_BY_N = 16u;

// PCG max_iter value that selects the reference experiments' cap per
// horizon (perform_experiment.sh:14-19: 173 at N = 32, 167 at 64 and 128,
// 118 at 256, 67 at 512), taken for the tabulated horizon at or above N
constexpr uint32_t PCG_CAP_BY_N = 0xFFFFFFFFu;
__host__ __device__ __forceinline__ uint32_t pcg_cap(uint32_t max_iter, uint32_t N) {
  if (max_iter != PCG_CAP_BY_N) return max_iter;
  return N <= 32u ? 173u : N <= 128u ? 167u : N <= 256u ? 118u : 67u;
}

// host-side launchers (implemented in the .hip files)
hipError_t launch_fk(uint32_t n, const float *q, uint32_t q_stride, float *ee, hipStream_t st);
hipError_t launch_dyn_grad(uint32_t n, const float *qqdu, float *qdd, float *df_du, hipStream_t st);
hipError_t launch_inverse_dynamics(uint32_t n, const float *qqdqdd, float *tau, hipStream_t st);
// gq (optional, NQ floats per knot): the tracking-cost gradient g_q of each
// knot, the factor of the cost Hessian Q = g_q g_q^T + diag(wx).
// Cl (optional, with gq): the compact form instead of G and C (neither is
// written): per knot CL_STRIDE floats, the lower block rows (rows NQ..NX-1)
// of C's A part column-major (CL_A), then of its B part (CL_B).  The upper
// rows are fixed by the Euler step: C_A(r, c) = -(delta_rc + dt delta_{c,r+NQ})
// and C_B(r, c) = 0 for r < NQ (cl_upper_a); G follows from gq and the
// weights.  launch_expand_gc writes G and C back from it.
constexpr int CL_A = NQ * NX;          // 98
constexpr int CL_B = NQ * NU;          // 49
constexpr int CL_STRIDE = 148;         // CL_A + CL_B, padded to 16 bytes
static_assert(CL_A + CL_B <= CL_STRIDE && CL_STRIDE % 4 == 0, "compact C rows");
// C_A(r, c) for r < NQ: the negated Euler-step rows, exactly as
// launch_form_kkt computes them (v = delta + dt, stored as -v)
__host__ __device__ __forceinline__ float cl_upper_a(int r, int c, float dt) {
  float v = (r == c) ? 1.f : 0.f;
  if (c == r + NQ) v += dt;
  return -v;
}
hipError_t launch_form_kkt(const Batch &b, const Params &p, const float *xu, const float *ee_ref,
                           const float *xs, const uint32_t *status, float *G, float *C, float *g,
                           float *c, hipStream_t st, float *gq = nullptr, float *Cl = nullptr);
// G (cost blocks) and C (dynamics blocks) of every knot from the compact
// form of launch_form_kkt (gq, Cl), bitwise as launch_form_kkt writes them
hipError_t launch_expand_gc(const Batch &b, const Params &p, const float *gq, const float *Cl, float *G, float *C,
                            hipStream_t st);
// mode: 0..7 line-search alphas (-2^-p, with the initial-state term); 8 the
// current iterate (alpha 0, without it).  Writes out[(p - p0) * total + gk].
hipError_t launch_merit(const Batch &b, const Params &p, const float *xu, const float *dz,
                        const float *ee_ref, const float *xs, const uint32_t *status, int p0, int np,
                        float *out, hipStream_t st);
// out_mode: SCHUR_RIGHT_BLOCK also writes S(b-1,b) = S(b,b-1)^T into the
// right slot of row b-1 (the reference layout; the solvers read the
// transpose of the next row's left block instead), SCHUR_ZERO_UNUSED
// writes zeros into the blocks that are structurally zero (S(0,-1), the
// right block of the last row, Pinv's off-diagonal blocks)
constexpr int SCHUR_ZERO_UNUSED = 1, SCHUR_RIGHT_BLOCK = 2;
hipError_t launch_form_schur(const Batch &b, const float *G, const float *C, const float *g,
                             const float *c, const float *rho_dev, float rho, const uint32_t *status,
                             float *S, float *Pinv, float *gamma, float *Ginv, int out_mode,
                             hipStream_t st);
// The split system (for launch_pcg_split) from the KKT blocks of
// launch_form_kkt, with Q~ = diag(wx + rho) + v v^T (v = (g_q, 0), gq of
// launch_form_kkt) and R~ = diag(wu + rho) inverted in closed form: L^_b
// and W_b into the rows of Ps and gamma^ into ghat with the Schur block
// rows (S, Pinv, gamma and G~^-1 are not formed); each active instance's
// rho into rho_used (optional).
// The dynamics blocks come in the compact form Cl of launch_form_kkt.
hipError_t launch_form_split(const Batch &b, const Params &p, const float *Cl, const float *g, const float *c,
                             const float *gq, const float *rho_dev, float rho, const uint32_t *status, float *Ps,
                             float *ghat, float *rho_used, hipStream_t st);
// dz as launch_compute_dz, with G~^-1 in the closed form of launch_form_split
// (rho_dev: one rho per instance, required, as gq) and the dynamics blocks
// in the compact form Cl of launch_form_kkt
hipError_t launch_compute_dz_lr(const Batch &b, const Params &p, const float *gq, const float *rho_dev,
                                const float *Cl, const float *g, const float *lambda, const uint32_t *status,
                                float *dz, hipStream_t st);
hipError_t launch_compute_dz(const Batch &b, const float *Ginv, const float *C, const float *g,
                             const float *lambda, const uint32_t *status, float *dz, hipStream_t st);
// PCG of every instance.  Outputs per instance: iters (S products), flag (0
// converged, 1 max_iter reached, PCG_FLAG_FAULT abandoned) and eta (the
// |r^T Pinv r| of the last exit test); each may be null.  max_iter ==
// PCG_CAP_BY_N takes the cap of pcg_cap per instance.
// PCG streaming S from memory every iteration (any N <= 512); inst_list
// (n_list entries) restricts the launch to some instances of the batch.
hipError_t launch_pcg_stream(const Batch &b, uint32_t max_n, const float *S, const float *Pinv,
                             const float *gamma, float *lambda, uint32_t max_iter, float tol,
                             const uint32_t *status, const uint32_t *inst_list, uint32_t n_list,
                             uint32_t *iters, uint32_t *flag, float *eta, bool stair, hipStream_t st);
// PCG with S resident on chip (N <= pcg_res_max_knots(nt)); one workgroup of
// nt lanes (512: 128 knots, one per CU; 256: 64 knots, two per CU) per bin
// of instances, bins bin0 .. bin0 + n_bins - 1.  items[bin_off[w] ..
// bin_off[w+1]) = {inst, first group, groups, 0}; items == nullptr: a single
// instance (0) in one bin.
hipError_t launch_pcg_res(const Batch &b, const uint4 *items, const uint32_t *bin_off, uint32_t bin0, uint32_t n_bins,
                          uint32_t nt, const float *S, const float *Pinv, const float *gamma, float *lambda,
                          uint32_t max_iter, float tol, const uint32_t *status, uint32_t *iters,
                          uint32_t *flag, float *eta, bool stair, hipStream_t st);
// PCG of 128 < N <= pcg_team_max_knots(q) on a team of q = 2 or 4
// workgroups (cooperative launches, pcg_team_launches of them); err
// receives 1 if a team wait timed out; done (one word per instance) counts
// the parts that finished each team, and a team short of q gets
// PCG_FLAG_FAULT.  poison (test hook, 0 = off): part poison - 1 of every
// team posts its exchange words under a wrong epoch, so every team ends
// with PCG_FLAG_FAULT.
hipError_t launch_pcg_team(const Batch &b, uint32_t q, const uint32_t *team_inst, uint32_t n_teams,
                           TeamMailbox *mb, uint32_t n_mb, const float *S, const float *Pinv, const float *gamma,
                           float *lambda, uint32_t max_iter, float tol, const uint32_t *status, uint32_t *iters,
                           uint32_t *flag, float *eta, uint32_t *done, uint32_t *err, bool stair, uint32_t poison,
                           hipStream_t st);
// One persistent cooperative launch (one workgroup per CU) for the
// two-workgroup teams team_inst[0 .. n_teams) and the 512-lane bins
// 0 .. n_bins - 1 of the bin table, taken from work queues (ctr): teams
// first, then bins.  Same outputs and fault handling as launch_pcg_team.
// Needs pcg_pool_ok(n_teams, max_iter).
hipError_t launch_pcg_pool(const Batch &b, const uint32_t *team_inst, uint32_t n_teams, TeamMailbox *mb,
                           uint32_t n_mb, PoolCounters *ctr, const uint4 *items, const uint32_t *bin_off,
                           uint32_t n_bins, const float *S, const float *Pinv, const float *gamma, float *lambda,
                           uint32_t max_iter, float tol, const uint32_t *status, uint32_t *iters, uint32_t *flag,
                           float *eta, uint32_t *done, uint32_t *err, bool stair, uint32_t poison, hipStream_t st);
uint32_t pcg_team_max_knots(uint32_t q);
uint32_t pcg_team_max_grid();
uint32_t pcg_team_per_launch(uint32_t q);
uint32_t pcg_team_launches(uint32_t n_teams, uint32_t q);
bool pcg_pool_ok(uint32_t n_teams, uint32_t max_iter);
// Block-Jacobi PCG in split-preconditioned form (k_pcg_split.hip), N <= 256:
// launch_split_prep turns S (left and diagonal blocks) and gamma into the
// split system (rows of Ps: L^_b at 0, W_b at NXX; ghat = W^T gamma), then
// launch_pcg_split solves bins of instances (one 512-lane workgroup per
// bin; items / bin_off as for the resident kernel, groups of
// pcg_split_groups(N)) and writes lambda = W lambda^.  Same outputs as launch_pcg_res.
hipError_t launch_split_prep(const Batch &b, const float *S, const float *gamma, const uint32_t *status, float *Ps,
                             float *ghat, hipStream_t st);
hipError_t launch_pcg_split(const Batch &b, const uint4 *items, const uint32_t *bin_off, uint32_t n_bins,
                            const float *Ps, const float *ghat, float *lambda, uint32_t max_iter, float tol,
                            const uint32_t *status, uint32_t *iters, uint32_t *flag, float *eta, hipStream_t st);
uint32_t pcg_split_max_knots();
uint32_t pcg_split_bin_groups();
uint32_t pcg_split_max_items();
uint32_t pcg_split_groups(uint32_t N);  // 16-lane groups of an instance in a split-form bin
uint32_t pcg_res_groups(uint32_t N);
uint32_t pcg_res_bin_groups(uint32_t nt);  // 16-lane groups per workgroup of nt lanes
uint32_t pcg_res_max_knots(uint32_t nt);
uint32_t pcg_res_max_items();
// a synthetic block-tridiagonal Schur system (diag -2.05 I, off-diagonal I,
// block-Jacobi Pinv, gamma = 1, lambda = 0) for timing launch plans
hipError_t launch_tune_system(const Batch &b, float *S, float *Pinv, float *gamma, float *lambda, hipStream_t st);
// bandwidth probes (k_bandwidth.hip)
hipError_t launch_bw_copy(const float *src, float *dst, size_t n, hipStream_t st);
hipError_t launch_bw_lds(uint32_t *out, uint32_t n_wg, uint32_t reps, hipStream_t st);
hipError_t launch_merit_reduce(const Batch &b, const float *knot_merits, int np, float *out,
                               hipStream_t st);
hipError_t launch_ls_update(const Batch &b, const Params &p, const float *knot_merits,
                            const float *dz, float *xu, InstState s, uint32_t *pcg_iter_log,
                            const uint32_t *pcg_iters_now, const uint32_t *pcg_flag_now,
                            uint32_t it, uint32_t log_stride, hipStream_t st);
hipError_t launch_axpy(uint32_t n, float alpha, const float *x, float *y, hipStream_t st);
// differentiable QP layer (k_adjoint.hip)
hipError_t launch_qp_loss(const Batch &b, const float *xu, const float *ee_ref, float *knot_loss, float *dLdx,
                          float *dLdref, hipStream_t st);
hipError_t launch_pack_neg(const Batch &b, const float *a, float *gz, hipStream_t st);
hipError_t launch_adj_params(const Batch &b, const Params &p, const float *xu, const float *ee_ref,
                             const float *dz, const float *v, float *knot_w, float *gref, hipStream_t st);
hipError_t launch_reduce_knots(const Batch &b, const float *vals, uint32_t S, float *out, hipStream_t st);
hipError_t launch_sub(uint32_t n, const float *a, const float *bb, float *o, hipStream_t st);
hipError_t launch_fill(uint32_t n, float v, float *o, hipStream_t st);
// direct block-tridiagonal solve of the Schur system (k_direct.hip); order:
// instance indices longest first (NULL: instance = list index); work: the
// factors E_b, D_b^-1 and a vector per knot in each knot's S_STRIDE slot;
// refine: iterative-refinement steps with an fp64 residual; do_factor =
// false reuses the factors a previous call left in work for the same S
hipError_t launch_direct_solve(const Batch &b, const uint32_t *order, uint32_t n_list, const float *S,
                               const float *gamma, const uint32_t *status, float *work, float *lambda,
                               uint32_t refine, bool do_factor, uint32_t *iters, uint32_t *flag, hipStream_t st);
// closed-loop harness (k_closed_loop.hip)
hipError_t launch_cl_init(uint32_t N, const float *x0, const float *path, float *win, float *xu, float *xp,
                          uint32_t *ctr, float *states, float *rho, float rho0, hipStream_t st);
hipError_t launch_cl_step(const ClStep &p, const float *path, const float *traj, float *win, float *xu, float *lam,
                          float *xp, uint32_t *ctr, float *states, float *err, const uint32_t *pcg_now,
                          uint32_t *pcg_log, hipStream_t st);
// iterative refinement of the layer's KKT solves (k_refine.hip)
hipError_t launch_kkt_residual(const Batch &b, const float *G, const float *C, const float *g, const float *c,
                               const float *rho_dev, const float *dz, const float *lam, float *rz, float *rc,
                               hipStream_t st);
hipError_t launch_rhs_normalise(const Batch &b, float *rz, float *rc, float *scale, hipStream_t st);
hipError_t launch_form_gamma(const Batch &b, const float *Ginv, const float *C, const float *g, const float *c,
                             float *gamma, hipStream_t st);
hipError_t launch_refine_update(const Batch &b, const float *scale, const float *ddz, const float *dlam, float *dz,
                                float *lam, hipStream_t st);
hipError_t launch_batch_init(const Batch &b, const float *xu, const float *knot_merits, InstState s,
                             hipStream_t st);

}  // namespace mpc
