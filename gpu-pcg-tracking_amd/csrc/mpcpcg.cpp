// mpcpcg.cpp -- host side of the C-ABI (include/mpcpcg.h): handle and
// workspace ownership, stage entry points, the plain single-instance SQP
// driver (drop-in for sqpSolve, toplevel.cuh:24-503) and the batched,
// fully device-side SQP driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"
#include "mpcpcg.h"

using namespace mpc;

#define HIP_TRY(expr)                          \
  do {                                         \
    hipError_t e_ = (expr);                    \
    if (e_ != hipSuccess) return MPC_ERR_HIP;  \
  } while (0)

// Launch plan of the batched PCG (mpc_batch_autotune / mpc_batch_set_plan)
struct PcgPlan {
  uint32_t small_nt = 512;  // workgroup of the bins of N <= 64: 512 or 256 lanes
  uint32_t max_items = 8;   // instances per bin
  bool mixed = true;        // two-workgroup teams and 512-lane bins in one persistent pool launch
  bool teams = true;        // 128 < N <= 512 on teams of two or four workgroups (else streamed)
  bool split = true;        // block Jacobi with every N <= 256: the split-form kernel (k_pcg_split.hip)
};

struct mpc_handle {
  mpc_config cfg{};
  int device = 0;
  // workspace, sized for cfg.max_total_knots (T) and cfg.max_instances (B)
  float *G = nullptr, *Ginv = nullptr, *C = nullptr, *g = nullptr, *c = nullptr;
  float *gq = nullptr;  // [NQ T] tracking-cost gradient per knot (Q = gq gq^T + diag(wx))
  float *Cl = nullptr;  // [CL_STRIDE T] compact dynamics rows of the split path (G, C not written)
  float *S = nullptr, *P = nullptr, *gamma = nullptr, *dz = nullptr;
  float *knot_merits = nullptr;  // [9 T]
  float *merits = nullptr;       // [9 B]
  float *xs = nullptr;           // [14 B]
  float *drho = nullptr, *merit0 = nullptr, *rho1 = nullptr;
  uint32_t *status_ws = nullptr, *sqp_done = nullptr, *pcg_iters = nullptr, *pcg_flag = nullptr;
  float *pcg_eta = nullptr;  // [B] |r^T Pinv r| at each instance's last PCG exit test
  unsigned long long *t0 = nullptr;  // [1] device clock at the start of a batched solve
  double clock_ticks_per_us = 0.0;   // device wall clock rate
  uint32_t *knot_inst = nullptr, *inst_k0 = nullptr;
  // PCG dispatch of a batch: bins of small instances (resident kernel) and
  // the list of instances beyond the teams (streaming kernel)
  uint4 *bin_items = nullptr;
  uint32_t *bin_off = nullptr, *big_list = nullptr;
  uint32_t n_bins = 0, n_bins_small = 0, n_big = 0, max_big_n = 0;
  // split-form PCG: its bins (512-lane workgroups) and gamma^ = W^T gamma
  // (the split system itself lives in P)
  uint4 *split_items = nullptr;
  uint32_t *split_off = nullptr;
  uint32_t n_split_bins = 0;
  bool split_ok = false;  // every horizon of the layout fits the split-form kernel
  float *ghat = nullptr;
  // the batched SQP in split form forms no S: each instance's rho of its last
  // formation is kept so that mpc_batch_direct_resolve can form S again
  float *rho_used = nullptr;
  bool schur_formed = true;  // the last batched solve left S and gamma in the workspace
  bool gc_compact = false;   // its linearisation left G and C in the compact form (gq, Cl)
  // PCG teams: two workgroups for 128 < N <= 256, four for 256 < N <= 512
  uint32_t *team_list = nullptr, *team_one = nullptr, *team_err = nullptr;
  uint32_t *pcg_done = nullptr;       // [B] parts that finished each team solve
  PoolCounters *pool_ctr = nullptr;   // work queues of the PCG pool
  uint32_t *dir_order = nullptr;  // direct solve: instances longest first
  TeamMailbox *mbox = nullptr;
  uint32_t n_teams = 0, n_quads = 0, n_mbox = 256;  // team_list: n_teams pairs, then n_quads quads
  // differentiable QP layer workspace: adjoint multipliers (14 T), per-knot
  // gradient terms (21 T) and per-knot loss terms (T)
  float *adj_mu = nullptr, *adj_knot = nullptr, *loss_knot = nullptr;
  // iterative refinement: residual right-hand side (21 T + 14 T), correction
  // (21 T + 14 T) and per-instance scale (B)
  float *ref_g = nullptr, *ref_c = nullptr, *ref_dz = nullptr, *ref_lam = nullptr, *ref_scale = nullptr;
  // closed loop: reference window (6 max_knots), plant state (16), rho,
  // step counter, SQP status, PCG log of one step; a stream for capture
  float *cl_win = nullptr, *cl_xp = nullptr, *cl_rho = nullptr;
  float *rho_one = nullptr;  // the plain driver's rho on the device (direct mode)
  uint32_t *cl_ctr = nullptr, *cl_status = nullptr, *cl_pcg = nullptr;
  hipStream_t cl_stream = nullptr;
  // the batched solve as a HIP graph (mpc_batch_set_graph): captured once per
  // key and replayed; graph_epoch changes with everything a captured launch
  // bakes in (configuration, weights, batch layout)
  struct GraphKey {
    uint64_t epoch = 0;
    const void *ptr[6]{};
    uint32_t sqp_iters = 0;
    bool operator==(const GraphKey &o) const {
      return epoch == o.epoch && sqp_iters == o.sqp_iters && std::memcmp(ptr, o.ptr, sizeof(ptr)) == 0;
    }
  };
  uint32_t team_launches = 0;  // cooperative team launches per PCG phase
  uint32_t fault_inject = 0;   // mpc_set_fault_injection (test hook)
  bool graph_on = false;
  uint64_t graph_epoch = 1;
  GraphKey graph_key;
  hipGraph_t bgraph = nullptr;
  hipGraphExec_t bexec = nullptr;
  // pinned staging for the host round trips of the plain driver
  float *h_merits = nullptr;
  uint32_t *h_pcg = nullptr;
  float *h_rho = nullptr;
  // batch layout
  std::vector<uint32_t> horizons, k0;
  std::vector<uint32_t> pcg_launch;  // PCG launch of each instance within a phase (mpc_batch_pcg_launch_map)
  PcgPlan plan;
  uint32_t n_inst = 0, total = 0, max_n = 0;
  bool batch_ready = false;
  // per-dimension cost weights (mpc_set_cost_weights; derived from cfg by default)
  float wx[NX]{}, wu[NU]{};
  // timing
  bool timing = false;
  hipEvent_t ev[11]{};  // 0..8: batched SQP phases; 9, 10: closed-loop total
  double phase_us[7]{};
};

// the reference's cost as per-dimension weights: wx = (0 x 7, QD x 7), wu = R x 7
static void weights_from_config(mpc_handle *h) {
  for (int j = 0; j < NX; ++j) h->wx[j] = j < NQ ? 0.f : h->cfg.qd_cost;
  for (int j = 0; j < NU; ++j) h->wu[j] = h->cfg.r_cost;
}

static Params make_params(const mpc_handle *h) {
  const mpc_config &c = h->cfg;
  Params p;
  p.dt = c.dt;
  p.q_ee = c.q_ee;
  for (int j = 0; j < NX; ++j) p.wx[j] = h->wx[j];
  for (int j = 0; j < NU; ++j) p.wu[j] = h->wu[j];
  p.mu = c.mu;
  p.rho_factor = c.rho_factor;
  p.rho_max = c.rho_max;
  p.rho_min = c.rho_min;
  p.rho_reset = c.rho_reset;
  p.pcg_tol = c.pcg_exit_tol;
  p.pcg_max_iter = c.pcg_max_iter;
  p.flags = c.flags;
  p.budget_ticks = c.sqp_max_time_us > 0.f
                       ? (unsigned long long)std::ceil((double)c.sqp_max_time_us * h->clock_ticks_per_us)
                       : 0ull;
  return p;
}

static Batch single(uint32_t N) { return Batch{nullptr, nullptr, 1u, N}; }

// PCG cap of the SQP solves: the configured one, or with
// MPC_FLAG_PCG_CAP_BY_N the reference experiments' cap per horizon
static uint32_t sqp_pcg_cap(const mpc_config &c) {
  return (c.flags & FLAG_PCG_CAP_BY_N) ? PCG_CAP_BY_N : c.pcg_max_iter;
}



// PCG of one instance: resident kernel when its Schur system fits one
// workgroup, a two-workgroup team up to twice that, streaming otherwise.
static hipError_t pcg_single(mpc_handle *h, uint32_t N, const float *S, const float *P, const float *gamma,
                             float *lambda, uint32_t max_iter, float tol, uint32_t *iters, uint32_t *flag,
                             hipStream_t st) {
  const uint32_t flags = h->cfg.flags;
  if (flags & FLAG_DIRECT_SOLVE)
    return launch_direct_solve(single(N), nullptr, 1, S, gamma, nullptr, const_cast<float *>(P), lambda,
                               h->cfg.direct_refine, true, iters, flag, st);
  const bool stair = flags & FLAG_PCG_STAIR;
  if (!(flags & FLAG_PCG_STREAM)) {
    if (N <= pcg_res_max_knots(512))
      return launch_pcg_res(single(N), nullptr, nullptr, 0, 1, 512, S, P, gamma, lambda, max_iter, tol, nullptr,
                            iters, flag, h->pcg_eta, stair, st);
    for (uint32_t q : {2u, 4u})
      if (N <= pcg_team_max_knots(q) && pcg_team_per_launch(q) > 0)
        return launch_pcg_team(single(N), q, h->team_one, 1, h->mbox, h->n_mbox, S, P, gamma, lambda, max_iter,
                               tol, nullptr, iters, flag, h->pcg_eta, h->pcg_done, h->team_err, stair,
                               h->fault_inject, st);
  }
  return launch_pcg_stream(single(N), N, S, P, gamma, lambda, max_iter, tol, nullptr, nullptr, 0, iters, flag,
                           h->pcg_eta, stair, st);
}

static Batch batch_of(const mpc_handle *h) {
  return Batch{h->knot_inst, h->inst_k0, h->n_inst, h->total};
}

static hipError_t pcg_batch(mpc_handle *h, float *d_lambda, const uint32_t *d_status, uint32_t max_iter, float tol,
                            hipEvent_t mid, hipStream_t st, bool prepped = false);

// the batch's block-Jacobi PCG runs in split form (k_pcg_split.hip)
static bool split_active(const mpc_handle *h) {
  return h->plan.split && h->split_ok && !(h->cfg.flags & (FLAG_PCG_STAIR | FLAG_PCG_STREAM | FLAG_DIRECT_SOLVE));
}

// `steps` steps of iterative refinement of K [dz; lam] = [g; c] (constraint
// row 0: dx_0 = 0; c == nullptr: c = 0) for the instances of layout b
// (total knots T) on the workspace's G, C, G~^-1, S and Pinv, with
// solve(lam, tol) the Schur solver of the layout; dz in the xu layout, lam
// 14 per knot, both in place.
template <class Solve>
static hipError_t refine_kkt_with(mpc_handle *h, const Batch &b, uint32_t T, const float *d_rho, const float *g,
                                  const float *c, float *dz, float *lam, uint32_t steps, float rel_tol, Solve solve,
                                  hipStream_t st) {
  const float tol = -(rel_tol > 0.f ? rel_tol : 1e-6f);  // negative: relative PCG exit
  for (uint32_t s = 0; s < steps; ++s) {
    hipError_t e = launch_kkt_residual(b, h->G, h->C, g, c, d_rho, dz, lam, h->ref_g, h->ref_c, st);
    if (e == hipSuccess) e = launch_rhs_normalise(b, h->ref_g, h->ref_c, h->ref_scale, st);
    if (e == hipSuccess) e = launch_form_gamma(b, h->Ginv, h->C, h->ref_g, h->ref_c, h->gamma, st);
    if (e == hipSuccess) e = launch_fill(NX * T, 0.f, h->ref_lam, st);
    if (e == hipSuccess) e = solve(h->ref_lam, tol);
    if (e == hipSuccess) e = launch_compute_dz(b, h->Ginv, h->C, h->ref_g, h->ref_lam, nullptr, h->ref_dz, st);
    if (e == hipSuccess) e = launch_refine_update(b, h->ref_scale, h->ref_dz, h->ref_lam, dz, lam, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// the same for the batch layout
static hipError_t refine_kkt(mpc_handle *h, const float *d_rho, const float *g, const float *c, float *dz,
                             float *lam, uint32_t steps, float rel_tol, hipStream_t st) {
  // direct mode: the corrections reuse the factors of the solve just made on S
  return refine_kkt_with(h, batch_of(h), h->total, d_rho, g, c, dz, lam, steps, rel_tol,
                         [&](float *l, float tol) {
                           if (h->cfg.flags & FLAG_DIRECT_SOLVE)
                             return launch_direct_solve(batch_of(h), h->dir_order, h->n_inst, h->S, h->gamma, nullptr,
                                                        h->P, l, h->cfg.direct_refine, false, nullptr, nullptr, st);
                           return pcg_batch(h, l, nullptr, h->cfg.pcg_max_iter, tol, nullptr, st);
                         },
                         st);
}


static bool valid_numbers(const mpc_config &c) {
  return c.dt > 0 && std::isfinite(c.dt) && c.q_ee >= 0 && c.qd_cost >= 0 && c.r_cost >= 0 &&
         c.mu >= 0 && c.rho_factor > 1 && c.rho_max > 0 && c.rho_min > 0 && c.pcg_exit_tol >= 0 &&
         c.sqp_max_iter <= MPC_MAX_SQP_ITERS && c.direct_refine <= 16 && c.direct_kkt_refine <= 16 &&
         c.sqp_max_time_us >= 0 && std::isfinite(c.sqp_max_time_us);
}

extern "C" {

void mpc_default_config(mpc_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->max_knots = 256;
  cfg->max_instances = 1;
  cfg->max_total_knots = 256;
  cfg->dt = 0.01f;
  cfg->q_ee = 1.f;
  cfg->qd_cost = 1e-2f;
  cfg->r_cost = 1e-4f;
  cfg->mu = 10.f;
  cfg->rho_factor = 1.2f;
  cfg->rho_max = 10.f;
  cfg->rho_min = 1e-3f;
  cfg->rho_reset = 1e-3f;
  cfg->pcg_exit_tol = 1e-5f;
  cfg->pcg_max_iter = 173;
  cfg->sqp_max_iter = 3;
  cfg->flags = 0;
  cfg->direct_refine = 0;
  cfg->direct_kkt_refine = 1;
}

const char *mpc_status_string(int s) {
  switch (s) {
    case MPC_OK: return "ok";
    case MPC_ERR_ARG: return "bad argument";
    case MPC_ERR_HIP: return "HIP runtime error";
    case MPC_ERR_NOT_READY: return "batch layout not set up";
    case MPC_ERR_NONFINITE: return "non-finite value";
    case MPC_ERR_NO_DEVICE: return "no usable GPU";
    default: return "unknown status";
  }
}

static void free_all(mpc_handle *h) {
  void *ptrs[] = {h->G, h->Ginv, h->gq, h->Cl, h->C, h->g, h->c, h->S, h->P, h->gamma, h->dz, h->knot_merits,
                  h->merits, h->xs, h->drho, h->merit0, h->rho1, h->status_ws, h->sqp_done,
                  h->pcg_iters, h->pcg_flag, h->pcg_eta, h->t0, h->knot_inst, h->inst_k0, h->bin_items,
                  h->bin_off, h->big_list, h->team_list, h->team_one, h->team_err, h->mbox, h->dir_order,
                  h->pcg_done, h->pool_ctr, h->split_items, h->split_off, h->ghat, h->rho_used,
                  h->adj_mu, h->adj_knot, h->loss_knot, h->ref_g, h->ref_c, h->ref_dz, h->ref_lam,
                  h->ref_scale, h->cl_win, h->cl_xp, h->cl_rho, h->cl_ctr, h->cl_status, h->cl_pcg,
                  h->rho_one};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (h->h_merits) (void)hipHostFree(h->h_merits);
  if (h->h_pcg) (void)hipHostFree(h->h_pcg);
  if (h->h_rho) (void)hipHostFree(h->h_rho);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->bexec) (void)hipGraphExecDestroy(h->bexec);
  if (h->bgraph) (void)hipGraphDestroy(h->bgraph);
  if (h->cl_stream) (void)hipStreamDestroy(h->cl_stream);
}

int mpc_create(const mpc_config *cfg, mpc_handle **out) {
  if (!cfg || !out) return MPC_ERR_ARG;
  *out = nullptr;
  if (cfg->max_knots < 2 || cfg->max_knots > 512 || cfg->max_instances < 1 ||
      cfg->max_total_knots < cfg->max_knots || !valid_numbers(*cfg))
    return MPC_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPC_ERR_NO_DEVICE;
  mpc_handle *h = new (std::nothrow) mpc_handle();
  if (!h) return MPC_ERR_ARG;
  h->cfg = *cfg;
  weights_from_config(h);
  int clock_khz = 0;
  if (hipGetDevice(&h->device) != hipSuccess ||
      hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, h->device) != hipSuccess ||
      clock_khz <= 0) {
    delete h;
    return MPC_ERR_HIP;
  }
  h->clock_ticks_per_us = clock_khz * 1e-3;
  const size_t T = cfg->max_total_knots, B = cfg->max_instances;
  auto alloc = [&](void **p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess; };
  bool ok = alloc((void **)&h->G, sizeof(float) * G_STRIDE * T) &&
            alloc((void **)&h->Ginv, sizeof(float) * G_STRIDE * T) &&
            alloc((void **)&h->gq, sizeof(float) * NQ * T) &&
            alloc((void **)&h->Cl, sizeof(float) * CL_STRIDE * T) &&
            alloc((void **)&h->C, sizeof(float) * C_STRIDE * T) &&
            alloc((void **)&h->g, sizeof(float) * NXU * T) &&
            alloc((void **)&h->c, sizeof(float) * NX * T) &&
            alloc((void **)&h->S, sizeof(float) * S_STRIDE * T) &&
            alloc((void **)&h->P, sizeof(float) * S_STRIDE * T) &&
            alloc((void **)&h->gamma, sizeof(float) * NX * T) &&
            alloc((void **)&h->dz, sizeof(float) * NXU * T) &&
            alloc((void **)&h->knot_merits, sizeof(float) * 9 * T) &&
            alloc((void **)&h->merits, sizeof(float) * 9 * B) &&
            alloc((void **)&h->xs, sizeof(float) * NX * B) &&
            alloc((void **)&h->drho, sizeof(float) * B) &&
            alloc((void **)&h->merit0, sizeof(float) * B) &&
            alloc((void **)&h->rho1, sizeof(float) * B) &&
            alloc((void **)&h->status_ws, sizeof(uint32_t) * B) &&
            alloc((void **)&h->sqp_done, sizeof(uint32_t) * B) &&
            alloc((void **)&h->pcg_iters, sizeof(uint32_t) * B) &&
            alloc((void **)&h->pcg_flag, sizeof(uint32_t) * B) &&
            alloc((void **)&h->pcg_eta, sizeof(float) * B) &&
            alloc((void **)&h->t0, sizeof(unsigned long long)) &&
            alloc((void **)&h->knot_inst, sizeof(uint32_t) * T) &&
            alloc((void **)&h->inst_k0, sizeof(uint32_t) * (B + 1)) &&
            alloc((void **)&h->bin_items, sizeof(uint4) * B) &&
            alloc((void **)&h->bin_off, sizeof(uint32_t) * (B + 1)) &&
            alloc((void **)&h->split_items, sizeof(uint4) * B) &&
            alloc((void **)&h->split_off, sizeof(uint32_t) * (B + 1)) &&
            alloc((void **)&h->ghat, sizeof(float) * NX * T) &&
            alloc((void **)&h->rho_used, sizeof(float) * B) &&
            alloc((void **)&h->big_list, sizeof(uint32_t) * B) &&
            alloc((void **)&h->team_list, sizeof(uint32_t) * B) &&
            alloc((void **)&h->dir_order, sizeof(uint32_t) * B) &&
            alloc((void **)&h->team_one, sizeof(uint32_t)) &&
            alloc((void **)&h->team_err, sizeof(uint32_t)) &&
            alloc((void **)&h->pcg_done, sizeof(uint32_t) * B) &&
            alloc((void **)&h->pool_ctr, sizeof(PoolCounters)) &&
            alloc((void **)&h->mbox, sizeof(TeamMailbox) * h->n_mbox) &&
            alloc((void **)&h->adj_mu, sizeof(float) * NX * T) &&
            alloc((void **)&h->adj_knot, sizeof(float) * NXU * T) &&
            alloc((void **)&h->loss_knot, sizeof(float) * T) &&
            alloc((void **)&h->ref_g, sizeof(float) * NXU * T) &&
            alloc((void **)&h->ref_c, sizeof(float) * NX * T) &&
            alloc((void **)&h->ref_dz, sizeof(float) * NXU * T) &&
            alloc((void **)&h->ref_lam, sizeof(float) * NX * T) &&
            alloc((void **)&h->ref_scale, sizeof(float) * B) &&
            alloc((void **)&h->cl_win, sizeof(float) * 6 * cfg->max_knots) &&
            alloc((void **)&h->cl_xp, sizeof(float) * 16) && alloc((void **)&h->cl_rho, sizeof(float) * 4) &&
            alloc((void **)&h->rho_one, sizeof(float) * 4) &&
            alloc((void **)&h->cl_ctr, sizeof(uint32_t) * 4) &&
            alloc((void **)&h->cl_status, sizeof(uint32_t) * 4) &&
            alloc((void **)&h->cl_pcg, sizeof(uint32_t) * MPC_MAX_SQP_ITERS) &&
            hipStreamCreate(&h->cl_stream) == hipSuccess &&
            hipHostMalloc((void **)&h->h_merits, sizeof(float) * 16) == hipSuccess &&
            hipHostMalloc((void **)&h->h_pcg, sizeof(uint32_t) * 4) == hipSuccess &&
            hipHostMalloc((void **)&h->h_rho, sizeof(float) * 4) == hipSuccess;
  for (auto &e : h->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && hipMemset(h->team_one, 0, sizeof(uint32_t)) == hipSuccess &&
       hipMemset(h->team_err, 0, sizeof(uint32_t)) == hipSuccess &&
       hipMemset(h->mbox, 0, sizeof(TeamMailbox) * h->n_mbox) == hipSuccess;
  if (!ok) {
    free_all(h);
    delete h;
    return MPC_ERR_HIP;
  }
  *out = h;
  return MPC_OK;
}

int mpc_destroy(mpc_handle *h) {
  if (!h) return MPC_ERR_ARG;
  (void)hipDeviceSynchronize();
  free_all(h);
  delete h;
  return MPC_OK;
}

int mpc_get_config(const mpc_handle *h, mpc_config *cfg) {
  if (!h || !cfg) return MPC_ERR_ARG;
  *cfg = h->cfg;
  return MPC_OK;
}

int mpc_set_config(mpc_handle *h, const mpc_config *cfg) {
  if (!h || !cfg || !valid_numbers(*cfg)) return MPC_ERR_ARG;
  mpc_config c = *cfg;
  c.max_knots = h->cfg.max_knots;
  c.max_instances = h->cfg.max_instances;
  c.max_total_knots = h->cfg.max_total_knots;
  h->cfg = c;
  weights_from_config(h);
  ++h->graph_epoch;
  return MPC_OK;
}

int mpc_set_cost_weights(mpc_handle *h, const float *wx14, const float *wu7) {
  if (!h || !wx14 || !wu7) return MPC_ERR_ARG;
  for (int j = 0; j < NX; ++j)
    if (!(wx14[j] >= 0.f) || !std::isfinite(wx14[j])) return MPC_ERR_ARG;
  for (int j = 0; j < NU; ++j)
    if (!(wu7[j] >= 0.f) || !std::isfinite(wu7[j])) return MPC_ERR_ARG;
  std::memcpy(h->wx, wx14, sizeof(h->wx));
  std::memcpy(h->wu, wu7, sizeof(h->wu));
  ++h->graph_epoch;
  return MPC_OK;
}

int mpc_set_fault_injection(mpc_handle *h, uint32_t part_plus_one) {
  if (!h || part_plus_one > 4) return MPC_ERR_ARG;
  h->fault_inject = part_plus_one;
  ++h->graph_epoch;  // captured team launches carry the value
  return MPC_OK;
}

int mpc_get_cost_weights(mpc_handle *h, float *wx14, float *wu7) {
  if (!h || !wx14 || !wu7) return MPC_ERR_ARG;
  std::memcpy(wx14, h->wx, sizeof(h->wx));
  std::memcpy(wu7, h->wu, sizeof(h->wu));
  return MPC_OK;
}

// ---- robot model ------------------------------------------------------------

int mpc_fk(mpc_handle *h, uint32_t n, const float *d_q, uint32_t q_stride, float *d_ee,
           mpc_stream_t stream) {
  if (!h || (n && (!d_q || !d_ee)) || q_stride < NQ) return MPC_ERR_ARG;
  HIP_TRY(launch_fk(n, d_q, q_stride, d_ee, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_inverse_dynamics(mpc_handle *h, uint32_t n, const float *d_qqdqdd, float *d_tau, mpc_stream_t stream) {
  if (!h || (n && (!d_qqdqdd || !d_tau))) return MPC_ERR_ARG;
  HIP_TRY(launch_inverse_dynamics(n, d_qqdqdd, d_tau, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_bw_copy(const float *d_src, float *d_dst, size_t n, mpc_stream_t stream) {
  if (n && (!d_src || !d_dst || n % 4 != 0 || (reinterpret_cast<uintptr_t>(d_src) & 15u) ||
            (reinterpret_cast<uintptr_t>(d_dst) & 15u)))
    return MPC_ERR_ARG;
  HIP_TRY(launch_bw_copy(d_src, d_dst, n, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_bw_lds(uint32_t *d_out, uint32_t n_wg, uint32_t reps, mpc_stream_t stream) {
  if (n_wg && !d_out) return MPC_ERR_ARG;
  HIP_TRY(launch_bw_lds(d_out, n_wg, reps, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_dynamics_gradient(mpc_handle *h, uint32_t n, const float *d_qqdu, float *d_qdd,
                          float *d_df_du, mpc_stream_t stream) {
  if (!h || (n && (!d_qqdu || !d_qdd || !d_df_du))) return MPC_ERR_ARG;
  HIP_TRY(launch_dyn_grad(n, d_qqdu, d_qdd, d_df_du, (hipStream_t)stream));
  return MPC_OK;
}

// ---- stages -----------------------------------------------------------------

static bool n_ok(const mpc_handle *h, uint32_t N) {
  return N >= 2 && N <= h->cfg.max_knots && N <= h->cfg.max_total_knots;
}

int mpc_form_kkt(mpc_handle *h, uint32_t N, const float *d_xu, const float *d_ee_ref,
                 const float *d_xs, float *d_G, float *d_C, float *d_g, float *d_c,
                 mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_xu || !d_ee_ref || !d_xs || !d_G || !d_C || !d_g || !d_c)
    return MPC_ERR_ARG;
  HIP_TRY(launch_form_kkt(single(N), make_params(h), d_xu, d_ee_ref, d_xs, nullptr, d_G, d_C,
                          d_g, d_c, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_form_schur(mpc_handle *h, uint32_t N, float *d_G, const float *d_C, const float *d_g,
                   const float *d_c, float *d_S, float *d_Pinv, float *d_gamma, float rho,
                   mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_G || !d_C || !d_g || !d_c || !d_S || !d_Pinv || !d_gamma ||
      !(rho >= 0))
    return MPC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(launch_form_schur(single(N), d_G, d_C, d_g, d_c, nullptr, rho, nullptr, d_S, d_Pinv, d_gamma,
                            h->Ginv, SCHUR_ZERO_UNUSED | SCHUR_RIGHT_BLOCK, st));
  HIP_TRY(hipMemcpyAsync(d_G, h->Ginv, sizeof(float) * (G_STRIDE * (size_t)N - 49),
                         hipMemcpyDeviceToDevice, st));
  return MPC_OK;
}

int mpc_pcg_solve(mpc_handle *h, uint32_t N, const float *d_S, const float *d_Pinv,
                  const float *d_gamma, float *d_lambda, uint32_t max_iter, float exit_tol,
                  uint32_t *d_iters, uint32_t *d_max_iter_exit, mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_S || !d_Pinv || !d_gamma || !d_lambda || !(exit_tol >= 0))
    return MPC_ERR_ARG;
  HIP_TRY(pcg_single(h, N, d_S, d_Pinv, d_gamma, d_lambda, max_iter, exit_tol, d_iters, d_max_iter_exit,
                     (hipStream_t)stream));
  return MPC_OK;
}

int mpc_direct_solve(mpc_handle *h, uint32_t N, const float *d_S, const float *d_gamma, float *d_lambda,
                     uint32_t refine_steps, float *d_work, mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_S || !d_gamma || !d_lambda || !d_work || refine_steps > 16) return MPC_ERR_ARG;
  HIP_TRY(launch_direct_solve(single(N), nullptr, 1, d_S, d_gamma, nullptr, d_work, d_lambda, refine_steps, true,
                              nullptr, nullptr, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_compute_dz(mpc_handle *h, uint32_t N, const float *d_Ginv, const float *d_C,
                   const float *d_g, const float *d_lambda, float *d_dz, mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_Ginv || !d_C || !d_g || !d_lambda || !d_dz) return MPC_ERR_ARG;
  HIP_TRY(launch_compute_dz(single(N), d_Ginv, d_C, d_g, d_lambda, nullptr, d_dz, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_compute_merit(mpc_handle *h, uint32_t N, const float *d_xu, const float *d_ee_ref,
                      float *d_merit_out, mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_xu || !d_ee_ref || !d_merit_out) return MPC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const Batch b = single(N);
  HIP_TRY(launch_merit(b, make_params(h), d_xu, nullptr, d_ee_ref, h->xs, nullptr, 8, 1,
                       h->knot_merits, st));
  HIP_TRY(launch_merit_reduce(b, h->knot_merits, 1, d_merit_out, st));
  return MPC_OK;
}

int mpc_line_search_merits(mpc_handle *h, uint32_t N, const float *d_xs, const float *d_xu,
                           const float *d_ee_ref, const float *d_dz, float *d_merits_out,
                           mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_xs || !d_xu || !d_ee_ref || !d_dz || !d_merits_out) return MPC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const Batch b = single(N);
  HIP_TRY(launch_merit(b, make_params(h), d_xu, d_dz, d_ee_ref, d_xs, nullptr, 0, 8,
                       h->knot_merits, st));
  HIP_TRY(launch_merit_reduce(b, h->knot_merits, 8, d_merits_out, st));
  return MPC_OK;
}

// ---- plain SQP (one instance, host decisions as in the reference) -----------

int mpc_sqp_solve(mpc_handle *h, uint32_t N, const float *d_ee_ref, float *d_lambda, float *d_xu,
                  float *rho, float rho_reset, mpc_sqp_stats *stats, mpc_stream_t stream) {
  if (!h || !n_ok(h, N) || !d_ee_ref || !d_lambda || !d_xu || !rho || !(*rho > 0)) return MPC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const mpc_config &cfg = h->cfg;
  const Params prm = make_params(h);
  const Batch b = single(N);
  mpc_sqp_stats local;
  mpc_sqp_stats &s = stats ? *stats : local;
  std::memset(&s, 0, sizeof(s));
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t len = NXU * N - NU;
  // sqpTimecheck (toplevel.cuh:223-231): the host clock since the start
  // against SQP_MAX_TIME_US
  auto over_budget = [&]() {
    return cfg.sqp_max_time_us > 0.f &&
           std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() >
               (double)cfg.sqp_max_time_us;
  };

  HIP_TRY(hipMemcpyAsync(h->xs, d_xu, sizeof(float) * NX, hipMemcpyDeviceToDevice, st));
  HIP_TRY(launch_merit(b, prm, d_xu, nullptr, d_ee_ref, h->xs, nullptr, 8, 1, h->knot_merits, st));
  HIP_TRY(launch_merit_reduce(b, h->knot_merits, 1, h->merits, st));
  HIP_TRY(hipMemcpyAsync(h->h_merits, h->merits, sizeof(float), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  float merit0 = h->h_merits[0];
  if (!std::isfinite(merit0)) {  // nothing a line search could improve on
    s.sqp_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return MPC_ERR_NONFINITE;
  }
  float drho = 1.f;
  const float rf = cfg.rho_factor;
  uint32_t sqp_iter = 0;
  bool not_aborted = true;
  bool time_exit = false;
  for (uint32_t it = 0; it < cfg.sqp_max_iter; ++it) {
    HIP_TRY(launch_form_kkt(b, prm, d_xu, d_ee_ref, h->xs, nullptr, h->G, h->C, h->g, h->c, st));
    h->gc_compact = false;  // G and C written in full
    if ((time_exit = over_budget())) break;  // toplevel.cuh:271
    HIP_TRY(launch_form_schur(b, h->G, h->C, h->g, h->c, nullptr, *rho, nullptr, h->S, h->P, h->gamma,
                              h->Ginv, 0, st));
    if ((time_exit = over_budget())) break;  // toplevel.cuh:283
    HIP_TRY(hipEventRecord(h->ev[0], st));
    HIP_TRY(pcg_single(h, N, h->S, h->P, h->gamma, d_lambda, sqp_pcg_cap(cfg), cfg.pcg_exit_tol, h->pcg_iters,
                       h->pcg_flag, st));
    HIP_TRY(hipEventRecord(h->ev[1], st));
    HIP_TRY(hipMemcpyAsync(h->h_pcg, h->pcg_iters, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(h->h_pcg + 1, h->pcg_flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(launch_compute_dz(b, h->Ginv, h->C, h->g, d_lambda, nullptr, h->dz, st));
    if (cfg.flags & FLAG_DIRECT_SOLVE) {
      h->h_rho[3] = *rho;
      HIP_TRY(hipMemcpyAsync(h->rho_one, h->h_rho + 3, sizeof(float), hipMemcpyHostToDevice, st));
      HIP_TRY(refine_kkt_with(h, b, N, h->rho_one, h->g, h->c, h->dz, d_lambda, h->cfg.direct_kkt_refine, 0.f,
                              [&](float *l, float) {
                                return launch_direct_solve(b, nullptr, 1, h->S, h->gamma, nullptr, h->P, l,
                                                           h->cfg.direct_refine, false, nullptr, nullptr, st);
                              },
                              st));
    }
    HIP_TRY(launch_merit(b, prm, d_xu, h->dz, d_ee_ref, h->xs, nullptr, 0, 8, h->knot_merits, st));
    HIP_TRY(launch_merit_reduce(b, h->knot_merits, 8, h->merits, st));
    HIP_TRY(hipMemcpyAsync(h->h_merits, h->merits, sizeof(float) * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    if (h->h_pcg[1] == PCG_FLAG_FAULT) return MPC_ERR_PCG_FAULT;
    s.pcg_iters[it] = h->h_pcg[0];
    s.pcg_maxiter[it] = h->h_pcg[1];
    s.linsys_us[it] = 1e3 * ms;
    if ((time_exit = over_budget())) break;  // toplevel.cuh:381, before the line-search decision
    int step = -1;
    float minm = merit0;
    for (int i = 0; i < 8; ++i)
      if (h->h_merits[i] < minm) {
        minm = h->h_merits[i];
        step = i;
      }
    s.ls_step[it] = step;
    if (step < 0 && !std::isfinite(h->h_merits[0])) {  // the step itself is not finite: stop, not taken
      s.sqp_iters = sqp_iter + 1;
      s.not_aborted = 1u;
      s.sqp_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
      return MPC_ERR_NONFINITE;
    }
    if (step < 0) {
      drho = std::fmax(drho * rf, rf);
      *rho = std::fmax(*rho * drho, cfg.rho_min);
      sqp_iter++;
      if (*rho > cfg.rho_max) {
        not_aborted = false;
        *rho = rho_reset;
        break;
      }
      continue;
    }
    const float alpha = -1.f / (float)(1u << step);
    drho = std::fmin(drho / rf, 1.f / rf);
    *rho = std::fmax(*rho * drho, cfg.rho_min);
    HIP_TRY(launch_axpy(len, alpha, h->dz, d_xu, st));
    sqp_iter++;
    merit0 = minm;
    if ((time_exit = over_budget())) break;  // toplevel.cuh:437
  }
  HIP_TRY(hipStreamSynchronize(st));
  s.sqp_iters = sqp_iter;
  s.not_aborted = not_aborted ? 1u : 0u;
  s.time_exit = time_exit ? 1u : 0u;
  s.sqp_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  return MPC_OK;
}

// PCG launch plan of the batch layout under h->plan.  Instances with
// N <= 128 are packed first-fit decreasing into bins of the resident kernel
// (whole waves per instance, at most plan.max_items per bin): workgroups of
// 512 lanes, or for N <= 64 of 256 lanes when plan.small_nt is 256 (listed
// after the 512-lane bins in the same tables).  128 < N <= 256 run on
// two-workgroup teams (plan.teams) and longer ones, or all of them without
// teams, on the streaming kernel.  Also records the launch of each instance
// within a PCG phase (mpc_batch_pcg_launch_map).
static int build_pcg_plan(mpc_handle *h) {
  const PcgPlan &pl = h->plan;
  const uint32_t n_inst = (uint32_t)h->horizons.size();
  const uint32_t *horizons = h->horizons.data();
  const uint32_t max_items = pl.max_items < pcg_res_max_items() ? pl.max_items : pcg_res_max_items();
  const uint32_t small_max = pl.small_nt == 256 ? pcg_res_max_knots(256) : 0u;
  std::vector<uint32_t> order, order_s, big, teams, quads;
  uint32_t max_big = 0;
  const bool pair_ok = pl.teams && pcg_team_per_launch(2) > 0, quad_ok = pl.teams && pcg_team_per_launch(4) > 0;
  for (uint32_t i = 0; i < n_inst; ++i) {
    if (horizons[i] <= small_max) order_s.push_back(i);
    else if (horizons[i] <= pcg_res_max_knots(512)) order.push_back(i);
    else if (pair_ok && horizons[i] <= pcg_team_max_knots(2)) teams.push_back(i);
    else if (quad_ok && horizons[i] <= pcg_team_max_knots(4)) quads.push_back(i);
    else {
      big.push_back(i);
      max_big = horizons[i] > max_big ? horizons[i] : max_big;
    }
  }
  struct BinB {
    uint32_t used = 0;
    std::vector<uint32_t> inst;
  };
  auto pack = [&](std::vector<uint32_t> &ord, uint32_t cap_groups, uint32_t (*groups)(uint32_t)) {
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) { return horizons[a] > horizons[c]; });
    std::vector<BinB> bins;
    for (uint32_t i : ord) {
      const uint32_t need = groups(horizons[i]);
      BinB *dst = nullptr;
      for (auto &bb : bins)
        if (bb.used + need <= cap_groups && bb.inst.size() < max_items) {
          dst = &bb;
          break;
        }
      if (!dst) {
        bins.emplace_back();
        dst = &bins.back();
      }
      dst->inst.push_back(i);
      dst->used += need;
    }
    return bins;
  };
  // split form: every instance in bins of the 512-lane kernel
  h->split_ok = n_inst > 0;
  for (uint32_t i = 0; i < n_inst; ++i) h->split_ok = h->split_ok && horizons[i] <= pcg_split_max_knots();
  h->n_split_bins = 0;
  if (h->split_ok) {
    std::vector<uint32_t> all(n_inst);
    for (uint32_t i = 0; i < n_inst; ++i) all[i] = i;
    const std::vector<BinB> sb = pack(all, pcg_split_bin_groups(), pcg_split_groups);
    std::vector<uint4> sitems;
    std::vector<uint32_t> soff(1, 0);
    for (auto &bb : sb) {
      uint32_t g0 = 0;
      for (uint32_t i : bb.inst) {
        const uint32_t ng = pcg_split_groups(horizons[i]);
        sitems.push_back(make_uint4(i, g0, ng, 0u));
        g0 += ng;
      }
      soff.push_back((uint32_t)sitems.size());
    }
    HIP_TRY(hipMemcpy(h->split_items, sitems.data(), sizeof(uint4) * sitems.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->split_off, soff.data(), sizeof(uint32_t) * soff.size(), hipMemcpyHostToDevice));
    h->n_split_bins = (uint32_t)sb.size();
  }
  std::vector<BinB> bins = pack(order, pcg_res_bin_groups(512), pcg_res_groups);
  const std::vector<BinB> bins_s = pack(order_s, pcg_res_bin_groups(256), pcg_res_groups);
  const uint32_t n512 = (uint32_t)bins.size();
  bins.insert(bins.end(), bins_s.begin(), bins_s.end());
  std::vector<uint4> items;
  std::vector<uint32_t> boff(1, 0);
  for (auto &bb : bins) {
    uint32_t g0 = 0;
    for (uint32_t i : bb.inst) {
      const uint32_t ng = pcg_res_groups(horizons[i]);
      items.push_back(make_uint4(i, g0, ng, 0u));
      g0 += ng;
    }
    boff.push_back((uint32_t)items.size());
  }
  if (!items.empty())
    HIP_TRY(hipMemcpy(h->bin_items, items.data(), sizeof(uint4) * items.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->bin_off, boff.data(), sizeof(uint32_t) * boff.size(), hipMemcpyHostToDevice));
  if (!big.empty())
    HIP_TRY(hipMemcpy(h->big_list, big.data(), sizeof(uint32_t) * big.size(), hipMemcpyHostToDevice));
  std::vector<uint32_t> tl(teams);
  tl.insert(tl.end(), quads.begin(), quads.end());
  if (!tl.empty()) HIP_TRY(hipMemcpy(h->team_list, tl.data(), sizeof(uint32_t) * tl.size(), hipMemcpyHostToDevice));
  // launch of each instance within a PCG phase: four-workgroup team
  // launches 0 .. nq-1; then with the pool (plan.mixed) one launch nq for
  // the two-workgroup teams and the 512-lane bins, else two-workgroup team
  // launches nq .. nq+np-1 and the 512-lane bins after the streaming
  // launch; the streaming launch; the 256-lane bins last
  const uint32_t nq = pcg_team_launches((uint32_t)quads.size(), 4);
  const bool pool = pl.mixed && pcg_pool_ok((uint32_t)teams.size(), h->cfg.pcg_max_iter) &&
                    (!teams.empty() || n512 > 0);
  const uint32_t np = pool ? 1u : pcg_team_launches((uint32_t)teams.size(), 2);
  h->pcg_launch.assign(n_inst, 0u);
  for (size_t t = 0; t < quads.size(); ++t) h->pcg_launch[quads[t]] = (uint32_t)(t / pcg_team_per_launch(4));
  for (size_t t = 0; t < teams.size(); ++t)
    h->pcg_launch[teams[t]] = nq + (pool ? 0u : (uint32_t)(t / pcg_team_per_launch(2)));
  const uint32_t nl = nq + np;
  uint32_t next = nl;
  if (!big.empty())
    for (uint32_t i : big) h->pcg_launch[i] = next;
  next += big.empty() ? 0u : 1u;
  const uint32_t res512 = pool ? nq : next;
  next += (!pool && n512 > 0) ? 1u : 0u;
  for (size_t w = 0; w < bins.size(); ++w)
    for (uint32_t i : bins[w].inst) h->pcg_launch[i] = w < n512 ? res512 : next;
  h->n_teams = (uint32_t)teams.size();
  h->n_quads = (uint32_t)quads.size();
  h->team_launches = nl;
  h->n_bins = n512;
  h->n_bins_small = (uint32_t)bins_s.size();
  h->n_big = (uint32_t)big.size();
  h->max_big_n = max_big;
  ++h->graph_epoch;
  return MPC_OK;
}

// ---- batched SQP ------------------------------------------------------------

int mpc_batch_setup(mpc_handle *h, uint32_t n_inst, const uint32_t *horizons) {
  if (!h || !horizons || n_inst == 0 || n_inst > h->cfg.max_instances) return MPC_ERR_ARG;
  ++h->graph_epoch;
  std::vector<uint32_t> k0(n_inst + 1, 0);
  uint32_t max_n = 0;
  for (uint32_t i = 0; i < n_inst; ++i) {
    if (horizons[i] < 2 || horizons[i] > h->cfg.max_knots) return MPC_ERR_ARG;
    k0[i + 1] = k0[i] + horizons[i];
    if (k0[i + 1] > h->cfg.max_total_knots) return MPC_ERR_ARG;
    max_n = horizons[i] > max_n ? horizons[i] : max_n;
  }
  const uint32_t T = k0[n_inst];
  h->batch_ready = false;
  std::vector<uint32_t> ki(T);
  for (uint32_t i = 0; i < n_inst; ++i)
    for (uint32_t k = k0[i]; k < k0[i + 1]; ++k) ki[k] = i;
  HIP_TRY(hipMemcpy(h->knot_inst, ki.data(), sizeof(uint32_t) * T, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->inst_k0, k0.data(), sizeof(uint32_t) * (n_inst + 1), hipMemcpyHostToDevice));
  {
    std::vector<uint32_t> all(n_inst);
    for (uint32_t i = 0; i < n_inst; ++i) all[i] = i;
    std::stable_sort(all.begin(), all.end(), [&](uint32_t a, uint32_t c) { return horizons[a] > horizons[c]; });
    HIP_TRY(hipMemcpy(h->dir_order, all.data(), sizeof(uint32_t) * n_inst, hipMemcpyHostToDevice));
  }
  h->horizons.assign(horizons, horizons + n_inst);
  h->k0 = k0;
  h->n_inst = n_inst;
  h->total = T;
  h->max_n = max_n;
  const int rc = build_pcg_plan(h);
  if (rc != MPC_OK) return rc;
  h->batch_ready = true;
  return MPC_OK;
}

// ---- differentiable QP layer --------------------------------------------------

int mpc_batch_qp_layer(mpc_handle *h, const float *d_xu, const float *d_ee_ref, const float *d_rho,
                       float *d_lambda, float pcg_tol, uint32_t refine_steps, float refine_tol, float *d_dz,
                       float *d_xstar, mpc_stream_t stream) {
  if (!h || !d_xu || !d_ee_ref || !d_rho || !d_lambda || !d_dz || !d_xstar || !(refine_tol == refine_tol))
    return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  hipStream_t st = (hipStream_t)stream;
  const Batch b = batch_of(h);
  const Params prm = make_params(h);
  const float tol = pcg_tol > 0.f ? pcg_tol : h->cfg.pcg_exit_tol;
  // c_0 (x_0 - x_s) is not used by the Schur step, so x_s is left as it is
  HIP_TRY(launch_form_kkt(b, prm, d_xu, d_ee_ref, h->xs, nullptr, h->G, h->C, h->g, h->c, st));
  h->gc_compact = false;  // G and C written in full
  HIP_TRY(launch_form_schur(b, h->G, h->C, h->g, h->c, d_rho, 0.f, nullptr, h->S, h->P, h->gamma, h->Ginv, 0, st));
  HIP_TRY(pcg_batch(h, d_lambda, nullptr, h->cfg.pcg_max_iter, tol, nullptr, st));
  HIP_TRY(launch_compute_dz(b, h->Ginv, h->C, h->g, d_lambda, nullptr, d_dz, st));
  HIP_TRY(refine_kkt(h, d_rho, h->g, h->c, d_dz, d_lambda, refine_steps, refine_tol, st));
  HIP_TRY(launch_sub(NXU * h->total - NU * h->n_inst, d_xu, d_dz, d_xstar, st));
  return MPC_OK;
}

int mpc_batch_qp_loss(mpc_handle *h, const float *d_xu, const float *d_ee_ref, float *d_loss, float *d_dLdx,
                      float *d_dLdref, mpc_stream_t stream) {
  if (!h || !d_xu || !d_ee_ref || !d_loss) return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  hipStream_t st = (hipStream_t)stream;
  const Batch b = batch_of(h);
  HIP_TRY(launch_qp_loss(b, d_xu, d_ee_ref, h->loss_knot, d_dLdx, d_dLdref, st));
  HIP_TRY(launch_reduce_knots(b, h->loss_knot, 1, d_loss, st));
  return MPC_OK;
}

int mpc_batch_qp_adjoint(mpc_handle *h, const float *d_xu, const float *d_ee_ref, const float *d_rho,
                         const float *d_dz, const float *d_a, float pcg_tol, uint32_t refine_steps,
                         float refine_tol, float *d_grad_w, float *d_grad_ref, mpc_stream_t stream) {
  if (!h || !d_xu || !d_ee_ref || !d_rho || !d_dz || !d_a || !d_grad_w || !d_grad_ref ||
      !(refine_tol == refine_tol))
    return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  hipStream_t st = (hipStream_t)stream;
  const Batch b = batch_of(h);
  const Params prm = make_params(h);
  const float tol = pcg_tol > 0.f ? pcg_tol : h->cfg.pcg_exit_tol;
  // recompute the linearisation at xu (nothing of the forward pass is kept
  // but dz), then the Schur system with g := -a and c := 0
  HIP_TRY(launch_form_kkt(b, prm, d_xu, d_ee_ref, h->xs, nullptr, h->G, h->C, h->g, h->c, st));
  h->gc_compact = false;  // G and C written in full
  HIP_TRY(launch_pack_neg(b, d_a, h->g, st));
  HIP_TRY(launch_form_schur(b, h->G, h->C, h->g, nullptr, d_rho, 0.f, nullptr, h->S, h->P, h->gamma, h->Ginv, 0,
                            st));
  HIP_TRY(launch_fill(NX * h->total, 0.f, h->adj_mu, st));
  HIP_TRY(pcg_batch(h, h->adj_mu, nullptr, h->cfg.pcg_max_iter, tol, nullptr, st));
  HIP_TRY(launch_compute_dz(b, h->Ginv, h->C, h->g, h->adj_mu, nullptr, h->dz, st));  // v_z
  HIP_TRY(refine_kkt(h, d_rho, h->g, nullptr, h->dz, h->adj_mu, refine_steps, refine_tol, st));
  HIP_TRY(launch_adj_params(b, prm, d_xu, d_ee_ref, d_dz, h->dz, h->adj_knot, d_grad_ref, st));
  HIP_TRY(launch_reduce_knots(b, h->adj_knot, NXU, d_grad_w, st));
  return MPC_OK;
}

int mpc_batch_pcg_layout(const mpc_handle *h, uint32_t *out4) {
  if (!h || !out4) return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  const bool split = split_active(h);
  out4[0] = split ? h->n_split_bins : h->n_bins + h->n_bins_small;
  out4[1] = split ? 0u : h->n_teams + h->n_quads;
  out4[2] = split ? 0u : h->n_big;
  out4[3] = split || (h->cfg.flags & (FLAG_PCG_STREAM | FLAG_DIRECT_SOLVE)) ? 0u : h->team_launches;
  return MPC_OK;
}

static void plan_to_c(const PcgPlan &p, float us, mpc_pcg_plan *o) {
  o->small_nt = p.small_nt;
  o->max_items = p.max_items;
  o->mixed = p.mixed ? 1u : 0u;
  o->teams = p.teams ? 1u : 0u;
  o->us = us;
  o->split = p.split ? 1u : 0u;
}

int mpc_batch_get_plan(const mpc_handle *h, mpc_pcg_plan *plan) {
  if (!h || !plan) return MPC_ERR_ARG;
  plan_to_c(h->plan, 0.f, plan);
  return MPC_OK;
}

int mpc_batch_set_plan(mpc_handle *h, const mpc_pcg_plan *plan) {
  if (!h || !plan || (plan->small_nt != 256 && plan->small_nt != 512) || plan->max_items < 1 ||
      plan->max_items > pcg_res_max_items() || plan->mixed > 1 || plan->teams > 1 || plan->split > 1)
    return MPC_ERR_ARG;
  h->plan.small_nt = plan->small_nt;
  h->plan.max_items = plan->max_items;
  h->plan.mixed = plan->mixed != 0;
  h->plan.teams = plan->teams != 0;
  h->plan.split = plan->split != 0;
  return h->batch_ready ? build_pcg_plan(h) : MPC_OK;
}

// Times every launch plan that differs for this layout on the synthetic
// system of launch_tune_system (`iters` PCG iterations everywhere, tol 0,
// the best of `reps` passes after one untimed pass) and keeps the fastest.
int mpc_batch_autotune(mpc_handle *h, uint32_t iters, uint32_t reps, mpc_pcg_plan *results, uint32_t max_results,
                       uint32_t *n_results, mpc_stream_t stream) {
  if (!h || iters == 0 || iters > 100000 || reps == 0 || reps > 100 || (max_results && !results)) return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  if (h->cfg.flags & (FLAG_PCG_STREAM | FLAG_DIRECT_SOLVE)) return MPC_ERR_ARG;  // no launch plan in these modes
  hipStream_t st = (hipStream_t)stream;
  const Batch b = batch_of(h);
  bool has_small = false, has_bins = false, has_team = false, has_pair = false;
  for (uint32_t n : h->horizons) {
    has_small |= n <= pcg_res_max_knots(256);
    has_bins |= n <= pcg_res_max_knots(512);
    has_team |= n > pcg_res_max_knots(512) && n <= pcg_team_max_knots(4);
    has_pair |= n > pcg_res_max_knots(512) && n <= pcg_team_max_knots(2);
  }
  has_team = has_team && pcg_team_per_launch(4) > 0;
  std::vector<PcgPlan> cands;
  const bool split_cand = h->split_ok && !(h->cfg.flags & FLAG_PCG_STAIR);
  if (split_cand)
    for (uint32_t mi : {8u, 4u, 2u}) {
      PcgPlan c;
      c.max_items = mi;
      c.split = true;
      cands.push_back(c);
    }
  for (uint32_t nt : {512u, 256u}) {
    if (nt == 256 && !has_small) continue;
    for (uint32_t mi : {8u, 4u, 2u}) {
      if (mi != 8 && !has_bins) continue;
      for (int teams = 1; teams >= 0; --teams) {
        if (!teams && !has_team) continue;
        for (int mixed = 1; mixed >= 0; --mixed) {  // the pool, or static launches
          if (!mixed && !((teams && has_pair) || has_bins)) continue;
          PcgPlan c;
          c.small_nt = nt;
          c.max_items = mi;
          c.teams = teams != 0;
          c.mixed = mixed != 0;
          c.split = false;
          cands.push_back(c);
        }
      }
    }
  }
  const PcgPlan saved = h->plan;
  PcgPlan best = saved;
  float best_ms = INFINITY;
  uint32_t nres = 0;
  hipError_t e = hipSuccess;
  for (const PcgPlan &c : cands) {
    h->plan = c;
    if (build_pcg_plan(h) != MPC_OK) {
      e = hipErrorUnknown;
      break;
    }
    float tmin = INFINITY;
    for (uint32_t r = 0; e == hipSuccess && r <= reps; ++r) {
      e = launch_tune_system(b, h->S, h->P, h->gamma, h->adj_mu, st);
      // the split system is formed in the Schur phase of a solve, outside the timed PCG
      if (e == hipSuccess && split_active(h)) e = launch_split_prep(b, h->S, h->gamma, nullptr, h->P, h->ghat, st);
      if (e == hipSuccess) e = hipEventRecord(h->ev[0], st);
      if (e == hipSuccess) e = pcg_batch(h, h->adj_mu, nullptr, iters, 0.f, nullptr, st, true);
      if (e == hipSuccess) e = hipEventRecord(h->ev[1], st);
      if (e == hipSuccess) e = hipEventSynchronize(h->ev[1]);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
      if (e == hipSuccess && r > 0) tmin = ms < tmin ? ms : tmin;  // pass 0 is the untimed warm-up
    }
    if (e != hipSuccess) break;
    if (nres < max_results) plan_to_c(c, 1e3f * tmin, &results[nres++]);
    if (tmin < best_ms) {
      best_ms = tmin;
      best = c;
    }
  }
  h->plan = e == hipSuccess ? best : saved;
  const int rc = build_pcg_plan(h);
  if (n_results) *n_results = nres;
  if (e != hipSuccess) return MPC_ERR_HIP;
  return rc;
}

int mpc_batch_pcg_launch_map(const mpc_handle *h, uint32_t *launch_of_inst) {
  if (!h || !launch_of_inst) return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  const bool one = split_active(h) || (h->cfg.flags & (FLAG_PCG_STREAM | FLAG_DIRECT_SOLVE));
  for (uint32_t i = 0; i < h->n_inst; ++i) launch_of_inst[i] = one ? 0u : h->pcg_launch[i];
  return MPC_OK;
}

int mpc_batch_set_timing(mpc_handle *h, int enable) {
  if (!h) return MPC_ERR_ARG;
  h->timing = enable != 0;
  return MPC_OK;
}

int mpc_batch_pcg_exit(mpc_handle *h, uint32_t *d_iters, uint32_t *d_flags, float *d_eta, mpc_stream_t stream) {
  if (!h) return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  hipStream_t st = (hipStream_t)stream;
  const size_t n = h->n_inst;
  if (d_iters) HIP_TRY(hipMemcpyAsync(d_iters, h->pcg_iters, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
  if (d_flags) HIP_TRY(hipMemcpyAsync(d_flags, h->pcg_flag, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
  if (d_eta) HIP_TRY(hipMemcpyAsync(d_eta, h->pcg_eta, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  return MPC_OK;
}

int mpc_batch_direct_resolve(mpc_handle *h, float *d_lambda_out, uint32_t refine_steps, mpc_stream_t stream) {
  if (!h || !d_lambda_out || refine_steps > 16) return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  if (h->gc_compact)  // G and C of the last linearisation from its compact form
    HIP_TRY(launch_expand_gc(batch_of(h), make_params(h), h->gq, h->Cl, h->G, h->C, (hipStream_t)stream));
  if (!h->schur_formed)  // the last solve was in split form: S and gamma at its last iterate and rho
    HIP_TRY(launch_form_schur(batch_of(h), h->G, h->C, h->g, h->c, h->rho_used, 0.f, nullptr, h->S, h->P, h->gamma,
                              h->Ginv, 0, (hipStream_t)stream));
  HIP_TRY(launch_direct_solve(batch_of(h), h->dir_order, h->n_inst, h->S, h->gamma, nullptr, h->P, d_lambda_out,
                              refine_steps, true, nullptr, nullptr, (hipStream_t)stream));
  return MPC_OK;
}

int mpc_batch_phase_times(mpc_handle *h, double *us7) {
  if (!h || !us7) return MPC_ERR_ARG;
  for (int i = 0; i < 7; ++i) us7[i] = h->phase_us[i];
  return MPC_OK;
}

// PCG of every instance of the batch on S, P, gamma of the workspace: long
// horizons first on two-workgroup teams (the last team launch also solves
// the first bins of the resident kernel in its spare workgroups), then any
// beyond their reach on the streaming kernel, then the remaining bins (or
// everything streaming with FLAG_PCG_STREAM).  mid (optional) is recorded
// after the team launches.
static hipError_t pcg_batch(mpc_handle *h, float *d_lambda, const uint32_t *d_status, uint32_t max_iter, float tol,
                            hipEvent_t mid, hipStream_t st, bool prepped) {
  const Batch b = batch_of(h);
  hipError_t e;
  if (split_active(h)) {
    // block Jacobi in split form: the split system from S and gamma (unless
    // the caller formed it already), then one launch over the bins
    e = prepped ? hipSuccess : launch_split_prep(b, h->S, h->gamma, d_status, h->P, h->ghat, st);
    if (e == hipSuccess)
      e = launch_pcg_split(b, h->split_items, h->split_off, h->n_split_bins, h->P, h->ghat, d_lambda, max_iter, tol,
                           d_status, h->pcg_iters, h->pcg_flag, h->pcg_eta, st);
    if (e == hipSuccess && mid) e = hipEventRecord(mid, st);
    return e;
  }
  if (h->cfg.flags & FLAG_DIRECT_SOLVE) {
    e = launch_direct_solve(b, h->dir_order, h->n_inst, h->S, h->gamma, d_status, h->P, d_lambda, h->cfg.direct_refine,
                            true, h->pcg_iters, h->pcg_flag, st);
    if (e == hipSuccess && mid) e = hipEventRecord(mid, st);
    return e;
  }
  const bool stair = h->cfg.flags & FLAG_PCG_STAIR;
  if (h->cfg.flags & FLAG_PCG_STREAM) {
    e = launch_pcg_stream(b, h->max_n, h->S, h->P, h->gamma, d_lambda, max_iter, tol, d_status, nullptr, 0,
                          h->pcg_iters, h->pcg_flag, h->pcg_eta, stair, st);
    if (e == hipSuccess && mid) e = hipEventRecord(mid, st);
    return e;
  }
  // four-workgroup teams, then the two-workgroup teams and the 512-lane bins
  // (one persistent pool launch, or team launches and a bin launch), then
  // the streamed instances and the 256-lane bins
  e = launch_pcg_team(b, 4, h->team_list + h->n_teams, h->n_quads, h->mbox, h->n_mbox, h->S, h->P, h->gamma,
                      d_lambda, max_iter, tol, d_status, h->pcg_iters, h->pcg_flag, h->pcg_eta, h->pcg_done,
                      h->team_err, stair, h->fault_inject, st);
  const bool pool = h->plan.mixed && pcg_pool_ok(h->n_teams, max_iter);
  if (e == hipSuccess && pool)
    e = launch_pcg_pool(b, h->team_list, h->n_teams, h->mbox, h->n_mbox, h->pool_ctr, h->bin_items, h->bin_off,
                        h->n_bins, h->S, h->P, h->gamma, d_lambda, max_iter, tol, d_status, h->pcg_iters, h->pcg_flag,
                        h->pcg_eta, h->pcg_done, h->team_err, stair, h->fault_inject, st);
  if (e == hipSuccess && !pool)
    e = launch_pcg_team(b, 2, h->team_list, h->n_teams, h->mbox, h->n_mbox, h->S, h->P, h->gamma, d_lambda,
                        max_iter, tol, d_status, h->pcg_iters, h->pcg_flag, h->pcg_eta, h->pcg_done, h->team_err,
                        stair, h->fault_inject, st);
  if (e == hipSuccess && mid) e = hipEventRecord(mid, st);
  if (e == hipSuccess)
    e = launch_pcg_stream(b, h->max_big_n, h->S, h->P, h->gamma, d_lambda, max_iter, tol, d_status, h->big_list,
                          h->n_big, h->pcg_iters, h->pcg_flag, h->pcg_eta, stair, st);
  if (e == hipSuccess && !pool)
    e = launch_pcg_res(b, h->bin_items, h->bin_off, 0, h->n_bins, 512, h->S, h->P, h->gamma, d_lambda, max_iter,
                       tol, d_status, h->pcg_iters, h->pcg_flag, h->pcg_eta, stair, st);
  if (e == hipSuccess)
    e = launch_pcg_res(b, h->bin_items, h->bin_off, h->n_bins, h->n_bins_small, 256, h->S, h->P, h->gamma,
                       d_lambda, max_iter, tol, d_status, h->pcg_iters, h->pcg_flag, h->pcg_eta, stair, st);
  return e;
}

// sqp_iters batched SQP iterations, stream-ordered on the device; with
// h->timing, per-phase event times are added to h->phase_us (one
// synchronisation per iteration).
static hipError_t batch_sqp_run(mpc_handle *h, float *d_xu, float *d_lambda, const float *d_ee_ref, float *d_rho,
                                uint32_t sqp_iters, uint32_t *d_status, uint32_t *d_pcg_iters, hipStream_t st) {
  const Batch b = batch_of(h);
  const Params prm = make_params(h);
  InstState s{d_rho, h->drho, h->merit0, d_status, h->sqp_done, h->xs, h->t0};
  const bool tm = h->timing;
  hipError_t e;
#define MPC_RUN(expr)           \
  do {                          \
    e = (expr);                 \
    if (e != hipSuccess) return e; \
  } while (0)
  MPC_RUN(launch_merit(b, prm, d_xu, nullptr, d_ee_ref, h->xs, nullptr, 8, 1, h->knot_merits, st));
  MPC_RUN(launch_batch_init(b, d_xu, h->knot_merits, s, st));
  for (uint32_t it = 0; it < sqp_iters; ++it) {
    if (tm) MPC_RUN(hipEventRecord(h->ev[0], st));
    // split form: G and C in the compact form (gq, Cl), the split system
    // with the Schur block rows (no S, no G~^-1: Q~^-1 and R~^-1 in closed
    // form from gq)
    const bool split = split_active(h);
    h->schur_formed = !split;
    h->gc_compact = split;
    MPC_RUN(launch_form_kkt(b, prm, d_xu, d_ee_ref, h->xs, d_status, h->G, h->C, h->g, h->c, st, h->gq,
                            split ? h->Cl : nullptr));
    if (tm) MPC_RUN(hipEventRecord(h->ev[1], st));
    if (split)
      MPC_RUN(launch_form_split(b, prm, h->Cl, h->g, h->c, h->gq, d_rho, 0.f, d_status, h->P, h->ghat, h->rho_used,
                                st));
    else
      MPC_RUN(launch_form_schur(b, h->G, h->C, h->g, h->c, d_rho, 0.f, d_status, h->S, h->P, h->gamma, h->Ginv, 0,
                                st));
    if (tm) MPC_RUN(hipEventRecord(h->ev[2], st));
    MPC_RUN(pcg_batch(h, d_lambda, d_status, sqp_pcg_cap(h->cfg), h->cfg.pcg_exit_tol, tm ? h->ev[3] : nullptr, st,
                      split));
    if (tm) MPC_RUN(hipEventRecord(h->ev[4], st));
    if (split)
      MPC_RUN(launch_compute_dz_lr(b, prm, h->gq, d_rho, h->Cl, h->g, d_lambda, d_status, h->dz, st));
    else
      MPC_RUN(launch_compute_dz(b, h->Ginv, h->C, h->g, d_lambda, d_status, h->dz, st));
    if (h->cfg.flags & FLAG_DIRECT_SOLVE)
      MPC_RUN(refine_kkt(h, d_rho, h->g, h->c, h->dz, d_lambda, h->cfg.direct_kkt_refine, 0.f, st));
    if (tm) MPC_RUN(hipEventRecord(h->ev[5], st));
    MPC_RUN(launch_merit(b, prm, d_xu, h->dz, d_ee_ref, h->xs, d_status, 0, 8, h->knot_merits, st));
    if (tm) MPC_RUN(hipEventRecord(h->ev[6], st));
    MPC_RUN(launch_ls_update(b, prm, h->knot_merits, h->dz, d_xu, s, d_pcg_iters, h->pcg_iters, h->pcg_flag, it,
                             sqp_iters, st));
    if (tm) {
      MPC_RUN(hipEventRecord(h->ev[7], st));
      MPC_RUN(hipEventSynchronize(h->ev[7]));
      for (int ph = 0; ph < 7; ++ph) {
        float ms = 0.f;
        MPC_RUN(hipEventElapsedTime(&ms, h->ev[ph], h->ev[ph + 1]));
        h->phase_us[ph] += 1e3 * ms;
      }
    }
  }
#undef MPC_RUN
  return hipSuccess;
}

static void drop_batch_graph(mpc_handle *h) {
  if (h->bexec) (void)hipGraphExecDestroy(h->bexec);
  if (h->bgraph) (void)hipGraphDestroy(h->bgraph);
  h->bexec = nullptr;
  h->bgraph = nullptr;
}

int mpc_batch_set_graph(mpc_handle *h, int enable) {
  if (!h) return MPC_ERR_ARG;
  h->graph_on = enable != 0;
  if (!h->graph_on) drop_batch_graph(h);
  return MPC_OK;
}

int mpc_batch_sqp(mpc_handle *h, float *d_xu, float *d_lambda, const float *d_ee_ref, float *d_rho,
                  uint32_t sqp_iters, uint32_t *d_status, uint32_t *d_pcg_iters, mpc_stream_t stream) {
  if (!h || !d_xu || !d_lambda || !d_ee_ref || !d_rho || !d_status) return MPC_ERR_ARG;
  if (!h->batch_ready) return MPC_ERR_NOT_READY;
  if (sqp_iters > MPC_MAX_SQP_ITERS) return MPC_ERR_ARG;
  for (double &v : h->phase_us) v = 0;
  if (h->graph_on && !h->timing) {
    mpc_handle::GraphKey k;
    k.epoch = h->graph_epoch;
    k.sqp_iters = sqp_iters;
    const void *p[6] = {d_xu, d_lambda, d_ee_ref, d_rho, d_status, d_pcg_iters};
    std::memcpy(k.ptr, p, sizeof(p));
    if (!h->bexec || !(k == h->graph_key)) {
      drop_batch_graph(h);
      hipStream_t cs = h->cl_stream;
      HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      const hipError_t er = batch_sqp_run(h, d_xu, d_lambda, d_ee_ref, d_rho, sqp_iters, d_status, d_pcg_iters, cs);
      const hipError_t ee = hipStreamEndCapture(cs, &h->bgraph);
      HIP_TRY(er);
      HIP_TRY(ee);
      HIP_TRY(hipGraphInstantiate(&h->bexec, h->bgraph, nullptr, nullptr, 0));
      h->graph_key = k;
    }
    HIP_TRY(hipGraphLaunch(h->bexec, (hipStream_t)stream));
    return MPC_OK;
  }
  HIP_TRY(batch_sqp_run(h, d_xu, d_lambda, d_ee_ref, d_rho, sqp_iters, d_status, d_pcg_iters,
                        (hipStream_t)stream));
  return MPC_OK;
}

// ---- closed loop ----------------------------------------------------------------

void mpc_default_closed_loop(mpc_closed_loop_config *cl) {
  if (!cl) return;
  std::memset(cl, 0, sizeof(*cl));
  cl->steps = 2000;
  cl->sqp_iters = 3;
  cl->substeps = 50;
  cl->sim_dt = 2e-4f;
  cl->noise_sigma = 1e-3f;
  cl->noise_seed = 12345;
  cl->use_graph = 1;
}

int mpc_closed_loop(mpc_handle *h, const mpc_closed_loop_config *cl, uint32_t N, const float *d_x0,
                    const float *d_ee_path, const float *d_xu_traj, float *d_xu, float *d_lambda, float *d_states,
                    float *d_ee_err, uint32_t *d_pcg_iters, mpc_closed_loop_stats *stats, mpc_stream_t stream) {
  if (!h || !cl || !d_x0 || !d_ee_path || !d_xu || !d_lambda || !d_states || !d_ee_err) return MPC_ERR_ARG;
  if (!n_ok(h, N) || cl->steps < 1 || cl->sqp_iters < 1 || cl->sqp_iters > MPC_MAX_SQP_ITERS ||
      cl->substeps < 1 || !(cl->sim_dt > 0.f) || !std::isfinite(cl->sim_dt) || !(cl->noise_sigma >= 0.f) ||
      !std::isfinite(cl->noise_sigma) || (cl->use_graph && cl->timing))
    return MPC_ERR_ARG;
  const uint32_t horizon[1] = {N};
  int rc = mpc_batch_setup(h, 1, horizon);
  if (rc != MPC_OK) return rc;
  hipStream_t user = (hipStream_t)stream;
  hipStream_t st = cl->use_graph ? h->cl_stream : user;
  const ClStep step{N, cl->substeps, cl->sqp_iters, cl->sim_dt, cl->noise_sigma, cl->noise_seed};
  double solve_us = 0, linsys_us = 0, plant_us = 0;
  const bool tm = cl->timing != 0;
  const bool saved_timing = h->timing;
  h->timing = tm;
  auto one_step = [&]() -> hipError_t {
    if (tm)
      for (double &v : h->phase_us) v = 0;
    hipError_t e = batch_sqp_run(h, d_xu, d_lambda, h->cl_win, h->cl_rho, cl->sqp_iters, h->cl_status, h->cl_pcg,
                                 st);
    if (e == hipSuccess && tm) e = hipEventRecord(h->ev[0], st);
    if (e == hipSuccess)
      e = launch_cl_step(step, d_ee_path, d_xu_traj, h->cl_win, d_xu, d_lambda, h->cl_xp, h->cl_ctr, d_states,
                         d_ee_err, h->cl_pcg, d_pcg_iters, st);
    if (e == hipSuccess && tm) {
      e = hipEventRecord(h->ev[1], st);
      if (e == hipSuccess) e = hipEventSynchronize(h->ev[1]);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
      plant_us += 1e3 * ms;
      for (int ph = 0; ph < 7; ++ph) solve_us += h->phase_us[ph];
      linsys_us += h->phase_us[1] + h->phase_us[2] + h->phase_us[3] + h->phase_us[4];
    }
    return e;
  };
  hipError_t e = hipSuccess;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t t0 = h->ev[9], t1 = h->ev[10];
  if (cl->use_graph) {
    // order the private stream after the caller's work
    e = hipEventRecord(t0, user);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, t0, 0);
  }
  if (e == hipSuccess)
    e = launch_cl_init(N, d_x0, d_ee_path, h->cl_win, d_xu, h->cl_xp, h->cl_ctr, d_states, h->cl_rho,
                       h->cfg.rho_reset, st);
  if (e == hipSuccess && cl->use_graph) {
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t ec = one_step();
      e = hipStreamEndCapture(st, &graph);
      if (ec != hipSuccess) e = ec;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  }
  if (e == hipSuccess) e = hipEventRecord(t0, st);
  for (uint32_t k = 0; e == hipSuccess && k < cl->steps; ++k) e = cl->use_graph ? hipGraphLaunch(exec, st) : one_step();
  if (e == hipSuccess) e = hipEventRecord(t1, st);
  if (e == hipSuccess && cl->use_graph) e = hipStreamWaitEvent(user, t1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(t1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  h->timing = saved_timing;
  if (e != hipSuccess) return MPC_ERR_HIP;
  if (stats) {
    stats->total_us = 1e3 * ms;
    stats->solve_us = solve_us;
    stats->linsys_us = linsys_us;
    stats->plant_us = plant_us;
    stats->graph = cl->use_graph ? 1u : 0u;
    stats->reserved = 0;
  }
  return MPC_OK;
}

// ---- synthetic workload -------------------------------------------------------

static inline uint64_t splitmix64(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double uniform01(uint64_t &state) { return (double)(splitmix64(state) >> 11) * 0x1.0p-53; }
static inline double normal01(uint64_t &state) {
  const double u1 = 1.0 - uniform01(state), u2 = uniform01(state);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

void mpc_synth_instance(uint64_t seed, uint32_t N, float dt, float *h_x0, float *h_qref) {
  uint64_t st = seed;
  double q0[NQ], qd0[NQ], f[NQ], ph[NQ];
  for (int j = 0; j < NQ; ++j) q0[j] = -0.5 + uniform01(st);
  for (int j = 0; j < NQ; ++j) qd0[j] = 0.1 * normal01(st);
  for (int j = 0; j < NQ; ++j) f[j] = 0.2 + 0.8 * uniform01(st);
  for (int j = 0; j < NQ; ++j) ph[j] = 2.0 * M_PI * uniform01(st);
  if (h_x0)
    for (int j = 0; j < NQ; ++j) {
      h_x0[j] = (float)q0[j];
      h_x0[NQ + j] = (float)qd0[j];
    }
  if (h_qref)
    for (uint32_t k = 0; k < N; ++k)
      for (int j = 0; j < NQ; ++j)
        h_qref[(size_t)NQ * k + j] =
            (float)(q0[j] + 0.3 * std::sin(2.0 * M_PI * f[j] * (double)k * (double)dt + ph[j]));
}

void mpc_synth_track(uint64_t seed, uint32_t K, float dt, float *h_x0, float *h_qqdqdd) {
  uint64_t st = seed;
  double q0[NQ], qd0[NQ], f[NQ], ph[NQ];
  for (int j = 0; j < NQ; ++j) q0[j] = -0.5 + uniform01(st);
  for (int j = 0; j < NQ; ++j) qd0[j] = 0.1 * normal01(st);
  for (int j = 0; j < NQ; ++j) f[j] = 0.2 + 0.8 * uniform01(st);
  for (int j = 0; j < NQ; ++j) ph[j] = 2.0 * M_PI * uniform01(st);
  if (h_x0)
    for (int j = 0; j < NQ; ++j) {
      h_x0[j] = (float)q0[j];
      h_x0[NQ + j] = (float)qd0[j];
    }
  if (h_qqdqdd)
    for (uint32_t k = 0; k < K; ++k)
      for (int j = 0; j < NQ; ++j) {
        const double w = 2.0 * M_PI * f[j], a = w * (double)k * (double)dt + ph[j];
        float *o = h_qqdqdd + (size_t)NXU * k;
        o[j] = (float)(q0[j] + 0.3 * std::sin(a));
        o[NQ + j] = (float)(0.3 * w * std::cos(a));
        o[NX + j] = (float)(-0.3 * w * w * std::sin(a));
      }
}

void mpc_synth_horizons(uint64_t seed, uint32_t n_inst, const uint32_t *choices, uint32_t n_choices,
                        uint32_t *horizons) {
  if (!choices || !horizons || n_choices == 0) return;
  uint64_t st = seed;
  for (uint32_t i = 0; i < n_inst; ++i) {
    uint32_t idx = (uint32_t)(uniform01(st) * n_choices);
    if (idx >= n_choices) idx = n_choices - 1;
    horizons[i] = choices[idx];
  }
}

}  // extern "C"
