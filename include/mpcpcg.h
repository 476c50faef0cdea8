/*
 * mpcpcg.h -- C-ABI of the MI355X-native SQP tracking-MPC inner loop
 * (KUKA iiwa 7-DOF end-effector tracking; nx = 14, nu = 7, fp32).
 *
 * Every entry point below replaces one piece of the reference call surface
 * of EmreAdabag/gpu-pcg-tracking (file:line cited per function).  Buffers
 * use the reference's layouts, all fp32, column-major 14x14 / 14x7 / 7x7
 * blocks, so fixtures and buffers are interchangeable:
 *
 *   xu      21N-7   [x_0 u_0 x_1 u_1 ... x_{N-1}]           (toplevel.cuh:48)
 *   ee_ref  6N      [xyz rpy] per knot; only xyz enters the cost
 *   lambda  14N     Schur multipliers
 *   G       245N-49 per knot [Q_k 196 | R_k 49], last knot Q only
 *   C       294(N-1) per knot [-A_k 196 | -B_k 98]
 *   g       21N-7   per knot [q_k 14 | r_k 7], last knot q only
 *   c       14N     c_0 = x_0 - x_s, c_{k+1} = x_{k+1} - (x_k + dt f(x_k,u_k))
 *   S, Pinv 588N    block row b = [S(b,b-1) | S(b,b) | S(b,b+1)], 196 each
 *   gamma   14N
 *   dz      21N-7
 *
 * All device pointers are hipMalloc'd device memory owned by the caller.
 * The compute entry points are stream-ordered on the given stream (NULL =
 * default stream) and none of them synchronises the device; the set-up
 * calls without a stream argument (mpc_create, mpc_destroy, mpc_batch_setup,
 * mpc_batch_set_plan) copy and synchronise like hipMemcpy.  A handle is
 * used by one host thread at a time.  Errors are returned as MPC_* status codes; no
 * entry point exits the process.
 */
#ifndef MPCPCG_H
#define MPCPCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *mpc_stream_t; /* == hipStream_t */

enum {
  MPC_OK = 0,
  MPC_ERR_ARG = 1,         /* bad argument (null pointer, N < 2, size beyond capacity) */
  MPC_ERR_HIP = 2,         /* a HIP runtime call failed */
  MPC_ERR_NOT_READY = 3,   /* batch layout not set up */
  MPC_ERR_NONFINITE = 4,   /* non-finite value detected */
  MPC_ERR_NO_DEVICE = 5,   /* no usable GPU */
  MPC_ERR_PCG_FAULT = 6    /* a PCG team exchange timed out (the solve was abandoned) */
};

/* per-instance status bits (batched solve) */
enum {
  MPC_ST_ACTIVE = 1,       /* still iterating */
  MPC_ST_PCG_MAXITER = 2,  /* some PCG solve hit max_iter */
  MPC_ST_RHO_ABORT = 4,    /* line search failed with rho > rho_max (toplevel.cuh:396-407) */
  MPC_ST_NONFINITE = 8,    /* a merit or step was not finite */
  MPC_ST_PCG_FAULT = 16,   /* a PCG team exchange timed out: the instance was stopped, its step not taken */
  MPC_ST_TIME_BUDGET = 32  /* the SQP time budget ran out: the instance was stopped, its last step not taken */
};

/* Runtime configuration.  Reference equivalents: settings.cuh and the -D
 * flags of perform_experiment.sh:42-59; pcg_config (toplevel.cuh:573-576). */
typedef struct mpc_config {
  uint32_t max_knots;        /* largest N of one instance */
  uint32_t max_instances;    /* batch capacity */
  uint32_t max_total_knots;  /* capacity of sum N_i over a batch */
  float dt;                  /* integrator step (runme.cu:65) */
  float q_ee;                /* end-effector position weight (reference: 1) */
  float qd_cost;             /* QD_COST */
  float r_cost;              /* R_COST */
  float mu;                  /* merit penalty, 10 (toplevel.cuh:54) */
  float rho_factor;          /* RHO_FACTOR */
  float rho_max;             /* RHO_MAX */
  float rho_min;             /* RHO_MIN (settings.cuh:50) */
  float rho_reset;           /* rho after an abort in batched solves (track passes 1e-3, toplevel.cuh:579) */
  float pcg_exit_tol;        /* pcg_config.pcg_exit_tol */
  uint32_t pcg_max_iter;     /* pcg_config.pcg_max_iter */
  uint32_t sqp_max_iter;     /* SQP_MAX_ITER */
  uint32_t flags;            /* MPC_FLAG_* */
  /* MPC_FLAG_DIRECT_SOLVE: iterative-refinement steps (fp64 residuals) of the
   * Schur solve and of the whole KKT step (dz, lambda) after it (<= 16 each;
   * defaults 0 and 1).  The dz recovery cancels, so the step needs the
   * latter: with one KKT step the SQP follows the fp64 direct SQP to ~3e-5,
   * with two to ~1e-6. */
  uint32_t direct_refine;
  uint32_t direct_kkt_refine;
  /* SQP_MAX_TIME_US (perform_experiment.sh:45, sqpTimecheck at
   * toplevel.cuh:223-231): wall-clock budget of one solve in microseconds,
   * 0 = none.  The plain driver checks its host clock where the reference
   * does (after the linearisation and Schur launches, and after the
   * line-search merits: the iteration then ends without its step); the
   * batched solve checks the device clock against the solve's start at
   * each line search, and an instance past the budget stops there with
   * MPC_ST_TIME_BUDGET without taking that step. */
  float sqp_max_time_us;
} mpc_config;

/* The terminal knot's cost uses x_{N-1} instead of the reference's x_{N-2}
 * (rbdfiles/iiwa_eepos_plant.cuh:399 passes the knot N-2 state). */
#define MPC_FLAG_LAST_KNOT_OWN_STATE 1u
/* Always use the streaming PCG kernel (S read from memory every iteration)
 * instead of the on-chip resident kernel; for kernel comparisons. */
#define MPC_FLAG_PCG_STREAM 2u
/* Solve the Schur system directly (block LDL^T of -S on the GPU, exact up
 * to fp32 rounding) instead of by PCG, in the plain and batched SQP and the
 * QP layer: the GPU counterpart of the reference's QDLDL path
 * (PCG_SOLVE = 0, toplevel.cuh:160-215). */
#define MPC_FLAG_DIRECT_SOLVE 4u
/* PCG with the stair preconditioner M = D^-1 - D^-1 (L + L^T) D^-1 (D, L: the
 * block diagonal and the lower block part of S) instead of block Jacobi
 * D^-1: the reference's gato_form_ss_inner (oldschur.cuh:334-494), present
 * but unused there.  Applied as M r = D^-1 (2 r - S D^-1 r), it halves the
 * iteration count on the block-tridiagonal Schur system for one more block
 * product (and one more neighbour exchange) per iteration. */
#define MPC_FLAG_PCG_STAIR 8u
/* The SQP solves' PCG cap per horizon as in the reference's experiments
 * (perform_experiment.sh:14-19: 173 at N = 32, 167 at N = 64 and 128, 118 at
 * N = 256, 67 at N = 512), taken for the tabulated horizon at or above N,
 * instead of pcg_max_iter for all horizons. */
#define MPC_FLAG_PCG_CAP_BY_N 16u

typedef struct mpc_handle mpc_handle;

/* Statistics of one plain SQP solve; mirrors the tuple returned by sqpSolve
 * (toplevel.cuh:502): per-iteration PCG iterations and max-iter flags,
 * per-iteration linear-system time (PCG kernel, device events), SQP time,
 * SQP iterations, and whether the solve ended without a rho abort. */
#define MPC_MAX_SQP_ITERS 64
typedef struct mpc_sqp_stats {
  uint32_t sqp_iters;
  uint32_t not_aborted;
  double sqp_us;
  uint32_t pcg_iters[MPC_MAX_SQP_ITERS];
  uint32_t pcg_maxiter[MPC_MAX_SQP_ITERS];
  double linsys_us[MPC_MAX_SQP_ITERS];
  int32_t ls_step[MPC_MAX_SQP_ITERS];   /* accepted alpha index, -1 = failure */
  uint32_t time_exit;                   /* 1: the solve ended on sqp_max_time_us */
  uint32_t reserved;
} mpc_sqp_stats;

/* Fill *cfg with the defaults of the published workload: dt = 0.01,
 * Q_ee = 1, QD = 1e-2, R = 1e-4, mu = 10, rho_factor 1.2, rho_max 10,
 * rho_min 1e-3, rho_reset 1e-3, tol 1e-5, max_iter 173, 3 SQP iterations. */
void mpc_default_config(mpc_config *cfg);

/* Create / destroy a handle; all workspace for max_knots / max_instances /
 * max_total_knots is allocated here, none inside the solves.  Replaces the
 * per-call allocations of sqpSolve (toplevel.cuh:67-215) and
 * gato_plant::initializeDynamicsConstMem (rbdfiles/iiwa_eepos_plant.cuh:66-74). */
int mpc_create(const mpc_config *cfg, mpc_handle **out);
int mpc_destroy(mpc_handle *h);
int mpc_get_config(const mpc_handle *h, mpc_config *cfg);
/* Update the numeric parameters (not the capacities) of a handle. */
int mpc_set_config(mpc_handle *h, const mpc_config *cfg);
const char *mpc_status_string(int status);

/* ---- robot model (GRiD-level) -------------------------------------------- */

/* End-effector xyz + rpy of n joint vectors (stride q_stride floats).
 * grid::end_effector_positions_kernel (rbdfiles/iiwa_eepos_grid.cuh:2141-2174). */
int mpc_fk(mpc_handle *h, uint32_t n, const float *d_q, uint32_t q_stride, float *d_ee,
           mpc_stream_t stream);

/* tau = M(q) qdd + c(q, qd) (7 per point) for n points [q qd qdd] (21 floats
 * each), gravity 0 as the plant.  grid::inverse_dynamics_kernel
 * (rbdfiles/iiwa_eepos_grid.cuh, called with gato_plant::GRAVITY = 0,
 * rbdfiles/iiwa_eepos_plant.cuh:51). */
int mpc_inverse_dynamics(mpc_handle *h, uint32_t n, const float *d_qqdqdd, float *d_tau, mpc_stream_t stream);

/* qdd (7 per point) and df_du (7x21 per point: dqdd/dq, dqdd/dqd, dqdd/du)
 * for n points [q qd u] (21 floats each).
 * gato_plant::forwardDynamicsAndGradient (rbdfiles/iiwa_eepos_plant.cuh:126-156). */
int mpc_dynamics_gradient(mpc_handle *h, uint32_t n, const float *d_qqdu, float *d_qdd,
                          float *d_df_du, mpc_stream_t stream);

/* ---- stage entry points (one instance, reference layouts) ---------------- */

/* KKT blocks G, C, g, c from the trajectory.  gato_form_kkt (schur.cuh:391-524). */
int mpc_form_kkt(mpc_handle *h, uint32_t N, const float *d_xu, const float *d_ee_ref,
                 const float *d_xs, float *d_G, float *d_C, float *d_g, float *d_c,
                 mpc_stream_t stream);

/* Schur complement S, block-Jacobi Pinv (off-diagonal blocks written as 0),
 * gamma, and G overwritten with the regularised inverses (Q_k+rho I)^-1,
 * (R_k+rho I)^-1 as the reference does.  form_schur (schur.cuh:545-559). */
int mpc_form_schur(mpc_handle *h, uint32_t N, float *d_G, const float *d_C, const float *d_g,
                   const float *d_c, float *d_S, float *d_Pinv, float *d_gamma, float rho,
                   mpc_stream_t stream);

/* PCG on S lambda = gamma, warm-started from lambda.  Exit when
 * |r^T Pinv r| < exit_tol (also tested before the first iteration); writes
 * the iteration count and a max-iter flag (1 = max_iter reached, 2 = a
 * PCG team's exchange timed out and the solve was abandoned).
 * GPU-PCG pcg<T,14,N> (call site toplevel.cuh:143-157, :292-294). */
int mpc_pcg_solve(mpc_handle *h, uint32_t N, const float *d_S, const float *d_Pinv,
                  const float *d_gamma, float *d_lambda, uint32_t max_iter, float exit_tol,
                  uint32_t *d_iters, uint32_t *d_max_iter_exit, mpc_stream_t stream);

/* Direct solve of S lambda = gamma (one instance of N knots) by block LDL^T
 * of -S on the GPU, then refine_steps (<= 16) steps of iterative refinement
 * with an fp64 residual; d_work: 588 N floats of scratch (the layout of S:
 * the elimination factors and a vector per block row).  GPU counterpart of
 * the reference's QDLDL solve (qdldl_helper.cuh:62-180, toplevel.cuh:314-335);
 * the SQP drivers use 2 refinement steps. */
int mpc_direct_solve(mpc_handle *h, uint32_t N, const float *d_S, const float *d_gamma, float *d_lambda,
                     uint32_t refine_steps, float *d_work, mpc_stream_t stream);

/* mpc_direct_solve without a handle and without d_work: allocates nothing
 * (no hipMalloc, no host allocation), for callers that cannot allocate
 * device memory, such as code under HIP graph capture or tight serving
 * loops.  No counterpart in the reference.  The factor workspace is a
 * library-owned device array of 588 x MPC_DIRECT_STATIC_MAX_KNOTS floats
 * (1.2 MB), one per GPU, held until the process ends; N beyond that limit
 * (or N < 2, refine_steps > 16, a null pointer) returns MPC_ERR_ARG.  All
 * calls on one GPU share that array, so calls that overlap in time on one
 * GPU (other streams, concurrent graph replays) race; calls ordered on one
 * stream do not.  Results are bitwise those of mpc_direct_solve. */
#define MPC_DIRECT_STATIC_MAX_KNOTS 512u
int mpc_direct_solve_static(uint32_t N, const float *d_S, const float *d_gamma, float *d_lambda,
                            uint32_t refine_steps, mpc_stream_t stream);

/* dz from G^-1, C, g, lambda.  compute_dz (schur.cuh:562-566). */
int mpc_compute_dz(mpc_handle *h, uint32_t N, const float *d_Ginv, const float *d_C,
                   const float *d_g, const float *d_lambda, float *d_dz, mpc_stream_t stream);

/* Merit of xu (no initial-state term).  compute_merit (merit.cuh:99-145). */
int mpc_compute_merit(mpc_handle *h, uint32_t N, const float *d_xu, const float *d_ee_ref,
                      float *d_merit_out, mpc_stream_t stream);

/* The 8 line-search merits of xu + alpha_p dz, alpha_p = -2^-p, including
 * mu |x_0 + alpha_p dz_0 - x_s|_1.  ls_gato_compute_merit x 8
 * (merit.cuh:17-94, toplevel.cuh:357-374). */
int mpc_line_search_merits(mpc_handle *h, uint32_t N, const float *d_xs, const float *d_xu,
                           const float *d_ee_ref, const float *d_dz, float *d_merits_out,
                           mpc_stream_t stream);

/* ---- full solves ----------------------------------------------------------- */

/* One SQP solve of one instance; drop-in for sqpSolve (toplevel.cuh:24-503).
 * d_xu (in/out; x_0 is the current state), d_lambda (in/out warm start),
 * *rho (in/out).  Host-synchronising once per SQP iteration, like the
 * reference, to report per-iteration statistics.  Returns
 * MPC_ERR_NONFINITE (the iterate left as it was, stats up to that point)
 * when the initial merit, or the merit of a step the line search rejected,
 * is not finite, and MPC_ERR_PCG_FAULT when a PCG team exchange timed
 * out. */
int mpc_sqp_solve(mpc_handle *h, uint32_t N, const float *d_ee_ref, float *d_lambda,
                  float *d_xu, float *rho, float rho_reset, mpc_sqp_stats *stats,
                  mpc_stream_t stream);

/* Batched solves of many instances with uneven horizons (no reference
 * counterpart: the reference solves one instance per call).  Instance i has
 * horizon horizons[i]; its buffers are packed back to back in the
 * reference per-instance layouts: xu at 21*K_i - 7*i, lambda at 14*K_i,
 * ee_ref at 6*K_i, with K_i = sum_{j<i} horizons[j].  A set-up call: it
 * uploads the layout tables with synchronous copies (hipMemcpy), so it
 * waits for the device's earlier work; the solves that follow are
 * stream-ordered. */
int mpc_batch_setup(mpc_handle *h, uint32_t n_inst, const uint32_t *horizons);

/* Run sqp_iters SQP iterations of every instance, entirely stream-ordered
 * on the device (line-search acceptance and rho rule on the device).
 * d_rho (n_inst floats, in/out), d_status (n_inst uint32, out: MPC_ST_*),
 * d_pcg_iters (n_inst x sqp_iters uint32, out, may be NULL). */
int mpc_batch_sqp(mpc_handle *h, float *d_xu, float *d_lambda, const float *d_ee_ref, float *d_rho,
                  uint32_t sqp_iters, uint32_t *d_status, uint32_t *d_pcg_iters,
                  mpc_stream_t stream);

/* Device-event timing of the phases of the last mpc_batch_sqp call made
 * with timing enabled (mpc_batch_set_timing(h, 1)), summed over its SQP
 * iterations, in microseconds: linearisation, Schur formation, PCG team
 * launches (with the bins that share the last one), the rest of PCG, dz,
 * line-search merits, update. */
int mpc_batch_set_timing(mpc_handle *h, int enable);
/* With enable != 0, mpc_batch_sqp runs its whole launch sequence (initial
 * merit, then sqp_iters x {linearise, Schur, PCG, dz, merits, line search})
 * as a HIP graph: captured on the first call and replayed on the caller's
 * stream while the buffers, sqp_iters, configuration, cost weights and batch
 * layout stay the same (any change re-captures).  In the graph the PCG team
 * launches are ordinary kernel nodes of grids that fit the device at once;
 * an instance whose team exchange timed out ends with MPC_ST_PCG_FAULT.
 * Not used while timing is enabled.  No counterpart in the reference (its
 * SQP loop decides on the host between launches, toplevel.cuh:254-446). */
int mpc_batch_set_graph(mpc_handle *h, int enable);
/* PCG packing of the batch layout: out4 = {resident-kernel bins, teams
 * (two workgroups for N <= 256, four for N <= 512), instances on the
 * streaming kernel, cooperative launches per PCG phase before the first
 * non-cooperative one (four-workgroup team launches, then the pool or the
 * two-workgroup team launches)}. */
int mpc_batch_pcg_layout(const mpc_handle *h, uint32_t *out4);
/* The PCG launch that solves each instance (n_inst values) within one PCG
 * phase: 0 .. L-1 the cooperative launches (four-workgroup teams first,
 * then the pool of two-workgroup teams and 512-lane bins, or the
 * two-workgroup team launches), L the streaming launch
 * when there are instances beyond the teams' reach, then the launches of the
 * remaining resident bins (512-lane bins, then 256-lane bins); L = out4[3]
 * of mpc_batch_pcg_layout.  All 0 with
 * MPC_FLAG_PCG_STREAM or MPC_FLAG_DIRECT_SOLVE (one launch).  No
 * counterpart in the reference (its batch is one instance); for timing
 * analyses such as the per-launch roofline. */
int mpc_batch_pcg_launch_map(const mpc_handle *h, uint32_t *launch_of_inst);

/* Launch plan of the batched PCG.  Instances with N <= 128 are packed into
 * bins of the resident kernel (workgroups of 512 lanes = 128 knots, one per
 * CU; with small_nt = 256, instances with N <= 64 go to bins of 256 lanes =
 * 64 knots, two per CU), at most max_items instances per bin; 128 < N <= 256
 * run on two-workgroup and 256 < N <= 512 on four-workgroup teams
 * (teams = 1) or on the streaming kernel.  mixed = 1: the two-workgroup
 * teams and the 512-lane bins share one persistent launch of one workgroup
 * per CU that takes them from work queues (teams first, bins largest
 * first), so a CU whose item converged early takes the next one; mixed = 0:
 * static team launches, then one bin launch.  Every plan gives the same
 * arithmetic per instance (bins and teams: bitwise), only the packing and
 * the schedule change.  split = 1 (block Jacobi, every N <= 256): all
 * instances instead run on the split-form kernel, bins of up to max_items
 * instances on 512-lane workgroups (256 knots, one per CU); block-Jacobi
 * PCG there is CG on W^T S W (W = chol(-S(b,b))^-1 blockwise, formed with
 * the Schur block rows; in mpc_batch_sqp the regularised cost blocks are
 * inverted in closed form, Q~ being diagonal plus rank one), the same
 * iteration in exact arithmetic with its own fp32 rounding, and the other
 * fields do not apply.  Default
 * {512, 8, 1, 1, split 1}.  No counterpart in the reference (one instance
 * per solve). */
typedef struct mpc_pcg_plan {
  uint32_t small_nt;
  uint32_t max_items;
  uint32_t mixed;
  uint32_t teams;
  float us;          /* mpc_batch_autotune: PCG phase time of this plan, microseconds */
  uint32_t split;
} mpc_pcg_plan;
int mpc_batch_get_plan(const mpc_handle *h, mpc_pcg_plan *plan);
/* Sets the plan (us ignored) and re-packs the current layout. */
int mpc_batch_set_plan(mpc_handle *h, const mpc_pcg_plan *plan);
/* Autotuner: times every plan that differs for the current layout on a
 * synthetic block-tridiagonal system (diag -2.05 I, off-diagonal I), iters
 * PCG iterations in every instance (tol 0), best of reps passes after one
 * untimed pass, and keeps the fastest for later solves.  Writes up to
 * max_results timed plans to results and their count to *n_results.
 * Overwrites the Schur workspace (call it between solves); not in
 * MPC_FLAG_PCG_STREAM or MPC_FLAG_DIRECT_SOLVE mode. */
int mpc_batch_autotune(mpc_handle *h, uint32_t iters, uint32_t reps, mpc_pcg_plan *results, uint32_t max_results,
                       uint32_t *n_results, mpc_stream_t stream);
int mpc_batch_phase_times(mpc_handle *h, double *us7);

/* Exit state of each instance's last PCG solve in the batch layout (the
 * last SQP iteration's after mpc_batch_sqp): iterations (S products), the
 * max-iter flag (0 converged, 1 max_iter reached, 2 team exchange timed
 * out) and |r^T Pinv r| of its last exit test, n_inst values each (any
 * pointer may be NULL; stream-ordered device copies).  The reference
 * returns the first two per solve (toplevel.cuh:292-294, :502); the third
 * has no counterpart.  Not written in MPC_FLAG_DIRECT_SOLVE mode. */
int mpc_batch_pcg_exit(mpc_handle *h, uint32_t *d_iters, uint32_t *d_flags, float *d_eta, mpc_stream_t stream);

/* The Schur systems of the last batched SQP iteration (the workspace's S
 * and gamma) solved directly, block LDL^T of -S with refine_steps (<= 16)
 * steps of fp64-residual refinement, into d_lambda_out (14 per knot, the
 * batch's lambda layout): the accuracy reference for the PCG multipliers
 * that solve returned (the reference's QDLDL solve of the same system,
 * toplevel.cuh:314-335).  After a split-form solve (mpc_pcg_plan.split),
 * which forms no S, S and gamma are first formed from the last iteration's
 * KKT blocks and rho (its cost and dynamics blocks expanded from the
 * compact form that path keeps: g_q and the lower dynamics rows).
 * Overwrites the preconditioner workspace, so it is for
 * use after a solve, before the next one; instances that were no longer
 * active in that iteration get the solve of a stale system. */
int mpc_batch_direct_resolve(mpc_handle *h, float *d_lambda_out, uint32_t refine_steps, mpc_stream_t stream);

/* ---- closed-loop tracking ----------------------------------------------------
 * Receding-horizon tracking of one instance entirely on the device: per
 * control step the batched SQP (sqp_iters iterations from the warm start),
 * then `substeps` explicit Euler plant steps of sim_dt under the plan's
 * first control, Gaussian state noise (sigma, seeded splitmix64 +
 * Box-Muller), the end-effector error against the reference and a
 * one-knot shift of the plan, the multipliers and the reference window.
 * Replaces the tracking loop of track() (toplevel.cuh:505-775, plant
 * integrator.cuh:283-398), with one knot per control step (substeps *
 * sim_dt should equal dt).  Sets the batch layout to one instance of
 * horizon N.  Buffers (device): d_x0 (14), d_ee_path (6 (steps + N)),
 * d_xu_traj (precomputed trajectory in the xu layout over steps + N knots,
 * may be NULL: the shift fills the new horizon tail from it, else with a
 * zero control and the kept last state), d_xu (21N-7, in: initial plan,
 * out: last plan), d_lambda (14N, in/out), d_states (14 (steps + 1), out:
 * plant states), d_ee_err (2 steps, out: Euclidean and L1 norm of the xyz
 * error per step), d_pcg_iters (steps x sqp_iters, out, may be NULL).  use_graph captures
 * one control step in a HIP graph and replays it (N <= 128 with PCG, where
 * the PCG has no cooperative launch; any N with MPC_FLAG_DIRECT_SOLVE); timing adds per-phase HIP event timing
 * (one synchronisation per step, no graph). */
typedef struct mpc_closed_loop_config {
  uint32_t steps;
  uint32_t sqp_iters;
  uint32_t substeps;
  float sim_dt;
  float noise_sigma;
  uint32_t use_graph;
  uint64_t noise_seed;
  uint32_t timing;
  uint32_t reserved;
} mpc_closed_loop_config;

typedef struct mpc_closed_loop_stats {
  double total_us;   /* device time of the whole loop (HIP events) */
  double solve_us;   /* timing = 1: SQP phases, summed over the steps */
  double linsys_us;  /* timing = 1: Schur + PCG + dz phases, summed */
  double plant_us;   /* timing = 1: plant / shift kernel, summed */
  uint32_t graph;    /* 1 when the loop ran as graph replays */
  uint32_t reserved;
} mpc_closed_loop_stats;

void mpc_default_closed_loop(mpc_closed_loop_config *cl);
int mpc_closed_loop(mpc_handle *h, const mpc_closed_loop_config *cl, uint32_t N, const float *d_x0,
                    const float *d_ee_path, const float *d_xu_traj, float *d_xu, float *d_lambda, float *d_states,
                    float *d_ee_err, uint32_t *d_pcg_iters, mpc_closed_loop_stats *stats, mpc_stream_t stream);

/* ---- cost weights and the differentiable QP layer --------------------------
 * Neither has a counterpart in the reference (its QD_COST / R_COST are
 * scalar compile-time macros, settings.cuh, perform_experiment.sh:44; it has
 * no differentiation).  Stage cost with per-dimension weights:
 *   0.5 sum_j wx_j x_j^2 + 0.5 sum_j wu_j u_j^2 + 0.5 q_ee |ee(q) - ref|^2;
 * mpc_create and mpc_set_config set the reference's cost, wx = (0 x 7,
 * qd_cost x 7) and wu = (r_cost x 7).  Weights must be finite and >= 0. */
int mpc_set_cost_weights(mpc_handle *h, const float *wx14, const float *wu7);
int mpc_get_cost_weights(mpc_handle *h, float *wx14, float *wu7);

/* QP layer at the iterate d_xu of every instance of the batch layout: the
 * KKT system of the SQP step (KKT blocks, Schur complement with the
 * instance's rho, PCG from the warm start d_lambda with exit tolerance
 * pcg_tol, <= 0: the configured one, and the dz recovery) and its
 * full-step solution d_xstar = d_xu - d_dz (both in the xu layout).
 * refine_steps > 0 adds that many steps of iterative refinement: the KKT
 * residual is evaluated in fp64 from the fp32 blocks and solved again with
 * the same Schur complement and preconditioner, each correction's PCG
 * exiting at refine_tol (<= 0: 1e-6) relative to its first r^T Pinv r.
 * d_lambda returns the refined multipliers. */
int mpc_batch_qp_layer(mpc_handle *h, const float *d_xu, const float *d_ee_ref, const float *d_rho,
                       float *d_lambda, float pcg_tol, uint32_t refine_steps, float refine_tol, float *d_dz,
                       float *d_xstar, mpc_stream_t stream);

/* Tracking loss L_i = sum_k |ee(q_k) - ref_k|^2 (xyz) of each instance of the
 * trajectory d_xu (d_loss, n_inst floats), with dL/dxu (xu layout) and the
 * direct dL/dref (6 floats per knot, xyz used); the gradients may be NULL. */
int mpc_batch_qp_loss(mpc_handle *h, const float *d_xu, const float *d_ee_ref, float *d_loss, float *d_dLdx,
                      float *d_dLdref, mpc_stream_t stream);

/* Adjoint of mpc_batch_qp_layer at its solution d_dz, for d_a = dL/dx*
 * (xu layout): recomputes the linearisation at d_xu, solves
 * K v = [-a; 0] with the same Schur complement, preconditioner and PCG
 * (exit tolerance pcg_tol, <= 0: the configured one; refine_steps and
 * refine_tol as for the layer), and writes d_grad_w (21 floats per
 * instance: dL/dwx then dL/dwu) and ADDS the implicit part of dL/dref to
 * d_grad_ref (6 floats per knot). */
int mpc_batch_qp_adjoint(mpc_handle *h, const float *d_xu, const float *d_ee_ref, const float *d_rho,
                         const float *d_dz, const float *d_a, float pcg_tol, uint32_t refine_steps,
                         float refine_tol, float *d_grad_w, float *d_grad_ref, mpc_stream_t stream);

/* ---- test hook ---------------------------------------------------------------
 * No counterpart in the reference (its failures exit the process,
 * include/gpuassert.cuh:5-14).  With part_plus_one = q + 1 (q < 4), part q
 * of every PCG team (the multi-workgroup solves of 128 < N <= 512) posts its
 * exchange words under a wrong epoch tag, so its partners' bounded waits
 * run out: every team solve ends with the max-iter flag 2, the plain driver
 * returns MPC_ERR_PCG_FAULT and the batched solve marks the instance
 * MPC_ST_PCG_FAULT without taking its step, while every kernel exits
 * normally.  0 (the default) turns it off.  For failure-path tests only. */
int mpc_set_fault_injection(mpc_handle *h, uint32_t part_plus_one);

/* ---- bandwidth probes --------------------------------------------------------
 * No counterpart in the reference: the device's own HBM and LDS bandwidth,
 * against which bench.py prices the per-phase roofline of the SQP
 * iteration.  Stream-ordered; time them with events on `stream`.
 * mpc_bw_copy: d_dst = d_src, n floats with 128-bit loads and stores
 *   (n % 4 == 0, both pointers 16-byte aligned; else MPC_ERR_ARG); moves
 *   8 n bytes.
 * mpc_bw_lds: n_wg workgroups of 256 lanes each fill a 16 KB LDS block with
 *   the words 0, 1, 2, ... and every lane reads reps x 4 consecutive
 *   128-bit words of it (word (lane + 256 k + r) mod 1024); reads
 *   64 x 256 x reps bytes per workgroup.  d_out (n_wg x 256 uint32) gets,
 *   per lane, (sum of the x and y parts) XOR (sum of the z and w parts) of
 *   the words it read, mod 2^32. */
int mpc_bw_copy(const float *d_src, float *d_dst, size_t n, mpc_stream_t stream);
int mpc_bw_lds(uint32_t *d_out, uint32_t n_wg, uint32_t reps, mpc_stream_t stream);

/* ---- synthetic workload (seeded, reproducible) ----------------------------- */

/* Joint-space data of one synthetic instance: splitmix64 stream seeded with
 * seed; q0 ~ U(-0.5,0.5)^7, qd0 ~ N(0,0.1^2)^7, f ~ U(0.2,1)^7 Hz,
 * phi ~ U(0,2 pi)^7; q_ref(k) = q0 + 0.3 sin(2 pi f k dt + phi), k < N.
 * Writes x0 (14) and q_ref (7N).  Host-only, no GPU needed. */
void mpc_synth_instance(uint64_t seed, uint32_t N, float dt, float *h_x0, float *h_qref);

/* The joint path of mpc_synth_instance with its derivatives: per knot
 * k < K, [q_ref, qd_ref, qdd_ref] (21 floats) at t = k dt; x0 as there.
 * Host-only. */
void mpc_synth_track(uint64_t seed, uint32_t K, float dt, float *h_x0, float *h_qqdqdd);

/* Horizons drawn uniformly from choices[0..n_choices) with a splitmix64
 * stream seeded with seed.  Host-only. */
void mpc_synth_horizons(uint64_t seed, uint32_t n_inst, const uint32_t *choices,
                        uint32_t n_choices, uint32_t *horizons);

#ifdef __cplusplus
}
#endif
#endif /* MPCPCG_H */
