// k_dynamics.hip -- linearisation (KKT blocks), forward kinematics,
// dynamics gradient and merit kernels for gfx950.
//
// Mapping: one 16-lane group per knot or point; k_dyn_grad runs 4 groups
// per 64-lane workgroup (one wave, so its barriers are single-wave),
// k_form_kkt 16 per 256-lane workgroup, with its single-lane steps one knot
// per lane (below).  Within a group the lanes split the rigid-body work by
// direction:
//   lanes 0..6   column j of M(q) (unit-acceleration recursion)
//   lane  7      bias forces c(q, qd) (RNEA at qdd = 0)
//   lanes 0..13  directional derivatives d tau / d(q_j | qd_j)
//   lanes 0..6   Jacobian column j of the end-effector position
// Shared per-knot quantities live in the group's LDS slice.
//
// Reference counterparts: gato_form_kkt (schur.cuh:391-524) with
// gato_plant::forwardDynamicsAndGradient (rbdfiles/iiwa_eepos_plant.cuh:
// 126-156), integratorAndGradient (integrator.cuh:159-183),
// trackingCostGradientAndHessian(_lastblock) (rbdfiles/iiwa_eepos_plant.cuh:
// 293-401), compute_merit / ls_gato_compute_merit (merit.cuh:17-145) and
// grid::end_effector_positions_kernel (rbdfiles/iiwa_eepos_grid.cuh:2141).
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int GROUP = 16;
constexpr int GROUPS_PER_WG = 4;
constexpr int WG = GROUP * GROUPS_PER_WG;

struct DynLds {
  float sn[8], cs[8];
  float M[49];     // joint-space inertia (column-major)
  float bias[8];   // c(q, qd)
  float Lc[49];    // Cholesky factor of M (column-major, lower)
  float Minv[49];  // M^-1 (column-major)
  float qdd[8];
  float bv[42], ba[42], bf[42], bxv[42], bxa[42];  // base RNEA at (q, qd, qdd)
  float df[98];    // dqdd/d(q,qd) (7 x 14 column-major)
};

struct KktLds {
  DynLds dyn;
  float x[16], u[8], xn[16], xc[16];  // x_k, u_k, x_{k+1}, cost state
  float ref[4];
  float csn[8], ccs[8];                // sin / cos of the cost state
  float zax[21], org[21];
  float gq[8];
  float ee[4];
};

// Forward dynamics and its gradient for one point (q, qd, u) held by a
// 16-lane group.  On return (after a barrier): D.qdd, D.Minv, D.df hold
// qdd, M^-1 and [dqdd/dq, dqdd/dqd].
__device__ __forceinline__ void group_dyn_grad(int lane, const float *q, const float *qd,
                                               const float *u, DynLds &D) {
  if (lane < NQ) {
    float s, c;
    sincosf(q[lane], &s, &c);
    D.sn[lane] = s;
    D.cs[lane] = c;
  }
  __syncthreads();
  float s[NQ], c[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    s[i] = D.sn[i];
    c[i] = D.cs[i];
  }
  // one RNEA on lanes 0..7 with lane-dependent inputs (no divergent paths):
  // lane j < 7: (qd, qdd) = (0, e_j) gives column j of M (the velocity terms
  // are exact zeros); lane 7: (qd, 0) gives the bias forces c(q, qd)
  if (lane <= NQ) {
    float qdv[NQ], qddv[NQ], tau[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      qdv[i] = lane == NQ ? qd[i] : 0.f;
      qddv[i] = i == lane ? 1.f : 0.f;
    }
    rnea<false>(s, c, qdv, qddv, tau, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (lane < NQ) {
#pragma unroll
      for (int r = 0; r < NQ; ++r) D.M[r + 7 * lane] = tau[r];
    } else {
#pragma unroll
      for (int i = 0; i < NQ; ++i) D.bias[i] = tau[i];
    }
  }
  __syncthreads();
  if (lane == 0) cholesky7(D.M, D.Lc);
  __syncthreads();
  if (lane < NQ) {
    float L[49];
#pragma unroll
    for (int i = 0; i < 49; ++i) L[i] = D.Lc[i];
    float e[NQ], x[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) e[i] = (i == lane) ? 1.f : 0.f;
    chol_solve7(L, e, x);
#pragma unroll
    for (int i = 0; i < NQ; ++i) D.Minv[i + 7 * lane] = x[i];
  }
  __syncthreads();
  if (lane < NQ) {
    float acc = 0.f;
#pragma unroll
    for (int m = 0; m < NQ; ++m) acc += D.Minv[lane + 7 * m] * (u[m] - D.bias[m]);
    D.qdd[lane] = acc;
  }
  __syncthreads();
  if (lane == 0) {
    float qdv[NQ], qddv[NQ], tau[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      qdv[i] = qd[i];
      qddv[i] = D.qdd[i];
    }
    rnea<true>(s, c, qdv, qddv, tau, D.bv, D.ba, D.bf, D.bxv, D.bxa);
  }
  __syncthreads();
  float dtau[NQ];
  if (lane < NX) {
    float qdv[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) qdv[i] = qd[i];
    rnea_dir(lane, s, c, qdv, D.bv, D.bf, D.bxv, D.bxa, dtau);
  }
  __syncthreads();  // every lane is done reading the base point
  if (lane < NX) {
    // dqdd/d(dir) = -M^-1 dtau
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      float acc = 0.f;
#pragma unroll
      for (int m = 0; m < NQ; ++m) acc += D.Minv[r + 7 * m] * dtau[m];
      D.df[r + 7 * lane] = -acc;
    }
  }
  __syncthreads();
}

// --------------------------------------------------------------------------
// forward kinematics: one thread per point
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fk(uint32_t n, const float *__restrict__ q, uint32_t q_stride,
                                            float *__restrict__ ee) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s[NQ], c[NQ];
#pragma unroll
  for (int j = 0; j < NQ; ++j) sincosf(q[(size_t)i * q_stride + j], &s[j], &c[j]);
  float zax[21], org[21], R[9];
  fk_frames(s, c, zax, org, R);
  float *o = ee + 6 * (size_t)i;
  o[0] = org[18];
  o[1] = org[19];
  o[2] = org[20];
  o[3] = atan2f(R[7], R[8]);
  o[4] = -atan2f(R[6], sqrtf(R[7] * R[7] + R[8] * R[8]));
  o[5] = atan2f(R[3], R[0]);
}

// --------------------------------------------------------------------------
// dynamics gradient of n points [q qd u]
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(WG) k_dyn_grad(uint32_t n, const float *__restrict__ qqdu,
                                                 float *__restrict__ qdd_out,
                                                 float *__restrict__ df_out) {
  __shared__ DynLds lds[GROUPS_PER_WG];
  __shared__ float pin[GROUPS_PER_WG][24];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t pt = blockIdx.x * GROUPS_PER_WG + grp;
  const bool valid = pt < n;
  const uint32_t ptc = valid ? pt : n - 1;
  for (int i = lane; i < NXU; i += GROUP) pin[grp][i] = qqdu[(size_t)ptc * NXU + i];
  __syncthreads();
  DynLds &D = lds[grp];
  group_dyn_grad(lane, pin[grp], pin[grp] + NQ, pin[grp] + NX, D);
  if (!valid) return;
  if (lane < NQ) qdd_out[(size_t)pt * NQ + lane] = D.qdd[lane];
  float *o = df_out + (size_t)pt * 147;
  for (int i = lane; i < 98; i += GROUP) o[i] = D.df[i];
  for (int i = lane; i < 49; i += GROUP) o[98 + i] = D.Minv[i];
}

// --------------------------------------------------------------------------
// inverse dynamics tau = M(q) qdd + c(q, qd) of n points [q qd qdd] (stride
// 21), one thread per point (grid::inverse_dynamics_kernel,
// rbdfiles/iiwa_eepos_grid.cuh, with the plant's gravity 0)
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_inverse_dynamics(uint32_t n, const float *__restrict__ qqdqdd,
                                                          float *__restrict__ tau) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *p = qqdqdd + (size_t)NXU * i;
  float s[NQ], c[NQ], qd[NQ], qdd[NQ], t[NQ];
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    sincosf(p[j], &s[j], &c[j]);
    qd[j] = p[NQ + j];
    qdd[j] = p[NX + j];
  }
  rnea<false>(s, c, qd, qdd, t, nullptr, nullptr, nullptr, nullptr, nullptr);
#pragma unroll
  for (int j = 0; j < NQ; ++j) tau[(size_t)NQ * i + j] = t[j];
}

// --------------------------------------------------------------------------
// KKT blocks, one group per global knot, 16 knots per 256-lane workgroup
// --------------------------------------------------------------------------
// The steps of group_dyn_grad, with its single-lane steps (the Cholesky
// factor of M, the RNEA base point) and the frames of the cost state run
// one knot per lane for all 16 knots of the workgroup: one instruction
// stream per 16 knots instead of one per 4.  Each knot's arithmetic is the
// same function on the same inputs as in group_dyn_grad.
constexpr int KKT_GROUPS = 16;
constexpr int KKT_WG = GROUP * KKT_GROUPS;
__global__ void __launch_bounds__(KKT_WG) k_form_kkt(Batch b, Params p, const float *__restrict__ xu,
                                                     const float *__restrict__ ee_ref,
                                                     const float *__restrict__ xs,
                                                     const uint32_t *__restrict__ status,
                                                     float *__restrict__ G, float *__restrict__ C,
                                                     float *__restrict__ g, float *__restrict__ cvec,
                                                     float *__restrict__ gq_out, float *__restrict__ Cl) {
  __shared__ KktLds lds[KKT_GROUPS];
  const int tid = threadIdx.x, grp = tid / GROUP, lane = tid % GROUP, w = tid >> 6, l64 = tid & 63;
  const uint32_t gk_raw = blockIdx.x * KKT_GROUPS + grp;
  bool valid = gk_raw < b.total_knots;
  const uint32_t gk = valid ? gk_raw : b.total_knots - 1;
  const KnotEnds ke = knot_ends(b, gk);
  if (status && !(status[ke.inst] & ST_ACTIVE)) valid = false;
  KktLds &L = lds[grp];
  DynLds &D = L.dyn;
  const float *xk = xu + (size_t)NXU * gk - (size_t)NU * ke.inst;  // x_k, u_k, x_{k+1}
  const bool dyn = ke.has_next;
  // the cost state: x_k, or at the last knot x_{k-1} unless its own state is asked for
  const float *xc = (dyn || (p.flags & FLAG_LAST_KNOT_OWN_STATE)) ? xk : xk - NXU;
  for (int i = lane; i < NX; i += GROUP) {
    L.x[i] = xk[i];
    L.xc[i] = xc[i];
    L.xn[i] = dyn ? xk[NXU + i] : 0.f;
  }
  for (int i = lane; i < NU; i += GROUP) L.u[i] = dyn ? xk[NX + i] : 0.f;
  if (lane < 3) L.ref[lane] = ee_ref[(size_t)6 * gk + lane];
  __syncthreads();

  // dynamics linearisation (every group runs it so barriers line up; the
  // terminal knot's results are discarded), and sin / cos of the cost state.
  // The phases that need at most eight lanes per knot run two knots to a
  // group: groups 0..7 (waves 0 and 1) serve knot 2g + h on lanes 8h ..
  // 8h + 7 while waves 2 and 3 wait, half the instructions of one knot per
  // group.
  const float *q = L.x, *qd = L.x + NQ;
  const bool packed = grp < KKT_GROUPS / 2;
  const int pl = lane & 7;
  KktLds &P = lds[packed ? 2 * grp + (lane >> 3) : grp];
  if (packed && pl < NQ) {
    float sv, cv;
    sincosf(P.x[pl], &sv, &cv);
    P.dyn.sn[pl] = sv;
    P.dyn.cs[pl] = cv;
  }
  // the cost state is x_k except at an instance's last knot (x_{k-1} there
  // unless FLAG_LAST_KNOT_OWN_STATE): its sin / cos are copied below
  const bool cost_own = dyn || (p.flags & FLAG_LAST_KNOT_OWN_STATE);
  if (!cost_own && lane < NQ) {
    float sv, cv;
    sincosf(L.xc[lane], &sv, &cv);
    L.csn[lane] = sv;
    L.ccs[lane] = cv;
  }
  __syncthreads();
  if (cost_own && lane < NQ) {
    L.csn[lane] = D.sn[lane];
    L.ccs[lane] = D.cs[lane];
  }
  // one RNEA on lanes 0..7 of a knot with lane-dependent inputs: lane j < 7:
  // (qd, qdd) = (0, e_j) gives column j of M; lane 7: (qd, 0) the bias
  // forces c(q, qd)
  if (packed) {
    float ps[NQ], pc[NQ], qdv[NQ], qddv[NQ], tau[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      ps[i] = P.dyn.sn[i];
      pc[i] = P.dyn.cs[i];
      qdv[i] = pl == NQ ? P.x[NQ + i] : 0.f;
      qddv[i] = i == pl ? 1.f : 0.f;
    }
    rnea<false>(ps, pc, qdv, qddv, tau, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (pl < NQ) {
#pragma unroll
      for (int r = 0; r < NQ; ++r) P.dyn.M[r + 7 * pl] = tau[r];
    } else {
#pragma unroll
      for (int i = 0; i < NQ; ++i) P.dyn.bias[i] = tau[i];
    }
  }
  __syncthreads();
  // one knot per lane: wave 0 factors M, wave 1 forms the cost state's frames
  if (w == 0 && l64 < KKT_GROUPS) cholesky7(lds[l64].dyn.M, lds[l64].dyn.Lc);
  if (w == 1 && l64 < KKT_GROUPS) {
    KktLds &K = lds[l64];
    float sk[NQ], ck[NQ], R[9];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      sk[i] = K.csn[i];
      ck[i] = K.ccs[i];
    }
    fk_frames(sk, ck, K.zax, K.org, R);
    K.ee[0] = K.org[18];
    K.ee[1] = K.org[19];
    K.ee[2] = K.org[20];
  }
  __syncthreads();
  // M^-1 by columns, two knots to a group as above
  if (packed && pl < NQ) {
    float Lf[49];
#pragma unroll
    for (int i = 0; i < 49; ++i) Lf[i] = P.dyn.Lc[i];
    float e[NQ], x[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) e[i] = (i == pl) ? 1.f : 0.f;
    chol_solve7(Lf, e, x);
#pragma unroll
    for (int i = 0; i < NQ; ++i) P.dyn.Minv[i + 7 * pl] = x[i];
  }
  __syncthreads();
  if (lane < NQ) {
    float acc = 0.f;
#pragma unroll
    for (int m = 0; m < NQ; ++m) acc += D.Minv[lane + 7 * m] * (L.u[m] - D.bias[m]);
    D.qdd[lane] = acc;
  }
  __syncthreads();
  // one knot per lane: the RNEA base point at (q, qd, qdd)
  if (w == 0 && l64 < KKT_GROUPS) {
    KktLds &K = lds[l64];
    float sk[NQ], ck[NQ], qdv[NQ], qddv[NQ], tau[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      sk[i] = K.dyn.sn[i];
      ck[i] = K.dyn.cs[i];
      qdv[i] = K.x[NQ + i];
      qddv[i] = K.dyn.qdd[i];
    }
    rnea<true>(sk, ck, qdv, qddv, tau, K.dyn.bv, K.dyn.ba, K.dyn.bf, K.dyn.bxv, K.dyn.bxa);
  }
  __syncthreads();
  float dtau[NQ];
  if (lane < NX) {
    float s[NQ], c[NQ], qdv[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      s[i] = D.sn[i];
      c[i] = D.cs[i];
      qdv[i] = qd[i];
    }
    rnea_dir(lane, s, c, qdv, D.bv, D.bf, D.bxv, D.bxa, dtau);
  }
  __syncthreads();  // every lane is done reading the base point
  if (lane < NX) {
    // dqdd/d(dir) = -M^-1 dtau
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      float acc = 0.f;
#pragma unroll
      for (int m = 0; m < NQ; ++m) acc += D.Minv[r + 7 * m] * dtau[m];
      D.df[r + 7 * lane] = -acc;
    }
  }
  // tracking cost gradient at the cost state: J^T Q_ee (ee - ref)
  if (lane < NQ) {
    const float d0 = L.ee[0] - L.org[3 * lane], d1 = L.ee[1] - L.org[3 * lane + 1],
                d2 = L.ee[2] - L.org[3 * lane + 2];
    const float z0 = L.zax[3 * lane], z1 = L.zax[3 * lane + 1], z2 = L.zax[3 * lane + 2];
    const float j0 = z1 * d2 - z2 * d1, j1 = z2 * d0 - z0 * d2, j2 = z0 * d1 - z1 * d0;
    const float e0 = p.q_ee * (L.ee[0] - L.ref[0]), e1 = p.q_ee * (L.ee[1] - L.ref[1]),
                e2 = p.q_ee * (L.ee[2] - L.ref[2]);
    L.gq[lane] = j0 * e0 + j1 * e1 + j2 * e2;
  }
  __syncthreads();
  if (!valid) return;

  const size_t knot = gk;
  const bool compact = Cl != nullptr;  // G and C in the compact form (gq, Cl)
  float *Gk = G + (size_t)G_STRIDE * knot;
  if (!compact)
    for (int i = lane; i < NXX; i += GROUP) {
      const int r = i % NX, col = i / NX;
      const float w = r == col ? p.wx[r] : 0.f;
      Gk[i] = (r < NQ && col < NQ) ? L.gq[r] * L.gq[col] + w : w;
    }
  float *gk_ = g + (size_t)NXU * knot;
  if (lane < NX) gk_[lane] = (lane < NQ ? L.gq[lane] : 0.f) + p.wx[lane] * L.xc[lane];
  if (gq_out && lane < NQ) gq_out[(size_t)NQ * knot + lane] = L.gq[lane];  // Q = gq gq^T + diag(wx)
  if (ke.first && lane < NX) cvec[(size_t)NX * gk + lane] = L.x[lane] - xs[(size_t)NX * ke.inst + lane];
  if (!dyn) return;
  if (!compact)
    for (int i = lane; i < 49; i += GROUP) Gk[NXX + i] = (i % NU == i / NU) ? p.wu[i % NU] : 0.f;
  if (lane < NU) gk_[NX + lane] = p.wu[lane] * L.u[lane];
  const float dt = p.dt;
  if (compact) {
    // the lower block rows only (the upper ones are cl_upper_a and zero),
    // lane l its column l: 7 contiguous floats of A's (l < 14) and of B's
    float *Clk = Cl + (size_t)CL_STRIDE * knot;
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NQ; ++r) {
        float v = (NQ + r == lane) ? 1.f : 0.f;
        v += dt * L.dyn.df[r + NQ * lane];
        Clk[r + NQ * lane] = -v;
      }
    }
    if (lane < NU) {
#pragma unroll
      for (int r = 0; r < NQ; ++r) Clk[CL_A + r + NQ * lane] = -dt * L.dyn.Minv[r + NQ * lane];
    }
  } else {
    float *Ck = C + (size_t)C_STRIDE * knot;
    for (int i = lane; i < NXX; i += GROUP) {
      const int r = i % NX, col = i / NX;
      float v = (r == col) ? 1.f : 0.f;
      if (r < NQ && col == r + NQ) v += dt;
      if (r >= NQ) v += dt * L.dyn.df[(r - NQ) + NQ * col];
      Ck[i] = -v;
    }
    for (int i = lane; i < NX * NU; i += GROUP) {
      const int r = i % NX, col = i / NX;
      Ck[NXX + i] = r >= NQ ? -dt * L.dyn.Minv[(r - NQ) + NQ * col] : 0.f;
    }
  }
  if (lane < NX) {
    float v;
    if (lane < NQ) v = L.xn[lane] - (L.x[lane] + dt * L.x[NQ + lane]);
    else v = L.xn[lane] - (L.x[lane] + dt * L.dyn.qdd[lane - NQ]);
    cvec[(size_t)NX * (knot + 1) + lane] = v;
  }
}

// --------------------------------------------------------------------------
// merit terms, one lane per (alpha, global knot)
// --------------------------------------------------------------------------
// A wave serves 64 consecutive global knots of one step length.  Their
// trial points x_k, u_k, x_{k+1} form one contiguous range of the batch's xu
// (instances lie back to back), staged once through LDS with coalesced
// loads as xu + alpha dz; each lane then evaluates its knot alone: the lower
// triangle of M, the bias forces, the end-effector position, the Cholesky
// solve for qdd and the cost, with no divergence between lanes.
constexpr int MERIT_NT = 64;
constexpr int MERIT_SPAN = NXU * (MERIT_NT - 1) + NXU + NX;  // floats staged per wave

__global__ void __launch_bounds__(MERIT_NT) k_merit(Batch b, Params p, const float *__restrict__ xu,
                                                    const float *__restrict__ dz,
                                                    const float *__restrict__ ee_ref,
                                                    const float *__restrict__ xs,
                                                    const uint32_t *__restrict__ status, int p0,
                                                    float *__restrict__ out) {
  __shared__ float Z[MERIT_SPAN];
  const uint32_t lane = threadIdx.x;
  const uint32_t T = b.total_knots;
  const uint32_t gk_raw = blockIdx.x * MERIT_NT + lane;
  const bool valid = gk_raw < T;
  const uint32_t gk = valid ? gk_raw : T - 1;
  const uint32_t slot = blockIdx.y;
  const int pidx = p0 + (int)slot;
  const float alpha = (pidx >= 8) ? 0.f : -1.f / (float)(1u << pidx);
  const bool with_x0 = pidx < 8;
  const KnotEnds ke = knot_ends(b, gk);
  const uint32_t o = NXU * gk - NU * ke.inst;  // x_k in the batch xu
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)o);
  const uint32_t n_xu = NXU * T - NU * b.n_inst;
  const uint32_t n_stage = n_xu - lo < (uint32_t)MERIT_SPAN ? n_xu - lo : (uint32_t)MERIT_SPAN;
  for (uint32_t j = lane; j < n_stage; j += MERIT_NT) Z[j] = dz ? xu[lo + j] + alpha * dz[lo + j] : xu[lo + j];
  __syncthreads();
  if (!valid) return;
  const bool active = !(status && !(status[ke.inst] & ST_ACTIVE));
  if (!active) return;
  const float *z = Z + (o - lo);
  const bool dyn = ke.has_next;
  float s[NQ], c[NQ], ee[3];
#pragma unroll
  for (int i = 0; i < NQ; ++i) sincosf(z[i], &s[i], &c[i]);
  fk_pos(s, c, ee);
  // cost: 0.5 sum wx x^2 + 0.5 sum wu u^2 + 0.5 w |ee - ref|^2 (reference: wx = (0, QD), wu = R)
  float J = 0.f;
#pragma unroll
  for (int i = 0; i < NX; ++i) J += 0.5f * (p.wx[i] * z[i] * z[i]);
  if (dyn) {
#pragma unroll
    for (int i = 0; i < NU; ++i) J += 0.5f * (p.wu[i] * z[NX + i] * z[NX + i]);
  }
  const float *r = ee_ref + (size_t)6 * gk;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float e = ee[i] - r[i];
    J += 0.5f * p.q_ee * e * e;
  }
  float ck = 0.f;
  if (dyn) {
    float M[49], Lc[49], bias[NQ], zero[NQ], qd[NQ], rhs[NQ], qdd[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) mass_column_lower(j, s, c, M + 7 * j);
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      qd[i] = z[NQ + i];
      zero[i] = 0.f;
    }
    rnea<false>(s, c, qd, zero, bias, nullptr, nullptr, nullptr, nullptr, nullptr);
    cholesky7(M, Lc);
#pragma unroll
    for (int i = 0; i < NQ; ++i) rhs[i] = z[NX + i] - bias[i];
    chol_solve7(Lc, rhs, qdd);
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      ck += fabsf(z[NXU + i] - (z[i] + p.dt * z[NQ + i]));
      ck += fabsf(z[NXU + NQ + i] - (z[NQ + i] + p.dt * qdd[i]));
    }
  } else if (with_x0) {
    // the terminal slot also carries the initial-state constraint of the instance
    const size_t k0 = b.knot_inst ? b.inst_k0[ke.inst] : 0u;  // the instance's first knot
    const size_t o0 = (size_t)NXU * k0 - (size_t)NU * ke.inst;
    const float *x0s = xs + (size_t)NX * ke.inst;
    for (int i = 0; i < NX; ++i) {
      const float z0 = dz ? xu[o0 + i] + alpha * dz[o0 + i] : xu[o0 + i];
      ck += fabsf(z0 - x0s[i]);
    }
  }
  out[(size_t)slot * T + gk] = J + p.mu * ck;
}

}  // namespace

hipError_t launch_fk(uint32_t n, const float *q, uint32_t q_stride, float *ee, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fk, dim3((n + 255) / 256), dim3(256), 0, st, n, q, q_stride, ee);
  return hipGetLastError();
}

hipError_t launch_inverse_dynamics(uint32_t n, const float *qqdqdd, float *tau, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_inverse_dynamics, dim3((n + 255) / 256), dim3(256), 0, st, n, qqdqdd, tau);
  return hipGetLastError();
}

hipError_t launch_dyn_grad(uint32_t n, const float *qqdu, float *qdd, float *df_du, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dyn_grad, dim3((n + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG), 0, st, n,
                     qqdu, qdd, df_du);
  return hipGetLastError();
}

hipError_t launch_form_kkt(const Batch &b, const Params &p, const float *xu, const float *ee_ref,
                           const float *xs, const uint32_t *status, float *G, float *C, float *g,
                           float *c, hipStream_t st, float *gq, float *Cl) {
  if (b.total_knots == 0) return hipSuccess;
  if (Cl && !gq) return hipErrorInvalidValue;  // the compact form needs gq for G
  hipLaunchKernelGGL(k_form_kkt, dim3((b.total_knots + KKT_GROUPS - 1) / KKT_GROUPS), dim3(KKT_WG), 0, st, b, p,
                     xu, ee_ref, xs, status, G, C, g, c, gq, Cl);
  return hipGetLastError();
}

// G and C of every knot from gq and the compact dynamics rows, with the
// expressions of k_form_kkt; one 16-lane group per knot
__global__ void __launch_bounds__(WG) k_expand_gc(Batch b, Params p, const float *__restrict__ gq,
                                                  const float *__restrict__ Cl, float *__restrict__ G,
                                                  float *__restrict__ C) {
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t gk = blockIdx.x * GROUPS_PER_WG + grp;
  if (gk >= b.total_knots) return;
  const KnotEnds ke = knot_ends(b, gk);
  const size_t knot = gk;
  const float *gqk = gq + (size_t)NQ * knot;
  float *Gk = G + (size_t)G_STRIDE * knot;
  for (int i = lane; i < NXX; i += GROUP) {
    const int r = i % NX, col = i / NX;
    const float w = r == col ? p.wx[r] : 0.f;
    Gk[i] = (r < NQ && col < NQ) ? gqk[r] * gqk[col] + w : w;
  }
  if (!ke.has_next) return;
  for (int i = lane; i < 49; i += GROUP) Gk[NXX + i] = (i % NU == i / NU) ? p.wu[i % NU] : 0.f;
  const float *Clk = Cl + (size_t)CL_STRIDE * knot;
  float *Ck = C + (size_t)C_STRIDE * knot;
  for (int i = lane; i < NXX; i += GROUP) {
    const int r = i % NX, col = i / NX;
    Ck[i] = r < NQ ? cl_upper_a(r, col, p.dt) : Clk[(r - NQ) + NQ * col];
  }
  for (int i = lane; i < NX * NU; i += GROUP) {
    const int r = i % NX, col = i / NX;
    Ck[NXX + i] = r < NQ ? 0.f : Clk[CL_A + (r - NQ) + NQ * col];
  }
}

hipError_t launch_expand_gc(const Batch &b, const Params &p, const float *gq, const float *Cl, float *G, float *C,
                            hipStream_t st) {
  if (b.total_knots == 0) return hipSuccess;
  if (!gq || !Cl || !G || !C) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_expand_gc, dim3((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG), 0, st, b, p,
                     gq, Cl, G, C);
  return hipGetLastError();
}

hipError_t launch_merit(const Batch &b, const Params &p, const float *xu, const float *dz,
                        const float *ee_ref, const float *xs, const uint32_t *status, int p0, int np,
                        float *out, hipStream_t st) {
  if (b.total_knots == 0 || np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_merit, dim3((b.total_knots + MERIT_NT - 1) / MERIT_NT, (uint32_t)np), dim3(MERIT_NT), 0, st,
                     b, p, xu, dz, ee_ref, xs, status, p0, out);
  return hipGetLastError();
}

}  // namespace mpc
