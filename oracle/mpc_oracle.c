/*
 * mpc_oracle.c -- CPU restatement of the SQP tracking-MPC inner loop of
 * EmreAdabag/gpu-pcg-tracking, used ONLY as test infrastructure: the parity
 * checker for tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py.  Nothing in the product links or calls this file.
 *
 * The file is precision-generic: it is compiled twice, with REAL=double
 * (suffix _d) and REAL=float (suffix _f).  Every public function carries
 * the suffix.
 *
 * What each part follows in the reference (paths relative to the reference
 * tree) and how it is pinned (see tests/test_oracle_*.py):
 *   - model constants: rbdfiles/iiwa_eepos_grid.cuh:909-1672 (X, I, Xhom);
 *     joint transforms as in :1770-1851 and :1857-1914.
 *     Pinned by the reference golden FK file testfiles/0_0_eepos.traj.
 *   - FK / FK gradient: rbdfiles/iiwa_eepos_grid.cuh:2015-2099, :2255-2351.
 *     Pinned by the golden file and fp64 central differences.
 *   - RNEA / mass matrix / FD / FD gradient: rbdfiles/iiwa_eepos_grid.cuh
 *     :2743-2972, :3753-4184, :4351-4362, :2975-3198, :4558-4969 and
 *     rbdfiles/iiwa_eepos_plant.cuh:126-156.  Pinned by invariants (M
 *     symmetric positive definite, M = sum J_i^T I_i J_i from independent
 *     numpy kinematics, Lagrangian identity for the bias term, FD(ID(qdd)) =
 *     qdd) and fp64 central differences of FD.
 *   - Euler integrator: integrator.cuh:25-126, :186-219.  Closed form,
 *     pinned by finite differences of the Euler map.
 *   - KKT blocks: schur.cuh:391-524, rbdfiles/iiwa_eepos_plant.cuh:293-401.
 *   - Schur complement + block-Jacobi: oldschur.cuh:518-963.  Pinned by a
 *     dense numpy assembly of -C G^-1 C^T and a dense solve of the KKT system.
 *   - PCG: GPU-PCG submodule (absent from the reference snapshot; call site
 *     toplevel.cuh:143-157, :292-294).  Exit rule |r^T Pinv r| < tol as
 *     documented in the package; pinned by convergence to numpy.linalg.solve.
 *   - dz recovery: oldschur.cuh:966-1105.  Pinned by the dense KKT solve.
 *   - merit / line search / rho rule: merit.cuh:17-145, toplevel.cuh:234-446.
 *     Pinned by brute-force numpy evaluation from the pinned FK / FD.
 *   - differentiable QP layer and its adjoint (growth direction of the
 *     build; the reference has none): pinned by fp64 central differences.
 *   - closed-loop harness: toplevel.cuh:505-775, integrator.cuh:283-398,
 *     restated with one knot shift per control step.  Pinned by the noise
 *     statistics, the Euler plant against orc_fd, the shift invariants and
 *     convergence to a stationary reachable target.
 *   - sparse LDL^T (QDLDL-equivalent, published up-looking algorithm) with the
 *     reference's CSC pattern: utils.cuh:54-118, qdldl_helper.cuh:62-180,
 *     toplevel.cuh:160-215.  Pinned by L D L^T = S reconstruction and
 *     numpy.linalg.solve.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#define SFX d
#endif
typedef REAL real;

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SFX)

#define NQ 7
#define NX 14
#define NU 7
#define NXU 21
#define NXX 196

/* ------------------------------------------------------------------------
 * Model data.  Joint i: fixed parent->joint rotation RF[i] (row-major,
 * entries in {0,+-1}) and origin PF[i] (parent coordinates), then a rotation
 * q_i about the joint z axis.  Link i inertial parameters about the link
 * frame origin: {m, m*cx, m*cy, m*cz, Ixx, Iyy, Izz, Ixy, Ixz, Iyz}.
 * Values follow rbdfiles/iiwa_eepos_grid.cuh:909-1672.
 * ---------------------------------------------------------------------- */
static const double RF[7][9] = {
    {1, 0, 0, 0, 1, 0, 0, 0, 1},
    {-1, 0, 0, 0, 0, 1, 0, 1, 0},
    {-1, 0, 0, 0, 0, 1, 0, 1, 0},
    {1, 0, 0, 0, 0, -1, 0, 1, 0},
    {-1, 0, 0, 0, 0, 1, 0, 1, 0},
    {1, 0, 0, 0, 0, -1, 0, 1, 0},
    {-1, 0, 0, 0, 0, 1, 0, 1, 0},
};
static const double PF[7][3] = {
    {0, 0, 0.1575}, {0, 0, 0.2025}, {0, 0.2045, 0}, {0, 0, 0.2155},
    {0, 0.1845, 0}, {0, 0, 0.2155}, {0, 0.081, 0},
};
static const double LINK[7][10] = {
    {5.76, 0.0, -0.1728, 0.6912, 0.209128, 0.198944, 0.022684, 0.0, 0.0, -0.000264},
    {6.35, 0.001905, 0.37465, 0.2667, 0.09710575, 0.0528019715, 0.0552049215, -1.2395e-05, -1e-08, -3.53e-05},
    {3.5, 0.0, 0.105, 0.455, 0.1496, 0.14215, 0.01395, 0.0, 0.0, 0.00035},
    {3.5, 0.0, 0.2345, 0.119, 0.0565575, 0.024496, 0.0374215, 0.0, 0.0, 2.7e-05},
    {3.5, 0.00035, 0.0735, 0.266, 0.0535595, 0.049132035, 0.007543535, -3.5e-07, 4e-07, 0.0},
    {1.8, 0.0, 0.00108, 0.00072, 0.004900936, 0.004700288, 0.003600648, 0.0, 0.0, -4.32e-07},
    {1.2, 0.0, 0.0, 0.024, 0.00598, 0.00598, 0.005, 0.0, 0.0, 0.0},
};

/* ------------------------------------------------------------------------
 * small vector helpers
 * ---------------------------------------------------------------------- */
static void cross3(const real *a, const real *b, real *o) {
  real x = a[1] * b[2] - a[2] * b[1];
  real y = a[2] * b[0] - a[0] * b[2];
  real z = a[0] * b[1] - a[1] * b[0];
  o[0] = x; o[1] = y; o[2] = z;
}

/* Joint i transform: parent<-child rotation R (row-major) for angle q. */
static void joint_rot(int i, real s, real c, real *R) {
  /* R = RF * Rz(q); Rz = [c -s 0; s c 0; 0 0 1] */
  for (int r = 0; r < 3; ++r) {
    real a = (real)RF[i][3 * r + 0], b = (real)RF[i][3 * r + 1], d = (real)RF[i][3 * r + 2];
    R[3 * r + 0] = a * c + b * s;
    R[3 * r + 1] = -a * s + b * c;
    R[3 * r + 2] = d;
  }
}

/* Spatial transform X_i (child <- parent) is kept as E = R^T and r = PF[i].
 * Motion vectors are [w; v], force vectors [n; f]. */
typedef struct { real E[9]; real r[3]; } xform;

static void make_xforms(const real *q, xform *X) {
  for (int i = 0; i < NQ; ++i) {
    real R[9];
    joint_rot(i, (real)sin(q[i]), (real)cos(q[i]), R);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) X[i].E[3 * a + b] = R[3 * b + a];
    for (int a = 0; a < 3; ++a) X[i].r[a] = (real)PF[i][a];
  }
}

static void mat3v(const real *M, const real *v, real *o) {
  real x = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
  real y = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
  real z = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
  o[0] = x; o[1] = y; o[2] = z;
}
static void mat3tv(const real *M, const real *v, real *o) {
  real x = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
  real y = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
  real z = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
  o[0] = x; o[1] = y; o[2] = z;
}

/* X m = [E w; E (v - r x w)] */
static void X_mv(const xform *X, const real *m, real *o) {
  real t[3], u[3];
  cross3(X->r, m, t);
  u[0] = m[3] - t[0]; u[1] = m[4] - t[1]; u[2] = m[5] - t[2];
  real w[3];
  mat3v(X->E, m, w);
  mat3v(X->E, u, o + 3);
  o[0] = w[0]; o[1] = w[1]; o[2] = w[2];
}
/* X^T f = [E^T n + r x (E^T f); E^T f] */
static void XT_fv(const xform *X, const real *f, real *o) {
  real n[3], ff[3], t[3];
  mat3tv(X->E, f, n);
  mat3tv(X->E, f + 3, ff);
  cross3(X->r, ff, t);
  o[0] = n[0] + t[0]; o[1] = n[1] + t[1]; o[2] = n[2] + t[2];
  o[3] = ff[0]; o[4] = ff[1]; o[5] = ff[2];
}
/* motion cross: v x m = [w x wm; w x vm + v x wm] */
static void crm(const real *v, const real *m, real *o) {
  real a[3], b[3], c[3];
  cross3(v, m, a);
  cross3(v, m + 3, b);
  cross3(v + 3, m, c);
  o[0] = a[0]; o[1] = a[1]; o[2] = a[2];
  o[3] = b[0] + c[0]; o[4] = b[1] + c[1]; o[5] = b[2] + c[2];
}
/* force cross: v x* f = [w x n + v x f; w x f] */
static void crf(const real *v, const real *f, real *o) {
  real a[3], b[3], c[3];
  cross3(v, f, a);
  cross3(v + 3, f + 3, b);
  cross3(v, f + 3, c);
  o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2];
  o[3] = c[0]; o[4] = c[1]; o[5] = c[2];
}
/* I m = [Io w + mc x v; m v - mc x w] */
static void inertia_mv(int i, const real *m, real *o) {
  const double *L = LINK[i];
  real mass = (real)L[0];
  real mc[3] = {(real)L[1], (real)L[2], (real)L[3]};
  real Io[9] = {(real)L[4], (real)L[7], (real)L[8],
                (real)L[7], (real)L[5], (real)L[9],
                (real)L[8], (real)L[9], (real)L[6]};
  real a[3], b[3], c[3];
  mat3v(Io, m, a);
  cross3(mc, m + 3, b);
  cross3(mc, m, c);
  o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2];
  o[3] = mass * m[3] - c[0]; o[4] = mass * m[4] - c[1]; o[5] = mass * m[5] - c[2];
}

/* ------------------------------------------------------------------------
 * Forward kinematics (xyz + rpy of the link-7 frame origin), its xyz
 * gradient.  rbdfiles/iiwa_eepos_grid.cuh:2015-2099, :2255-2351.
 * ---------------------------------------------------------------------- */
static void fk_frames(const real *q, real Rw[8][9], real pw[8][3]) {
  for (int a = 0; a < 9; ++a) Rw[0][a] = (a % 4 == 0) ? 1 : 0;
  pw[0][0] = pw[0][1] = pw[0][2] = 0;
  for (int i = 0; i < NQ; ++i) {
    real R[9];
    joint_rot(i, (real)sin(q[i]), (real)cos(q[i]), R);
    real t[3], p[3] = {(real)PF[i][0], (real)PF[i][1], (real)PF[i][2]};
    mat3v(Rw[i], p, t);
    for (int a = 0; a < 3; ++a) pw[i + 1][a] = pw[i][a] + t[a];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        Rw[i + 1][3 * r + c] = Rw[i][3 * r + 0] * R[0 + c] + Rw[i][3 * r + 1] * R[3 + c] +
                               Rw[i][3 * r + 2] * R[6 + c];
  }
}

void F(orc_fk)(const real *q, real *ee) {
  real Rw[8][9], pw[8][3];
  fk_frames(q, Rw, pw);
  const real *R = Rw[7];
  ee[0] = pw[7][0]; ee[1] = pw[7][1]; ee[2] = pw[7][2];
  ee[3] = (real)atan2(R[7], R[8]);
  ee[4] = -(real)atan2(R[6], (real)sqrt(R[7] * R[7] + R[8] * R[8]));
  ee[5] = (real)atan2(R[3], R[0]);
}

/* J (3 x 7, column-major): d ee_xyz / d q_j = z_j x (p_ee - p_j). */
void F(orc_fk_grad)(const real *q, real *J) {
  real Rw[8][9], pw[8][3];
  fk_frames(q, Rw, pw);
  for (int j = 0; j < NQ; ++j) {
    real z[3] = {Rw[j + 1][2], Rw[j + 1][5], Rw[j + 1][8]};
    real d[3] = {pw[7][0] - pw[j + 1][0], pw[7][1] - pw[j + 1][1], pw[7][2] - pw[j + 1][2]};
    cross3(z, d, J + 3 * j);
  }
}

/* ------------------------------------------------------------------------
 * Rigid-body dynamics, gravity 0 (rbdfiles/iiwa_eepos_plant.cuh:49-51).
 * ---------------------------------------------------------------------- */
typedef struct {
  xform X[NQ];
  real v[NQ][6], a[NQ][6], f[NQ][6]; /* f: total after the backward pass */
  real Xv[NQ][6], Xa[NQ][6];         /* X_i v_{i-1}, X_i a_{i-1} */
} rnea_state;

static void rnea_full(const real *q, const real *qd, const real *qdd, real *tau, rnea_state *st) {
  rnea_state loc;
  rnea_state *s = st ? st : &loc;
  make_xforms(q, s->X);
  real vp[6] = {0, 0, 0, 0, 0, 0}, ap[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < NQ; ++i) {
    real t[6], Iv[6], c[6];
    X_mv(&s->X[i], vp, s->Xv[i]);
    X_mv(&s->X[i], ap, s->Xa[i]);
    for (int k = 0; k < 6; ++k) { s->v[i][k] = s->Xv[i][k]; s->a[i][k] = s->Xa[i][k]; }
    s->v[i][2] += qd[i];
    s->a[i][2] += qdd[i];
    real Sq[6] = {0, 0, qd[i], 0, 0, 0};
    crm(s->v[i], Sq, t);
    for (int k = 0; k < 6; ++k) s->a[i][k] += t[k];
    inertia_mv(i, s->a[i], s->f[i]);
    inertia_mv(i, s->v[i], Iv);
    crf(s->v[i], Iv, c);
    for (int k = 0; k < 6; ++k) s->f[i][k] += c[k];
    memcpy(vp, s->v[i], sizeof vp);
    memcpy(ap, s->a[i], sizeof ap);
  }
  for (int i = NQ - 1; i >= 0; --i) {
    tau[i] = s->f[i][2];
    if (i > 0) {
      real t[6];
      XT_fv(&s->X[i], s->f[i], t);
      for (int k = 0; k < 6; ++k) s->f[i - 1][k] += t[k];
    }
  }
}

void F(orc_rnea)(const real *q, const real *qd, const real *qdd, real *tau) {
  rnea_full(q, qd, qdd, tau, NULL);
}

/* M (7x7 column-major): column j = RNEA(q, 0, e_j) (gravity 0). */
void F(orc_mass_matrix)(const real *q, real *M) {
  real z[NQ] = {0}, e[NQ];
  for (int j = 0; j < NQ; ++j) {
    for (int k = 0; k < NQ; ++k) e[k] = (k == j);
    rnea_full(q, z, e, M + NQ * j, NULL);
  }
}

/* Cholesky-based inverse of an SPD n x n column-major matrix. */
static int spd_inverse(int n, const real *A, real *Ainv) {
  real L[NX * NX];
  memset(L, 0, sizeof(real) * n * n);
  for (int j = 0; j < n; ++j) {
    real d = A[j + n * j];
    for (int k = 0; k < j; ++k) d -= L[j + n * k] * L[j + n * k];
    if (!(d > 0)) return -1;
    d = (real)sqrt(d);
    L[j + n * j] = d;
    for (int i = j + 1; i < n; ++i) {
      real s = A[i + n * j];
      for (int k = 0; k < j; ++k) s -= L[i + n * k] * L[j + n * k];
      L[i + n * j] = s / d;
    }
  }
  for (int c = 0; c < n; ++c) {
    real y[NX];
    for (int i = 0; i < n; ++i) {
      real s = (i == c);
      for (int k = 0; k < i; ++k) s -= L[i + n * k] * y[k];
      y[i] = s / L[i + n * i];
    }
    for (int i = n - 1; i >= 0; --i) {
      real s = y[i];
      for (int k = i + 1; k < n; ++k) s -= L[k + n * i] * Ainv[k + n * c];
      Ainv[i + n * c] = s / L[i + n * i];
    }
  }
  return 0;
}

void F(orc_minv)(const real *q, real *Minv) {
  real M[NQ * NQ];
  F(orc_mass_matrix)(q, M);
  spd_inverse(NQ, M, Minv);
}

/* qdd = Minv (u - c(q, qd)) */
void F(orc_fd)(const real *q, const real *qd, const real *u, real *qdd) {
  real z[NQ] = {0}, c[NQ], Minv[NQ * NQ];
  rnea_full(q, qd, z, c, NULL);
  F(orc_minv)(q, Minv);
  for (int i = 0; i < NQ; ++i) {
    real s = 0;
    for (int k = 0; k < NQ; ++k) s += Minv[i + NQ * k] * (u[k] - c[k]);
    qdd[i] = s;
  }
}

/* Forward-mode directional derivative of RNEA at (q, qd, qdd).  dir < 7:
 * d/dq_dir; 7 <= dir < 14: d/dqd_(dir-7).  dtau = d tau / d dir. */
static void rnea_dir(const rnea_state *s, const real *qd, int dir, real *dtau) {
  int jq = dir < NQ ? dir : -1;
  int jd = dir >= NQ ? dir - NQ : -1;
  real dv[NQ][6], da[NQ][6], df[NQ][6];
  real dvp[6] = {0, 0, 0, 0, 0, 0}, dap[6] = {0, 0, 0, 0, 0, 0};
  const real Sz[6] = {0, 0, 1, 0, 0, 0};
  for (int i = 0; i < NQ; ++i) {
    real t[6];
    X_mv(&s->X[i], dvp, dv[i]);
    X_mv(&s->X[i], dap, da[i]);
    if (i == jq) {
      /* d(X_i m)/dq_i = -S x (X_i m) */
      crm(Sz, s->Xv[i], t);
      for (int k = 0; k < 6; ++k) dv[i][k] -= t[k];
      crm(Sz, s->Xa[i], t);
      for (int k = 0; k < 6; ++k) da[i][k] -= t[k];
    }
    if (i == jd) dv[i][2] += 1;
    real Sq[6] = {0, 0, qd[i], 0, 0, 0};
    crm(dv[i], Sq, t);
    for (int k = 0; k < 6; ++k) da[i][k] += t[k];
    if (i == jd) {
      crm(s->v[i], Sz, t);
      for (int k = 0; k < 6; ++k) da[i][k] += t[k];
    }
    real Iv[6], Idv[6], c1[6], c2[6];
    inertia_mv(i, da[i], df[i]);
    inertia_mv(i, s->v[i], Iv);
    inertia_mv(i, dv[i], Idv);
    crf(dv[i], Iv, c1);
    crf(s->v[i], Idv, c2);
    for (int k = 0; k < 6; ++k) df[i][k] += c1[k] + c2[k];
    memcpy(dvp, dv[i], sizeof dvp);
    memcpy(dap, da[i], sizeof dap);
  }
  for (int i = NQ - 1; i >= 0; --i) {
    dtau[i] = df[i][2];
    if (i > 0) {
      real t[6];
      XT_fv(&s->X[i], df[i], t);
      for (int k = 0; k < 6; ++k) df[i - 1][k] += t[k];
      if (i == jq) {
        /* (dX_i/dq_i)^T f_i = X_i^T (S x* f_i) */
        real sf[6];
        crf(Sz, s->f[i], sf);
        XT_fv(&s->X[i], sf, t);
        for (int k = 0; k < 6; ++k) df[i - 1][k] += t[k];
      }
    }
  }
}

/* dtau (7 x 14 column-major) = [d ID/dq, d ID/dqd] at (q, qd, qdd). */
void F(orc_rnea_grad)(const real *q, const real *qd, const real *qdd, real *dtau) {
  rnea_state s;
  real tau[NQ];
  rnea_full(q, qd, qdd, tau, &s);
  for (int d = 0; d < NX; ++d) rnea_dir(&s, qd, d, dtau + NQ * d);
}

/* qdd (7) and df_du (7 x 21 column-major) = [dqdd/dq, dqdd/dqd, dqdd/du]
 * (rbdfiles/iiwa_eepos_plant.cuh:126-156). */
void F(orc_fd_grad)(const real *q, const real *qd, const real *u, real *qdd, real *df_du) {
  real Minv[NQ * NQ], dtau[NQ * NX];
  F(orc_fd)(q, qd, u, qdd);
  F(orc_minv)(q, Minv);
  F(orc_rnea_grad)(q, qd, qdd, dtau);
  for (int c = 0; c < NX; ++c)
    for (int r = 0; r < NQ; ++r) {
      real s = 0;
      for (int k = 0; k < NQ; ++k) s += Minv[r + NQ * k] * dtau[k + NQ * c];
      df_du[r + NQ * c] = -s;
    }
  for (int c = 0; c < NQ; ++c)
    for (int r = 0; r < NQ; ++r) df_du[r + NQ * (NX + c)] = Minv[r + NQ * c];
}

/* ------------------------------------------------------------------------
 * Cost / solver configuration shared by the stage functions.
 * ---------------------------------------------------------------------- */
typedef struct {
  double dt;        /* integrator step */
  double q_ee;      /* end-effector position weight (reference: 1) */
  double qd_cost;   /* QD_COST */
  double r_cost;    /* R_COST */
  double mu;        /* merit constraint penalty (merit.cuh / toplevel.cuh:54: 10) */
  double rho_factor, rho_max, rho_min;
  double pcg_tol;
  int pcg_max_iter;
  int sqp_max_iter;
  int last_knot_own_state; /* 0: reference quirk (knot N-1 cost uses x_{N-2}) */
  /* per-dimension quadratic weights 0.5 sum_j wx_j x_j^2 + 0.5 sum_j wu_j u_j^2;
   * the reference's cost is wx = (0 x 7, QD x 7), wu = (R x 7) */
  double wx[14];
  double wu[7];
  int pcg_stair;    /* 1: stair preconditioner instead of block Jacobi */
} orc_cfg;

/* Cost gradient/Hessian at one knot (rbdfiles/iiwa_eepos_plant.cuh:293-378),
 * with per-dimension weights (reference: wx = (0, QD), wu = R):
 *   q_k = [Jq^T e + wx_q x_q ; wx_qd qd], Q_k = [g_e g_e^T + diag(wx_q), 0; 0, diag(wx_qd)]
 *   (g_e = Jq^T e, the rank-1 outer product of the end-effector gradient as in
 *   the reference), r = wu u, R_k = diag(wu). */
static void cost_grad_hess(const orc_cfg *cfg, const real *x, const real *u, const real *ref,
                           real *Q, real *qv, real *R, real *rv) {
  real ee[6], J[21];
  F(orc_fk)(x, ee);
  F(orc_fk_grad)(x, J);
  real w = (real)cfg->q_ee;
  real e[3] = {w * (ee[0] - ref[0]), w * (ee[1] - ref[1]), w * (ee[2] - ref[2])};
  real ge[NQ];
  for (int j = 0; j < NQ; ++j) ge[j] = J[3 * j] * e[0] + J[3 * j + 1] * e[1] + J[3 * j + 2] * e[2];
  for (int j = 0; j < NX; ++j) qv[j] = (j < NQ ? ge[j] : 0) + (real)cfg->wx[j] * x[j];
  for (int c = 0; c < NX; ++c)
    for (int r = 0; r < NX; ++r) {
      real val = (r == c) ? (real)cfg->wx[r] : 0;
      if (r < NQ && c < NQ) val += ge[r] * ge[c];
      Q[r + NX * c] = val;
    }
  if (R) {
    for (int j = 0; j < NU; ++j) rv[j] = (real)cfg->wu[j] * u[j];
    for (int c = 0; c < NU; ++c)
      for (int r = 0; r < NU; ++r) R[r + NU * c] = (r == c) ? (real)cfg->wu[r] : 0;
  }
}

/* KKT blocks (schur.cuh:391-524).  Layouts (per knot k):
 *   G: [Q_k (196) | R_k (49)] at 245k, Q_{N-1} at 245(N-1)
 *   g: [q_k (14) | r_k (7)] at 21k, q_{N-1} at 21(N-1)
 *   C: [-A_k (196) | -B_k (98)] at 294k
 *   c: c_0 = x_0 - x_s, c_{k+1} = x_{k+1} - (x_k + dt f(x_k, u_k)) */
void F(orc_form_kkt)(const orc_cfg *cfg, int N, const real *xu, const real *ee_ref, const real *xs,
                     real *G, real *C, real *g, real *c) {
  real dt = (real)cfg->dt;
  for (int k = 0; k < N - 1; ++k) {
    const real *x = xu + NXU * k, *u = x + NX, *xn = xu + NXU * (k + 1);
    real qdd[NQ], df[NQ * NXU];
    F(orc_fd_grad)(x, x + NQ, u, qdd, df);
    real *Cm = C + 294 * k;
    for (int col = 0; col < NX + NU; ++col)
      for (int r = 0; r < NX; ++r) {
        real val = (r == col) ? 1 : 0;
        if (r < NQ && col == r + NQ) val += dt;
        if (r >= NQ) val += dt * df[(r - NQ) + NQ * col];
        Cm[r + NX * col] = -val;
      }
    for (int i = 0; i < NQ; ++i) {
      c[NX * (k + 1) + i] = xn[i] - (x[i] + dt * x[NQ + i]);
      c[NX * (k + 1) + NQ + i] = xn[NQ + i] - (x[NQ + i] + dt * qdd[i]);
    }
    cost_grad_hess(cfg, x, u, ee_ref + 6 * k, G + 245 * k, g + NXU * k, G + 245 * k + NXX,
                   g + NXU * k + NX);
  }
  {
    int k = N - 1;
    const real *x = cfg->last_knot_own_state ? xu + NXU * k : xu + NXU * (k - 1);
    cost_grad_hess(cfg, x, NULL, ee_ref + 6 * k, G + 245 * k, g + NXU * k, NULL, NULL);
  }
  for (int i = 0; i < NX; ++i) c[i] = xu[i] - xs[i];
}

static real rmax(real a, real b) { return a > b ? a : b; }
static real rmin(real a, real b) { return a < b ? a : b; }

/* small dense helpers (column-major) */
static void mm(int m, int k, int n, const real *A, const real *B, real *C) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      real s = 0;
      for (int l = 0; l < k; ++l) s += A[i + m * l] * B[l + k * j];
      C[i + m * j] = s;
    }
}
/* C = A B^T, A m x k, B n x k */
static void mmt(int m, int k, int n, const real *A, const real *B, real *C) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      real s = 0;
      for (int l = 0; l < k; ++l) s += A[i + m * l] * B[j + n * l];
      C[i + m * j] = s;
    }
}
static void mv(int m, int n, const real *A, const real *x, real *y) {
  for (int i = 0; i < m; ++i) {
    real s = 0;
    for (int l = 0; l < n; ++l) s += A[i + m * l] * x[l];
    y[i] = s;
  }
}

/* Schur complement S = -C G^-1 C^T (block tridiagonal, rows of 3 blocks),
 * gamma, block-Jacobi Pinv and G^-1 (oldschur.cuh:518-963).
 * Returns 0, or -1 if a regularised block is not positive definite. */
int F(orc_form_schur)(int N, const real *G, const real *C, const real *g, const real *c, real rho,
                      real *S, real *Pinv, real *gamma, real *Ginv) {
  memset(S, 0, sizeof(real) * 588 * N);
  memset(Pinv, 0, sizeof(real) * 588 * N);
  /* regularised inverses of every Q_k and R_k */
  for (int k = 0; k < N; ++k) {
    real Qr[NXX];
    memcpy(Qr, G + 245 * k, sizeof Qr);
    for (int i = 0; i < NX; ++i) Qr[i + NX * i] += rho;
    if (spd_inverse(NX, Qr, Ginv + 245 * k)) return -1;
    if (k < N - 1) {
      real Rr[49];
      memcpy(Rr, G + 245 * k + NXX, sizeof Rr);
      for (int i = 0; i < NU; ++i) Rr[i + NU * i] += rho;
      if (spd_inverse(NU, Rr, Ginv + 245 * k + NXX)) return -1;
    }
    if (k == 0) {
      for (int i = 0; i < NXX; ++i) {
        S[NXX + i] = -Ginv[i];
        Pinv[NXX + i] = -Qr[i];
      }
      real t[NX];
      mv(NX, NX, Ginv, g, t);
      for (int i = 0; i < NX; ++i) gamma[i] = -t[i];
    }
  }
  for (int b = 1; b < N; ++b) {
    int k = b - 1;
    const real *QkI = Ginv + 245 * k, *RkI = Ginv + 245 * k + NXX, *Qk1I = Ginv + 245 * b;
    real A[NXX], B[98];
    for (int i = 0; i < NXX; ++i) A[i] = -C[294 * k + i];
    for (int i = 0; i < 98; ++i) B[i] = -C[294 * k + NXX + i];
    real AQ[NXX], BR[98], th[NXX], t1[NXX];
    mm(NX, NX, NX, A, QkI, AQ);
    mm(NX, NU, NU, B, RkI, BR);
    mmt(NX, NX, NX, AQ, A, th);
    mmt(NX, NU, NX, BR, B, t1);
    for (int i = 0; i < NXX; ++i) th[i] += t1[i] + Qk1I[i];
    for (int i = 0; i < NXX; ++i) {
      S[588 * b + i] = AQ[i];
      S[588 * b + NXX + i] = -th[i];
    }
    for (int r = 0; r < NX; ++r)
      for (int cc = 0; cc < NX; ++cc) S[588 * k + 2 * NXX + r + NX * cc] = AQ[cc + NX * r];
    real thI[NXX];
    if (spd_inverse(NX, th, thI)) return -1;
    for (int i = 0; i < NXX; ++i) Pinv[588 * b + NXX + i] = -thI[i];
    real v1[NX], v2[NX], v3[NX];
    mv(NX, NX, Qk1I, g + NXU * b, v1);
    mv(NX, NX, AQ, g + NXU * k, v2);
    mv(NX, NU, BR, g + NXU * k + NX, v3);
    for (int i = 0; i < NX; ++i) gamma[NX * b + i] = c[NX * b + i] - (v1[i] - v2[i] - v3[i]);
  }
  return 0;
}

/* y = S x for the 3-block-per-row layout. */
void F(orc_bd_mv)(int N, const real *S, const real *x, real *y) {
  for (int b = 0; b < N; ++b) {
    real acc[NX];
    mv(NX, NX, S + 588 * b + NXX, x + NX * b, acc);
    if (b > 0) {
      real t[NX];
      mv(NX, NX, S + 588 * b, x + NX * (b - 1), t);
      for (int i = 0; i < NX; ++i) acc[i] += t[i];
    }
    if (b < N - 1) {
      real t[NX];
      mv(NX, NX, S + 588 * b + 2 * NXX, x + NX * (b + 1), t);
      for (int i = 0; i < NX; ++i) acc[i] += t[i];
    }
    memcpy(y + NX * b, acc, sizeof acc);
  }
}

/* Preconditioned CG on S lambda = gamma with preconditioner Pinv (same
 * 3-block layout; the build writes only its diagonal blocks).  Warm start
 * from lambda.  Exit when |r^T Pinv r| < tol; the test is also applied once
 * before the first iteration (a warm start already inside the tolerance
 * returns 0 iterations).  *iters = number of S products taken in the loop;
 * *max_iter_exit = 1 if max_iter was reached without meeting the tolerance. */
/* Preconditioner application z = M r.  Block Jacobi (stair == 0): M = Pinv
 * (its block diagonal D^-1).  Stair (stair != 0; gato_form_ss_inner,
 * oldschur.cuh:334-494, where the reference stores the off-diagonal blocks
 * -D_b^-1 S(b,b-1) D_{b-1}^-1): M = D^-1 - D^-1 (L + L^T) D^-1 with L the
 * strictly lower block part of S, applied as z = D^-1 (2 r - S D^-1 r)
 * (since (L + L^T) D^-1 r = S D^-1 r - r).  t is n scratch values. */
static void F(orc_precond)(int N, const real *S, const real *Pinv, const real *r, real *z, real *t, int stair) {
  F(orc_bd_mv)(N, Pinv, r, z);
  if (!stair) return;
  F(orc_bd_mv)(N, S, z, t);
  for (int i = 0; i < NX * N; ++i) t[i] = 2 * r[i] - t[i];
  F(orc_bd_mv)(N, Pinv, t, z);
}

/* orc_pcg_trace: as orc_pcg_ex, and when eta_trace is not NULL it receives
 * |r^T z| before the first iteration and after each one (max_iter + 1
 * values; entries past the exit are left as they were). */
void F(orc_pcg_trace)(int N, const real *S, const real *Pinv, const real *gamma, real *lambda,
                      int max_iter, real tol, int stair, int *iters, int *max_iter_exit, real *eta_trace) {
  int n = NX * N;
  real *r = (real *)malloc(sizeof(real) * n), *z = (real *)malloc(sizeof(real) * n);
  real *p = (real *)malloc(sizeof(real) * n), *y = (real *)malloc(sizeof(real) * n);
  real *t = (real *)malloc(sizeof(real) * n);
  F(orc_bd_mv)(N, S, lambda, y);
  for (int i = 0; i < n; ++i) r[i] = gamma[i] - y[i];
  F(orc_precond)(N, S, Pinv, r, z, t, stair);
  real eta = 0;
  for (int i = 0; i < n; ++i) { p[i] = z[i]; eta += r[i] * z[i]; }
  int it = 0, conv = 0;
  if (eta_trace) eta_trace[0] = (real)fabs(eta);
  if ((real)fabs(eta) < tol) conv = 1;
  while (!conv && it < max_iter) {
    F(orc_bd_mv)(N, S, p, y);
    real py = 0;
    for (int i = 0; i < n; ++i) py += p[i] * y[i];
    real alpha = eta / py;
    for (int i = 0; i < n; ++i) { lambda[i] += alpha * p[i]; r[i] -= alpha * y[i]; }
    F(orc_precond)(N, S, Pinv, r, z, t, stair);
    real eta_new = 0;
    for (int i = 0; i < n; ++i) eta_new += r[i] * z[i];
    ++it;
    if (eta_trace) eta_trace[it] = (real)fabs(eta_new);
    if ((real)fabs(eta_new) < tol) { conv = 1; break; }
    real beta = eta_new / eta;
    eta = eta_new;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  *iters = it;
  *max_iter_exit = !conv;
  free(r); free(z); free(p); free(y); free(t);
}

void F(orc_pcg_ex)(int N, const real *S, const real *Pinv, const real *gamma, real *lambda,
                   int max_iter, real tol, int stair, int *iters, int *max_iter_exit) {
  F(orc_pcg_trace)(N, S, Pinv, gamma, lambda, max_iter, tol, stair, iters, max_iter_exit, NULL);
}

void F(orc_pcg)(int N, const real *S, const real *Pinv, const real *gamma, real *lambda,
                int max_iter, real tol, int *iters, int *max_iter_exit) {
  F(orc_pcg_ex)(N, S, Pinv, gamma, lambda, max_iter, tol, 0, iters, max_iter_exit);
}

/* dz (oldschur.cuh:966-1105):
 *   dx_k = Q_k^-1 (q_k - lambda_k + A_k^T lambda_{k+1})  (no A term at k = N-1)
 *   du_k = R_k^-1 (r_k + B_k^T lambda_{k+1})
 * with C holding -A_k, -B_k and Ginv the regularised inverses. */
void F(orc_compute_dz)(int N, const real *Ginv, const real *C, const real *g, const real *lambda,
                       real *dz) {
  for (int k = 0; k < N; ++k) {
    real t[NX];
    for (int i = 0; i < NX; ++i) {
      real s = 0;
      if (k < N - 1)
        for (int j = 0; j < NX; ++j) s += -C[294 * k + j + NX * i] * lambda[NX * (k + 1) + j];
      t[i] = g[NXU * k + i] - lambda[NX * k + i] + s;
    }
    mv(NX, NX, Ginv + 245 * k, t, dz + NXU * k);
    if (k < N - 1) {
      real tu[NU];
      for (int i = 0; i < NU; ++i) {
        real s = 0;
        for (int j = 0; j < NX; ++j) s += -C[294 * k + NXX + j + NX * i] * lambda[NX * (k + 1) + j];
        tu[i] = g[NXU * k + NX + i] + s;
      }
      mv(NU, NU, Ginv + 245 * k + NXX, tu, dz + NXU * k + NX);
    }
  }
}

/* Merit (merit.cuh:17-145 with rbdfiles/iiwa_eepos_plant.cuh:239-290 and
 * integrator.cuh:186-219) of the trajectory xu + alpha dz:
 *   sum_k [0.5 sum_j wx_j x_kj^2 + 0.5 sum_j wu_j u_kj^2 (k<N-1) + 0.5 w |ee(q_k) - ref_k|^2]
 *   (reference: wx = (0, QD), wu = R)
 *   + mu sum_{k<N-1} |x_{k+1} - x_k - dt f(x_k, u_k)|_1
 *   + (with_x0) mu |x_0 - x_s|_1
 * dz may be NULL (alpha ignored).  knot_out (N) receives per-knot terms. */
real F(orc_merit)(const orc_cfg *cfg, int N, const real *xu, const real *dz, real alpha,
                  const real *ee_ref, const real *xs, int with_x0, real *knot_out) {
  int len = NXU * N - NU;
  real *z = (real *)malloc(sizeof(real) * len);
  for (int i = 0; i < len; ++i) z[i] = xu[i] + (dz ? alpha * dz[i] : 0);
  real total = 0;
  real dt = (real)cfg->dt, mu = (real)cfg->mu;
  for (int k = 0; k < N; ++k) {
    const real *x = z + NXU * k;
    real J = 0;
    for (int i = 0; i < NX; ++i) J += (real)0.5 * ((real)cfg->wx[i] * x[i] * x[i]);
    if (k < N - 1)
      for (int i = 0; i < NU; ++i) J += (real)0.5 * ((real)cfg->wu[i] * x[NX + i] * x[NX + i]);
    real ee[6];
    F(orc_fk)(x, ee);
    for (int i = 0; i < 3; ++i) {
      real e = ee[i] - ee_ref[6 * k + i];
      J += (real)0.5 * (real)cfg->q_ee * e * e;
    }
    real ck = 0;
    if (k < N - 1) {
      real qdd[NQ];
      F(orc_fd)(x, x + NQ, x + NX, qdd);
      const real *xn = z + NXU * (k + 1);
      for (int i = 0; i < NQ; ++i) {
        ck += (real)fabs(xn[i] - (x[i] + dt * x[NQ + i]));
        ck += (real)fabs(xn[NQ + i] - (x[NQ + i] + dt * qdd[i]));
      }
    } else if (with_x0) {
      for (int i = 0; i < NX; ++i) ck += (real)fabs(z[i] - xs[i]);
    }
    real m = J + mu * ck;
    if (knot_out) knot_out[k] = m;
    total += m;
  }
  free(z);
  return total;
}

/* ------------------------------------------------------------------------
 * Sparse LDL^T with the reference's CSC pattern (upper triangle of S,
 * utils.cuh:84-118), up-looking factorisation after an elimination tree
 * (the published LDL / QDLDL algorithm).
 * ---------------------------------------------------------------------- */
#ifndef ORC_NO_COMMON
int orc_csc_nnz(int N) { return 196 * (N - 1) + 105 * N; }

void orc_csc_pattern(int N, int32_t *Ap, int32_t *Ai) {
  int off = 0;
  Ap[0] = 0;
  for (int b = 0; b < N; ++b)
    for (int r = 0; r < NX; ++r) {
      int first = b > 0 ? (b - 1) * NX : 0;
      int last = b * NX + r;
      for (int i = first; i <= last; ++i) Ai[off++] = i;
      Ap[b * NX + r + 1] = off;
    }
}
#else
int orc_csc_nnz(int N);
void orc_csc_pattern(int N, int32_t *Ap, int32_t *Ai);
#endif

/* CSC values of S in the pattern above (S in the 3-block layout). */
void F(orc_csc_values)(int N, const real *S, real *Ax) {
  int off = 0;
  for (int b = 0; b < N; ++b)
    for (int r = 0; r < NX; ++r) {
      if (b > 0)
        for (int cc = 0; cc < NX; ++cc) Ax[off++] = S[588 * b + r + NX * cc]; /* S[b,b-1](r, cc) */
      for (int cc = 0; cc <= r; ++cc) Ax[off++] = S[588 * b + NXX + r + NX * cc];
    }
}

#ifndef ORC_NO_COMMON
/* elimination tree; returns sum of L nonzeros */
int orc_ldl_etree(int n, const int32_t *Ap, const int32_t *Ai, int32_t *work, int32_t *Lnz,
                  int32_t *etree) {
  for (int i = 0; i < n; ++i) { work[i] = 0; Lnz[i] = 0; etree[i] = -1; }
  for (int j = 0; j < n; ++j) {
    work[j] = j;
    for (int p = Ap[j]; p < Ap[j + 1]; ++p) {
      int i = Ai[p];
      if (i > j) return -1;
      while (work[i] != j) {
        if (etree[i] == -1) etree[i] = j;
        Lnz[i]++;
        work[i] = j;
        i = etree[i];
      }
    }
  }
  int s = 0;
  for (int i = 0; i < n; ++i) s += Lnz[i];
  return s;
}
#else
int orc_ldl_etree(int n, const int32_t *Ap, const int32_t *Ai, int32_t *work, int32_t *Lnz,
                  int32_t *etree);
#endif

/* up-looking numeric factorisation: row k of L from the etree reach of
 * column k; D[k] = A_kk - sum L_ki^2 D_i.  Returns the number of positive D
 * entries or -1 on a zero pivot. */
int F(orc_ldl_factor)(int n, const int32_t *Ap, const int32_t *Ai, const real *Ax, int32_t *Lp,
                      int32_t *Li, real *Lx, real *D, real *Dinv, const int32_t *Lnz,
                      const int32_t *etree, int32_t *iwork, real *fwork) {
  int32_t *flag = iwork, *pattern = iwork + n, *Lnext = iwork + 2 * n;
  real *y = fwork;
  Lp[0] = 0;
  for (int i = 0; i < n; ++i) { Lp[i + 1] = Lp[i] + Lnz[i]; Lnext[i] = Lp[i]; flag[i] = -1; y[i] = 0; }
  int pos = 0;
  for (int k = 0; k < n; ++k) {
    int top = n;
    flag[k] = k;
    real dk = 0;
    for (int p = Ap[k]; p < Ap[k + 1]; ++p) {
      int i = Ai[p];
      if (i == k) { dk += Ax[p]; continue; }
      y[i] += Ax[p];
      int len = 0;
      while (flag[i] != k) {
        pattern[len++] = i;
        flag[i] = k;
        i = etree[i];
      }
      while (len > 0) pattern[--top] = pattern[--len];
    }
    for (; top < n; ++top) {
      int i = pattern[top];
      real yi = y[i];
      y[i] = 0;
      for (int p = Lp[i]; p < Lnext[i]; ++p) y[Li[p]] -= Lx[p] * yi;
      real lki = yi * Dinv[i];
      dk -= lki * yi;
      Li[Lnext[i]] = k;
      Lx[Lnext[i]] = lki;
      Lnext[i]++;
    }
    if (dk == 0) return -1;
    D[k] = dk;
    Dinv[k] = 1 / dk;
    if (dk > 0) pos++;
  }
  return pos;
}

void F(orc_ldl_solve)(int n, const int32_t *Lp, const int32_t *Li, const real *Lx, const real *Dinv,
                      real *x) {
  for (int i = 0; i < n; ++i)
    for (int p = Lp[i]; p < Lp[i + 1]; ++p) x[Li[p]] -= Lx[p] * x[i];
  for (int i = 0; i < n; ++i) x[i] *= Dinv[i];
  for (int i = n - 1; i >= 0; --i)
    for (int p = Lp[i]; p < Lp[i + 1]; ++p) x[i] -= Lx[p] * x[Li[p]];
}

/* One direct Schur solve S lambda = gamma (factor + solve).  Workspace is
 * allocated here; the timing helper below reuses it. */
typedef struct {
  int n, nnz, lnz;
  int32_t *Ap, *Ai, *Lp, *Li, *Lnz, *etree, *iwork;
  real *Ax, *Lx, *D, *Dinv, *fwork;
} F(orc_ldl_ws);

void *F(orc_ldl_create)(int N) {
  F(orc_ldl_ws) *w = (F(orc_ldl_ws) *)calloc(1, sizeof(*w));
  int n = NX * N, nnz = orc_csc_nnz(N);
  w->n = n; w->nnz = nnz;
  w->Ap = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  w->Ai = (int32_t *)malloc(sizeof(int32_t) * nnz);
  w->Ax = (real *)malloc(sizeof(real) * nnz);
  w->Lnz = (int32_t *)malloc(sizeof(int32_t) * n);
  w->etree = (int32_t *)malloc(sizeof(int32_t) * n);
  w->iwork = (int32_t *)malloc(sizeof(int32_t) * 3 * n);
  w->Lp = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  w->D = (real *)malloc(sizeof(real) * n);
  w->Dinv = (real *)malloc(sizeof(real) * n);
  w->fwork = (real *)malloc(sizeof(real) * n);
  orc_csc_pattern(N, w->Ap, w->Ai);
  w->lnz = orc_ldl_etree(n, w->Ap, w->Ai, w->iwork, w->Lnz, w->etree);
  w->Li = (int32_t *)malloc(sizeof(int32_t) * (w->lnz > 0 ? w->lnz : 1));
  w->Lx = (real *)malloc(sizeof(real) * (w->lnz > 0 ? w->lnz : 1));
  return w;
}

void F(orc_ldl_destroy)(void *p) {
  F(orc_ldl_ws) *w = (F(orc_ldl_ws) *)p;
  if (!w) return;
  free(w->Ap); free(w->Ai); free(w->Ax); free(w->Lnz); free(w->etree); free(w->iwork);
  free(w->Lp); free(w->D); free(w->Dinv); free(w->fwork); free(w->Li); free(w->Lx);
  free(w);
}

int F(orc_ldl_schur_solve)(void *p, int N, const real *S, const real *gamma, real *lambda) {
  F(orc_ldl_ws) *w = (F(orc_ldl_ws) *)p;
  F(orc_csc_values)(N, S, w->Ax);
  int r = F(orc_ldl_factor)(w->n, w->Ap, w->Ai, w->Ax, w->Lp, w->Li, w->Lx, w->D, w->Dinv, w->Lnz,
                            w->etree, w->iwork, w->fwork);
  if (r < 0) return r;
  for (int i = 0; i < w->n; ++i) lambda[i] = gamma[i];
  F(orc_ldl_solve)(w->n, w->Lp, w->Li, w->Lx, w->Dinv, lambda);
  return 0;
}

/* ------------------------------------------------------------------------
 * SQP driver (toplevel.cuh:24-503) with PCG (or the direct solve when
 * use_direct != 0).  xu, lambda and *rho are updated in place.  Outputs:
 * sqp_iters, per-iteration PCG iteration counts / max-iter flags (up to
 * sqp_max_iter entries), the accepted line-search steps (-1 = failure), and
 * whether the solve ended without a rho abort.
 * ---------------------------------------------------------------------- */
int F(orc_sqp_solve)(const orc_cfg *cfg, int N, const real *ee_ref, real *lambda, real *xu,
                     real *rho, real rho_reset, int use_direct, int *sqp_iters_out,
                     int *pcg_iters_out, int *pcg_maxiter_out, int *ls_step_out,
                     int *not_aborted_out) {
  int len = NXU * N - NU;
  real *G = (real *)malloc(sizeof(real) * 245 * N), *Ginv = (real *)malloc(sizeof(real) * 245 * N);
  real *C = (real *)malloc(sizeof(real) * 294 * N), *g = (real *)malloc(sizeof(real) * NXU * N);
  real *c = (real *)malloc(sizeof(real) * NX * N), *S = (real *)malloc(sizeof(real) * 588 * N);
  real *P = (real *)malloc(sizeof(real) * 588 * N), *gam = (real *)malloc(sizeof(real) * NX * N);
  real *dz = (real *)malloc(sizeof(real) * NXU * N);
  void *ldl = use_direct ? F(orc_ldl_create)(N) : NULL;
  real xs[NX];
  memcpy(xs, xu, sizeof xs);
  real merit0 = F(orc_merit)(cfg, N, xu, NULL, 0, ee_ref, xs, 0, NULL);
  real drho = 1, rf = (real)cfg->rho_factor, rho_lo = (real)cfg->rho_min;
  int sqp_iter = 0, not_aborted = 1, rc = 0;
  for (int it = 0; it < cfg->sqp_max_iter; ++it) {
    F(orc_form_kkt)(cfg, N, xu, ee_ref, xs, G, C, g, c);
    if (F(orc_form_schur)(N, G, C, g, c, *rho, S, P, gam, Ginv)) { rc = -1; break; }
    if (use_direct) {
      if (F(orc_ldl_schur_solve)(ldl, N, S, gam, lambda)) { rc = -2; break; }
      if (pcg_iters_out) pcg_iters_out[it] = 0;
      if (pcg_maxiter_out) pcg_maxiter_out[it] = 0;
    } else {
      int pit, pmx;
      F(orc_pcg_ex)(N, S, P, gam, lambda, cfg->pcg_max_iter, (real)cfg->pcg_tol, cfg->pcg_stair, &pit, &pmx);
      if (pcg_iters_out) pcg_iters_out[it] = pit;
      if (pcg_maxiter_out) pcg_maxiter_out[it] = pmx;
    }
    F(orc_compute_dz)(N, Ginv, C, g, lambda, dz);
    real merits[8];
    for (int pidx = 0; pidx < 8; ++pidx) {
      real alpha = (real)(-1.0 / (double)(1 << pidx));
      merits[pidx] = F(orc_merit)(cfg, N, xu, dz, alpha, ee_ref, xs, 1, NULL);
    }
    int step = 0;
    real minm = merit0;
    for (int i = 0; i < 8; ++i)
      if (merits[i] < minm) { minm = merits[i]; step = i; }
    if (minm == merit0) {
      if (ls_step_out) ls_step_out[it] = -1;
      drho = rmax(drho * rf, rf);
      *rho = rmax(*rho * drho, rho_lo);
      sqp_iter++;
      if (*rho > (real)cfg->rho_max) {
        not_aborted = 0;
        *rho = rho_reset;
        break;
      }
      continue;
    }
    if (ls_step_out) ls_step_out[it] = step;
    real alpha = (real)(-1.0 / (double)(1 << step));
    drho = rmin(drho / rf, 1 / rf);
    *rho = rmax(*rho * drho, rho_lo);
    for (int i = 0; i < len; ++i) xu[i] += alpha * dz[i];
    sqp_iter++;
    merit0 = minm;
  }
  if (sqp_iters_out) *sqp_iters_out = sqp_iter;
  if (not_aborted_out) *not_aborted_out = not_aborted;
  if (ldl) F(orc_ldl_destroy)(ldl);
  free(G); free(Ginv); free(C); free(g); free(c); free(S); free(P); free(gam); free(dz);
  return rc;
}

/* ------------------------------------------------------------------------
 * Differentiable QP layer at a fixed iterate (the final SQP iterate xu):
 * K w = b with K = [G~ C^T; C 0] (G~ = G + rho I), w = [dz; lambda],
 * b = [g; c] formed at xu as above, and the full-step solution
 * x* = xu - dz.  Loss L = sum_k |ee(q*_k) - ref_k|^2 (xyz).  Adjoint:
 * v = K^-1 [-dL/dx*; 0], obtained through the same Schur complement (gamma
 * of g := -dL/dx*, c := 0, then the dz recovery), and for any cost
 * parameter theta (the weights wx, wu and the reference points; the
 * dynamics rows C, c do not depend on them)
 *     dL/dtheta = v_z^T (dg/dtheta - dG/dtheta dz)  (+ the direct ref term).
 * This is the restatement the GPU adjoint is checked against; it is itself
 * pinned by fp64 central differences of orc_qp_layer / orc_qp_loss.
 * ---------------------------------------------------------------------- */

/* x* = xu - dz of the QP at xu, solved with the direct LDL^T of S; dz and
 * lambda (may be NULL) receive the solution.  Returns 0 or a negative error. */
int F(orc_qp_layer)(const orc_cfg *cfg, int N, const real *xu, const real *ee_ref, const real *xs, real rho,
                    real *xstar, real *dz_out, real *lambda_out) {
  int len = NXU * N - NU, rc = 0;
  real *G = (real *)malloc(sizeof(real) * 245 * N), *Ginv = (real *)malloc(sizeof(real) * 245 * N);
  real *C = (real *)malloc(sizeof(real) * 294 * N), *g = (real *)malloc(sizeof(real) * NXU * N);
  real *c = (real *)malloc(sizeof(real) * NX * N), *S = (real *)malloc(sizeof(real) * 588 * N);
  real *P = (real *)malloc(sizeof(real) * 588 * N), *gam = (real *)malloc(sizeof(real) * NX * N);
  real *dz = (real *)malloc(sizeof(real) * NXU * N), *lam = (real *)malloc(sizeof(real) * NX * N);
  void *ldl = F(orc_ldl_create)(N);
  F(orc_form_kkt)(cfg, N, xu, ee_ref, xs, G, C, g, c);
  if (F(orc_form_schur)(N, G, C, g, c, rho, S, P, gam, Ginv)) rc = -1;
  if (!rc && F(orc_ldl_schur_solve)(ldl, N, S, gam, lam)) rc = -2;
  if (!rc) {
    F(orc_compute_dz)(N, Ginv, C, g, lam, dz);
    for (int i = 0; i < len; ++i) xstar[i] = xu[i] - dz[i];
    if (dz_out) memcpy(dz_out, dz, sizeof(real) * len);
    if (lambda_out) memcpy(lambda_out, lam, sizeof(real) * NX * N);
  }
  F(orc_ldl_destroy)(ldl);
  free(G); free(Ginv); free(C); free(g); free(c); free(S); free(P); free(gam); free(dz); free(lam);
  return rc;
}

/* L = sum_k |ee(q_k) - ref_k|^2 of the trajectory xu (xyz, every knot at its
 * own state); dLdx (21N-7) and dLdref (6N, xyz entries) receive the partial
 * derivatives (either may be NULL). */
real F(orc_qp_loss)(int N, const real *xu, const real *ee_ref, real *dLdx, real *dLdref) {
  real L = 0;
  if (dLdx) memset(dLdx, 0, sizeof(real) * (NXU * N - NU));
  if (dLdref) memset(dLdref, 0, sizeof(real) * 6 * N);
  for (int k = 0; k < N; ++k) {
    const real *x = xu + NXU * k;
    real ee[6], J[21];
    F(orc_fk)(x, ee);
    F(orc_fk_grad)(x, J);
    for (int a = 0; a < 3; ++a) {
      const real e = ee[a] - ee_ref[6 * k + a];
      L += e * e;
      if (dLdref) dLdref[6 * k + a] = -2 * e;
      if (dLdx)
        for (int j = 0; j < NQ; ++j) dLdx[NXU * k + j] += 2 * e * J[3 * j + a];
    }
  }
  return L;
}

/* Adjoint of orc_qp_layer at its solution dz: given a = dL/dx* (21N-7),
 * writes the weight gradient gw (14 wx then 7 wu) and ADDS the implicit
 * part of dL/dref to gref (6N, xyz entries).  Returns 0 or a negative error. */
int F(orc_qp_adjoint)(const orc_cfg *cfg, int N, const real *xu, const real *ee_ref, const real *xs, real rho,
                      const real *dz, const real *a, real *gw, real *gref) {
  int len = NXU * N - NU, rc = 0;
  real *G = (real *)malloc(sizeof(real) * 245 * N), *Ginv = (real *)malloc(sizeof(real) * 245 * N);
  real *C = (real *)malloc(sizeof(real) * 294 * N), *g = (real *)malloc(sizeof(real) * NXU * N);
  real *c = (real *)malloc(sizeof(real) * NX * N), *S = (real *)malloc(sizeof(real) * 588 * N);
  real *P = (real *)malloc(sizeof(real) * 588 * N), *gam = (real *)malloc(sizeof(real) * NX * N);
  real *az = (real *)malloc(sizeof(real) * NXU * N), *mu = (real *)malloc(sizeof(real) * NX * N);
  real *v = (real *)malloc(sizeof(real) * NXU * N), *c0 = (real *)calloc(NX * N, sizeof(real));
  void *ldl = F(orc_ldl_create)(N);
  F(orc_form_kkt)(cfg, N, xu, ee_ref, xs, G, C, g, c);
  for (int i = 0; i < len; ++i) az[i] = -a[i];
  if (F(orc_form_schur)(N, G, C, az, c0, rho, S, P, gam, Ginv)) rc = -1;
  if (!rc && F(orc_ldl_schur_solve)(ldl, N, S, gam, mu)) rc = -2;
  if (!rc) {
    F(orc_compute_dz)(N, Ginv, C, az, mu, v);
    for (int j = 0; j < 21; ++j) gw[j] = 0;
    for (int k = 0; k < N; ++k) {
      const real *xc = (k < N - 1 || cfg->last_knot_own_state) ? xu + NXU * k : xu + NXU * (k - 1);
      const real *vx = v + NXU * k, *dzx = dz + NXU * k;
      for (int j = 0; j < NX; ++j) gw[j] += vx[j] * (xc[j] - dzx[j]);
      if (k < N - 1)
        for (int j = 0; j < NU; ++j) gw[NX + j] += vx[NX + j] * (xu[NXU * k + NX + j] - dzx[NX + j]);
      real ee[6], J[21], ge[NQ];
      F(orc_fk)(xc, ee);
      F(orc_fk_grad)(xc, J);
      const real w = (real)cfg->q_ee;
      for (int j = 0; j < NQ; ++j) {
        ge[j] = 0;
        for (int b = 0; b < 3; ++b) ge[j] += J[3 * j + b] * w * (ee[b] - ee_ref[6 * k + b]);
      }
      real vge = 0, gedz = 0;
      for (int j = 0; j < NQ; ++j) {
        vge += vx[j] * ge[j];
        gedz += ge[j] * dzx[j];
      }
      for (int b = 0; b < 3; ++b) {
        real vd = 0, ddz = 0;  /* d ge / d ref_b = -w J[b, :] */
        for (int j = 0; j < NQ; ++j) {
          const real dge = -w * J[3 * j + b];
          vd += vx[j] * dge;
          ddz += dge * dzx[j];
        }
        gref[6 * k + b] += vd - vd * gedz - vge * ddz;
      }
    }
  }
  F(orc_ldl_destroy)(ldl);
  free(G); free(Ginv); free(C); free(g); free(c); free(S); free(P); free(gam); free(az); free(mu); free(v); free(c0);
  return rc;
}

/* ------------------------------------------------------------------------
 * Closed-loop tracking harness: track() (toplevel.cuh:505-775) with the
 * plant of integrator.cuh:283-398, restated with one knot shift per
 * control step.  Control step s:
 *   SQP (cfg->sqp_max_iter iterations) from the warm start xu, lambda, rho
 *   on the reference window path[s .. s+N-1];
 *   plant: `substeps` explicit Euler steps of sim_dt (q += h qd,
 *   qd += h qdd(q, qd, u_0)) under the plan's first control;
 *   noise: x += sigma n, pair i (components 2i, 2i+1) of step s from
 *   splitmix64 seeded with seed ^ 0xD1B54A32D192ED03 (8 s + i + 1), two
 *   draws u1 in (0, 1], u2 in [0, 1), Box-Muller r cos / r sin;
 *   error: |ee(q_{s+1}) - path[s+1]| (xyz, Euclidean);
 *   shift: x_k <- x_{k+1}, u_k <- u_{k+1} (k < N-2), lambda_k <- lambda_{k+1},
 *   x_0 <- plant state; the new tail (u_{N-2}, x_{N-1}) is the precomputed
 *   trajectory's (u_{s+N-1}, x_{s+N}) when traj (xu layout over steps + N
 *   knots) is given, else u_{N-2} <- 0 and x_{N-1} is kept.
 * states: 14 (steps + 1); ee_err: 2 per step (Euclidean and L1 norm of the
 * xyz error); pcg_iters: steps x sqp_max_iter (may be NULL).  Returns 0 or
 * the first SQP error.
 * ---------------------------------------------------------------------- */
static uint64_t cl_sm64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void F(orc_cl_noise)(uint64_t seed, uint32_t step, real *n14) {
  for (uint32_t i = 0; i < 7; ++i) {
    uint64_t s = seed ^ (0xD1B54A32D192ED03ull * (8ull * step + i + 1ull));
    const double u1 = (double)((cl_sm64(&s) >> 11) + 1ull) * 0x1.0p-53;
    const double u2 = (double)(cl_sm64(&s) >> 11) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1)), t = 6.283185307179586 * u2;
    n14[2 * i] = (real)(float)(r * cos(t));
    n14[2 * i + 1] = (real)(float)(r * sin(t));
  }
}

int F(orc_closed_loop)(const orc_cfg *cfg, int N, int steps, int substeps, real sim_dt, real sigma,
                       uint64_t seed, real rho_reset, const real *x0, const real *path, const real *traj,
                       real *xu, real *lambda, real *rho, real *states, real *ee_err, int *pcg_iters) {
  const int K = cfg->sqp_max_iter;
  real x[NX];
  memcpy(x, x0, sizeof x);
  memcpy(states, x0, sizeof x);
  memcpy(xu, x0, sizeof x);
  int *pit = (int *)calloc((size_t)(K > 0 ? K : 1), sizeof(int));
  int rc = 0;
  for (int s = 0; s < steps && !rc; ++s) {
    int si, na;
    rc = F(orc_sqp_solve)(cfg, N, path + 6 * s, lambda, xu, rho, rho_reset, 0, &si, pit, NULL, NULL, &na);
    if (pcg_iters)
      for (int j = 0; j < K; ++j) pcg_iters[K * s + j] = j < si ? pit[j] : 0;
    const real *u = xu + NX;
    for (int k = 0; k < substeps; ++k) {
      real qdd[NQ];
      F(orc_fd)(x, x + NQ, u, qdd);
      for (int i = 0; i < NQ; ++i) {
        const real q = x[i], qd = x[NQ + i];
        x[i] = q + sim_dt * qd;
        x[NQ + i] = qd + sim_dt * qdd[i];
      }
    }
    if (sigma > 0) {
      real n[NX];
      F(orc_cl_noise)(seed, (uint32_t)s, n);
      for (int i = 0; i < NX; ++i) x[i] += sigma * n[i];
    }
    memcpy(states + NX * (s + 1), x, sizeof x);
    real ee[6];
    F(orc_fk)(x, ee);
    const real *r = path + 6 * (s + 1);
    const real d0 = ee[0] - r[0], d1 = ee[1] - r[1], d2 = ee[2] - r[2];
    ee_err[2 * s] = (real)sqrt((double)(d0 * d0 + d1 * d1 + d2 * d2));
    ee_err[2 * s + 1] = (real)(fabs((double)d0) + fabs((double)d1) + fabs((double)d2));
    memmove(xu, xu + NXU, sizeof(real) * (size_t)(NXU * N - NU - NXU));
    if (traj)
      memcpy(xu + NXU * (N - 2) + NX, traj + (size_t)NXU * (s + N - 1) + NX, sizeof(real) * NXU);
    else
      for (int i = 0; i < NU; ++i) xu[NXU * (N - 2) + NX + i] = 0;
    memmove(lambda, lambda + NX, sizeof(real) * (size_t)(NX * (N - 1)));
    memcpy(xu, x, sizeof x);
  }
  free(pit);
  return rc;
}

/* ------------------------------------------------------------------------
 * Timing helper for the CPU baseline: reps x (CSC value fill + LDL^T factor
 * + solve) of one Schur system, as in the reference's host linear-system
 * step (qdldl_helper.cuh:62-180).  Returns elapsed seconds.
 * ---------------------------------------------------------------------- */
#include <time.h>
double F(orc_ldl_bench)(int N, const real *S, const real *gamma, int reps, real *lambda) {
  void *w = F(orc_ldl_create)(N);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int r = 0; r < reps; ++r) F(orc_ldl_schur_solve)(w, N, S, gamma, lambda);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  F(orc_ldl_destroy)(w);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
