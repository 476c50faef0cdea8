// model.h -- KUKA iiwa 7-DOF model data and per-lane spatial algebra for
// the gfx950 kernels.
//
// Model data follows rbdfiles/iiwa_eepos_grid.cuh:909-1672 (the X, I and
// Xhom tables) and the transform updates of :1770-1851 / :1857-1914, kept
// here as physical parameters instead of 6x6 tables:
//   joint i: fixed parent->joint rotation kRF[i] (entries 0/+-1) and origin
//            kPF[i] in parent coordinates, then a rotation q_i about z;
//   link i:  {m, m*c (3), rotational inertia about the link origin
//            (Ixx, Iyy, Izz, Ixy, Ixz, Iyz)}.
// The spatial transform X_i (child <- parent) is E = (kRF[i] Rz(q_i))^T with
// translation r = kPF[i]; motion vectors are [w; v], forces [n; f].  All
// helpers are written for fully unrolled loops over joints so that the 0/+-1
// rotation entries fold into the instruction stream: +-1 factors by the
// compiler's exact rules, zero factors through pa / ps (below), since x * 0
// is not foldable under IEEE rules (x may be infinite).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace mpc {

constexpr float kRF[7][9] = {
    {1, 0, 0, 0, 1, 0, 0, 0, 1},  {-1, 0, 0, 0, 0, 1, 0, 1, 0}, {-1, 0, 0, 0, 0, 1, 0, 1, 0},
    {1, 0, 0, 0, 0, -1, 0, 1, 0}, {-1, 0, 0, 0, 0, 1, 0, 1, 0}, {1, 0, 0, 0, 0, -1, 0, 1, 0},
    {-1, 0, 0, 0, 0, 1, 0, 1, 0},
};
constexpr float kPF[7][3] = {
    {0, 0, 0.1575f}, {0, 0, 0.2025f}, {0, 0.2045f, 0}, {0, 0, 0.2155f},
    {0, 0.1845f, 0}, {0, 0, 0.2155f}, {0, 0.081f, 0},
};
constexpr float kLink[7][10] = {
    {5.76f, 0.0f, -0.1728f, 0.6912f, 0.209128f, 0.198944f, 0.022684f, 0.0f, 0.0f, -0.000264f},
    {6.35f, 0.001905f, 0.37465f, 0.2667f, 0.09710575f, 0.0528019715f, 0.0552049215f, -1.2395e-05f,
     -1e-08f, -3.53e-05f},
    {3.5f, 0.0f, 0.105f, 0.455f, 0.1496f, 0.14215f, 0.01395f, 0.0f, 0.0f, 0.00035f},
    {3.5f, 0.0f, 0.2345f, 0.119f, 0.0565575f, 0.024496f, 0.0374215f, 0.0f, 0.0f, 2.7e-05f},
    {3.5f, 0.00035f, 0.0735f, 0.266f, 0.0535595f, 0.049132035f, 0.007543535f, -3.5e-07f, 4e-07f, 0.0f},
    {1.8f, 0.0f, 0.00108f, 0.00072f, 0.004900936f, 0.004700288f, 0.003600648f, 0.0f, 0.0f,
     -4.32e-07f},
    {1.2f, 0.0f, 0.0f, 0.024f, 0.00598f, 0.00598f, 0.005f, 0.0f, 0.0f, 0.0f},
};

// Products in which either factor may be a compile-time zero (a model
// constant, or a literal zero of a sparse vector): such a product is an
// exact zero for finite operands, so it is replaced by the zero that folds
// out of the expression it sits in, -0 where it is added (-0 + y = y) and
// +0 where it is subtracted (y - +0 = y); other products are plain.  Only
// the sign of an exactly zero result can differ from the full product.
__device__ __forceinline__ bool czero(float v) { return __builtin_constant_p(v) && v == 0.f; }
__device__ __forceinline__ float pa(float x, float y) { return czero(x) || czero(y) ? -0.f : x * y; }
__device__ __forceinline__ float ps(float x, float y) { return czero(x) || czero(y) ? 0.f : x * y; }

__device__ __forceinline__ void cross3(const float *a, const float *b, float *o) {
  float x = pa(a[1], b[2]) - ps(a[2], b[1]);
  float y = pa(a[2], b[0]) - ps(a[0], b[2]);
  float z = pa(a[0], b[1]) - ps(a[1], b[0]);
  o[0] = x;
  o[1] = y;
  o[2] = z;
}

// Parent<-child rotation R = kRF[i] Rz(q) (row-major) from (s, c) of q.
__device__ __forceinline__ void joint_rot(int i, float s, float c, float *R) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float a = kRF[i][3 * r + 0], b = kRF[i][3 * r + 1], d = kRF[i][3 * r + 2];
    R[3 * r + 0] = pa(a, c) + pa(b, s);
    R[3 * r + 1] = pa(-a, s) + pa(b, c);
    R[3 * r + 2] = d;
  }
}

// o = X_i m = [E w; E (v - r x w)],  E = R^T
__device__ __forceinline__ void X_mv(int i, float s, float c, const float *m, float *o) {
  float R[9];
  joint_rot(i, s, c, R);
  const float r[3] = {kPF[i][0], kPF[i][1], kPF[i][2]};
  float t[3];
  cross3(r, m, t);
  const float u0 = m[3] - t[0], u1 = m[4] - t[1], u2 = m[5] - t[2];
  const float w0 = pa(R[0], m[0]) + pa(R[3], m[1]) + pa(R[6], m[2]);
  const float w1 = pa(R[1], m[0]) + pa(R[4], m[1]) + pa(R[7], m[2]);
  const float w2 = pa(R[2], m[0]) + pa(R[5], m[1]) + pa(R[8], m[2]);
  o[3] = pa(R[0], u0) + pa(R[3], u1) + pa(R[6], u2);
  o[4] = pa(R[1], u0) + pa(R[4], u1) + pa(R[7], u2);
  o[5] = pa(R[2], u0) + pa(R[5], u1) + pa(R[8], u2);
  o[0] = w0;
  o[1] = w1;
  o[2] = w2;
}

// o = X_i^T f = [E^T n + r x (E^T f); E^T f]
__device__ __forceinline__ void XT_fv(int i, float s, float c, const float *f, float *o) {
  float R[9];
  joint_rot(i, s, c, R);
  const float r[3] = {kPF[i][0], kPF[i][1], kPF[i][2]};
  const float n0 = pa(R[0], f[0]) + pa(R[1], f[1]) + pa(R[2], f[2]);
  const float n1 = pa(R[3], f[0]) + pa(R[4], f[1]) + pa(R[5], f[2]);
  const float n2 = pa(R[6], f[0]) + pa(R[7], f[1]) + pa(R[8], f[2]);
  float ff[3];
  ff[0] = pa(R[0], f[3]) + pa(R[1], f[4]) + pa(R[2], f[5]);
  ff[1] = pa(R[3], f[3]) + pa(R[4], f[4]) + pa(R[5], f[5]);
  ff[2] = pa(R[6], f[3]) + pa(R[7], f[4]) + pa(R[8], f[5]);
  float t[3];
  cross3(r, ff, t);
  o[0] = n0 + t[0];
  o[1] = n1 + t[1];
  o[2] = n2 + t[2];
  o[3] = ff[0];
  o[4] = ff[1];
  o[5] = ff[2];
}

// motion cross product v x m
__device__ __forceinline__ void crm(const float *v, const float *m, float *o) {
  float a[3], b[3], c[3];
  cross3(v, m, a);
  cross3(v, m + 3, b);
  cross3(v + 3, m, c);
  o[0] = a[0];
  o[1] = a[1];
  o[2] = a[2];
  o[3] = b[0] + c[0];
  o[4] = b[1] + c[1];
  o[5] = b[2] + c[2];
}

// force cross product v x* f
__device__ __forceinline__ void crf(const float *v, const float *f, float *o) {
  float a[3], b[3], c[3];
  cross3(v, f, a);
  cross3(v + 3, f + 3, b);
  cross3(v, f + 3, c);
  o[0] = a[0] + b[0];
  o[1] = a[1] + b[1];
  o[2] = a[2] + b[2];
  o[3] = c[0];
  o[4] = c[1];
  o[5] = c[2];
}

// z-axis (S = [e_z; 0]) cross products: S x m and S x* f
__device__ __forceinline__ void crm_z(const float *m, float *o) {
  o[0] = -m[1];
  o[1] = m[0];
  o[2] = 0.f;
  o[3] = -m[4];
  o[4] = m[3];
  o[5] = 0.f;
}

// o = I_i m = [Io w + mc x v; m v - mc x w]
__device__ __forceinline__ void inertia_mv(int i, const float *m, float *o) {
  const float mass = kLink[i][0];
  const float mc[3] = {kLink[i][1], kLink[i][2], kLink[i][3]};
  const float Ixx = kLink[i][4], Iyy = kLink[i][5], Izz = kLink[i][6];
  const float Ixy = kLink[i][7], Ixz = kLink[i][8], Iyz = kLink[i][9];
  float b[3], c[3];
  cross3(mc, m + 3, b);
  cross3(mc, m, c);
  o[0] = pa(Ixx, m[0]) + pa(Ixy, m[1]) + pa(Ixz, m[2]) + b[0];
  o[1] = pa(Ixy, m[0]) + pa(Iyy, m[1]) + pa(Iyz, m[2]) + b[1];
  o[2] = pa(Ixz, m[0]) + pa(Iyz, m[1]) + pa(Izz, m[2]) + b[2];
  o[3] = pa(mass, m[3]) - c[0];
  o[4] = pa(mass, m[4]) - c[1];
  o[5] = pa(mass, m[5]) - c[2];
}

// Rows j..6 of column j of M(q), the lower triangle that a Cholesky factor
// reads, by the unit-acceleration recursion (RNEA at qd = 0, qdd = e_j,
// gravity 0): link accelerations forward from joint j (those before j are
// exactly zero), forces I a backward, row i = the joint-i component of the
// accumulated force; the backward pass stops at joint j (rows above the
// diagonal are not needed).
__device__ __forceinline__ void mass_column_lower(int j, const float *s, const float *c, float *col) {
  float a[NQ][6];
#pragma unroll
  for (int i = j; i < NQ; ++i) {
    if (i == j) {
#pragma unroll
      for (int k = 0; k < 6; ++k) a[i][k] = k == 2 ? 1.f : 0.f;
    } else {
      X_mv(i, s[i], c[i], a[i - 1], a[i]);
    }
  }
  float f[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = NQ - 1; i >= j; --i) {
    float fi[6];
    inertia_mv(i, a[i], fi);
#pragma unroll
    for (int k = 0; k < 6; ++k) f[k] += fi[k];
    col[i] = f[2];
    if (i > j) {
      float t[6];
      XT_fv(i, s[i], c[i], f, t);
#pragma unroll
      for (int k = 0; k < 6; ++k) f[k] = t[k];
    }
  }
}

// RNEA with gravity 0.  Returns tau; optionally stores per-link v, a, the
// total (backward-accumulated) f, X_i v_{i-1} and X_i a_{i-1} (6 each,
// link-major, stride 6) for the derivative passes.
template <bool STORE>
__device__ __forceinline__ void rnea(const float *s, const float *c, const float *qd,
                                     const float *qdd, float *tau, float *v_out, float *a_out,
                                     float *f_out, float *xv_out, float *xa_out) {
  float f[NQ][6];
  float vp[6] = {0, 0, 0, 0, 0, 0}, ap[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    float v[6], a[6];
    X_mv(i, s[i], c[i], vp, v);
    X_mv(i, s[i], c[i], ap, a);
    if (STORE) {
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        xv_out[6 * i + k] = v[k];
        xa_out[6 * i + k] = a[k];
      }
    }
    v[2] += qd[i];
    a[2] += qdd[i];
    // a += v x (S qd_i)
    const float sq[6] = {0, 0, qd[i], 0, 0, 0};
    float t[6];
    crm(v, sq, t);
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] += t[k];
    float Ia[6], Iv[6], cf[6];
    inertia_mv(i, a, Ia);
    inertia_mv(i, v, Iv);
    crf(v, Iv, cf);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      f[i][k] = Ia[k] + cf[k];
      vp[k] = v[k];
      ap[k] = a[k];
    }
    if (STORE) {
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        v_out[6 * i + k] = v[k];
        a_out[6 * i + k] = a[k];
      }
    }
  }
#pragma unroll
  for (int i = NQ - 1; i >= 0; --i) {
    tau[i] = f[i][2];
    if (STORE) {
#pragma unroll
      for (int k = 0; k < 6; ++k) f_out[6 * i + k] = f[i][k];
    }
    if (i > 0) {
      float t[6];
      XT_fv(i, s[i], c[i], f[i], t);
#pragma unroll
      for (int k = 0; k < 6; ++k) f[i - 1][k] += t[k];
    }
  }
}

// Directional derivative of RNEA at the stored base point in direction
// dir (< 7: q_dir, 7..13: qd_{dir-7}).  Base arrays are link-major (stride
// 6): v, f (total), X_i v_{i-1}, X_i a_{i-1}.  dtau receives 7 values.
__device__ __forceinline__ void rnea_dir(int dir, const float *s, const float *c, const float *qd,
                                         const float *bv, const float *bf, const float *bxv,
                                         const float *bxa, float *dtau) {
  const int jq = dir < NQ ? dir : -1;
  const int jd = dir >= NQ ? dir - NQ : -1;
  float df[NQ][6];
  float dvp[6] = {0, 0, 0, 0, 0, 0}, dap[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    float dv[6], da[6], t[6];
    X_mv(i, s[i], c[i], dvp, dv);
    X_mv(i, s[i], c[i], dap, da);
    if (i == jq) {
      crm_z(bxv + 6 * i, t);
#pragma unroll
      for (int k = 0; k < 6; ++k) dv[k] -= t[k];
      crm_z(bxa + 6 * i, t);
#pragma unroll
      for (int k = 0; k < 6; ++k) da[k] -= t[k];
    }
    if (i == jd) dv[2] += 1.f;
    const float sq[6] = {0, 0, qd[i], 0, 0, 0};
    crm(dv, sq, t);
#pragma unroll
    for (int k = 0; k < 6; ++k) da[k] += t[k];
    if (i == jd) {
      // v_i x S  = -(S x v_i)
      crm_z(bv + 6 * i, t);
#pragma unroll
      for (int k = 0; k < 6; ++k) da[k] -= t[k];
    }
    float v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = bv[6 * i + k];
    float Ida[6], Iv[6], Idv[6], c1[6], c2[6];
    inertia_mv(i, da, Ida);
    inertia_mv(i, v, Iv);
    inertia_mv(i, dv, Idv);
    crf(dv, Iv, c1);
    crf(v, Idv, c2);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      df[i][k] = Ida[k] + c1[k] + c2[k];
      dvp[k] = dv[k];
      dap[k] = da[k];
    }
  }
#pragma unroll
  for (int i = NQ - 1; i >= 0; --i) {
    dtau[i] = df[i][2];
    if (i > 0) {
      float t[6];
      XT_fv(i, s[i], c[i], df[i], t);
#pragma unroll
      for (int k = 0; k < 6; ++k) df[i - 1][k] += t[k];
      if (i == jq) {
        // (dX_i/dq_i)^T f_i = X_i^T (S x* f_i)
        const float *fi = bf + 6 * i;
        const float sf[6] = {-fi[1], fi[0], 0.f, -fi[4], fi[3], 0.f};
        XT_fv(i, s[i], c[i], sf, t);
#pragma unroll
        for (int k = 0; k < 6; ++k) df[i - 1][k] += t[k];
      }
    }
  }
}

// World frames of the chain: for joint j (1-based frame j+1) the z axis
// and origin in world coordinates (3 each), and the link-7 rotation.
__device__ __forceinline__ void fk_frames(const float *s, const float *c, float *zax, float *org,
                                          float *Rend) {
  float Rw[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  float pw[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    float R[9];
    joint_rot(i, s[i], c[i], R);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      pw[a] += pa(Rw[3 * a + 0], kPF[i][0]) + pa(Rw[3 * a + 1], kPF[i][1]) + pa(Rw[3 * a + 2], kPF[i][2]);
    float Rn[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        Rn[3 * r + cc] = pa(Rw[3 * r + 0], R[0 + cc]) + pa(Rw[3 * r + 1], R[3 + cc]) + pa(Rw[3 * r + 2], R[6 + cc]);
#pragma unroll
    for (int k = 0; k < 9; ++k) Rw[k] = Rn[k];
    zax[3 * i + 0] = Rw[2];
    zax[3 * i + 1] = Rw[5];
    zax[3 * i + 2] = Rw[8];
    org[3 * i + 0] = pw[0];
    org[3 * i + 1] = pw[1];
    org[3 * i + 2] = pw[2];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) Rend[k] = Rw[k];
}

// End-effector position of the link-7 origin (rbdfiles/iiwa_eepos_grid.cuh
// :2015-2091; no tool offset).
__device__ __forceinline__ void fk_pos(const float *s, const float *c, float *p) {
  float Rw[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  float pw[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    float R[9];
    joint_rot(i, s[i], c[i], R);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      pw[a] += pa(Rw[3 * a + 0], kPF[i][0]) + pa(Rw[3 * a + 1], kPF[i][1]) + pa(Rw[3 * a + 2], kPF[i][2]);
    if (i < NQ - 1) {
      float Rn[9];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
          Rn[3 * r + cc] =
              pa(Rw[3 * r + 0], R[0 + cc]) + pa(Rw[3 * r + 1], R[3 + cc]) + pa(Rw[3 * r + 2], R[6 + cc]);
#pragma unroll
      for (int k = 0; k < 9; ++k) Rw[k] = Rn[k];
    }
  }
  p[0] = pw[0];
  p[1] = pw[1];
  p[2] = pw[2];
}

// Lower Cholesky factor L of a 7x7 symmetric positive definite M (both
// column-major; only the lower triangle of M is read), one lane.
__device__ __forceinline__ void cholesky7(const float *M, float *L) {
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    float d = M[j + 7 * j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[j + 7 * k] * L[j + 7 * k];
    d = sqrtf(d);
    L[j + 7 * j] = d;
    const float inv = 1.f / d;
#pragma unroll
    for (int i = j + 1; i < NQ; ++i) {
      float s = M[i + 7 * j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[i + 7 * k] * L[j + 7 * k];
      L[i + 7 * j] = s * inv;
    }
#pragma unroll
    for (int i = 0; i < j; ++i) L[i + 7 * j] = 0.f;
  }
}

// x = L^-T L^-1 b for a 7x7 lower Cholesky factor
__device__ __forceinline__ void chol_solve7(const float *L, const float *b, float *x) {
  float y[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    float s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i + 7 * k] * y[k];
    y[i] = s / L[i + 7 * i];
  }
#pragma unroll
  for (int i = NQ - 1; i >= 0; --i) {
    float s = y[i];
#pragma unroll
    for (int k = i + 1; k < NQ; ++k) s -= L[k + 7 * i] * x[k];
    x[i] = s / L[i + 7 * i];
  }
}

}  // namespace mpc
