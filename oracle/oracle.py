"""ctypes front end of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports this module.

Every function takes numpy arrays and a ``dtype`` (np.float64 -> the fp64
restatement, np.float32 -> the fp32 twin).  Layouts are the reference's
(column-major 14x14 / 14x7 / 7x7 blocks, see oracle/mpc_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NQ, NX, NU, NXU = 7, 14, 7, 21


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


class Cfg(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("q_ee", C.c_double),
        ("qd_cost", C.c_double),
        ("r_cost", C.c_double),
        ("mu", C.c_double),
        ("rho_factor", C.c_double),
        ("rho_max", C.c_double),
        ("rho_min", C.c_double),
        ("pcg_tol", C.c_double),
        ("pcg_max_iter", C.c_int),
        ("sqp_max_iter", C.c_int),
        ("last_knot_own_state", C.c_int),
        ("wx", C.c_double * 14),
        ("wu", C.c_double * 7),
        ("pcg_stair", C.c_int),
    ]


def default_weights(qd_cost=1e-2, r_cost=1e-4):
    """Per-dimension weights of the reference's cost: wx = (0 x 7, QD x 7), wu = R x 7."""
    return np.array([0.0] * 7 + [qd_cost] * 7), np.array([r_cost] * 7)


def make_cfg(dt=0.01, q_ee=1.0, qd_cost=1e-2, r_cost=1e-4, mu=10.0, rho_factor=1.2, rho_max=10.0,
             rho_min=1e-3, pcg_tol=1e-5, pcg_max_iter=173, sqp_max_iter=3, last_knot_own_state=0,
             wx=None, wu=None, pcg_stair=0):
    dwx, dwu = default_weights(qd_cost, r_cost)
    wx = dwx if wx is None else np.asarray(wx, np.float64)
    wu = dwu if wu is None else np.asarray(wu, np.float64)
    return Cfg(dt, q_ee, qd_cost, r_cost, mu, rho_factor, rho_max, rho_min, pcg_tol, pcg_max_iter,
               sqp_max_iter, last_knot_own_state, (C.c_double * 14)(*wx), (C.c_double * 7)(*wu), int(pcg_stair))


def _sfx(dtype):
    return "_d" if np.dtype(dtype) == np.float64 else "_f"


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _arr(x, dtype):
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


def _fn(name, dtype):
    return getattr(lib(), name + _sfx(dtype))


def fk(q, dtype=np.float64):
    q = _arr(q, dtype)
    out = np.zeros(6, dtype)
    _fn("orc_fk", dtype)(_ptr(q), _ptr(out))
    return out


def fk_grad(q, dtype=np.float64):
    """3 x 7 Jacobian of the end-effector position."""
    q = _arr(q, dtype)
    out = np.zeros(21, dtype)
    _fn("orc_fk_grad", dtype)(_ptr(q), _ptr(out))
    return out.reshape(7, 3).T


def rnea(q, qd, qdd, dtype=np.float64):
    q, qd, qdd = (_arr(a, dtype) for a in (q, qd, qdd))
    out = np.zeros(7, dtype)
    _fn("orc_rnea", dtype)(_ptr(q), _ptr(qd), _ptr(qdd), _ptr(out))
    return out


def rnea_grad(q, qd, qdd, dtype=np.float64):
    q, qd, qdd = (_arr(a, dtype) for a in (q, qd, qdd))
    out = np.zeros(98, dtype)
    _fn("orc_rnea_grad", dtype)(_ptr(q), _ptr(qd), _ptr(qdd), _ptr(out))
    return out.reshape(14, 7).T


def mass_matrix(q, dtype=np.float64):
    q = _arr(q, dtype)
    out = np.zeros(49, dtype)
    _fn("orc_mass_matrix", dtype)(_ptr(q), _ptr(out))
    return out.reshape(7, 7).T


def minv(q, dtype=np.float64):
    q = _arr(q, dtype)
    out = np.zeros(49, dtype)
    _fn("orc_minv", dtype)(_ptr(q), _ptr(out))
    return out.reshape(7, 7).T


def fd(q, qd, u, dtype=np.float64):
    q, qd, u = (_arr(a, dtype) for a in (q, qd, u))
    out = np.zeros(7, dtype)
    _fn("orc_fd", dtype)(_ptr(q), _ptr(qd), _ptr(u), _ptr(out))
    return out


def fd_grad(q, qd, u, dtype=np.float64):
    """Returns (qdd (7), df_du (7 x 21) = [dqdd/dq, dqdd/dqd, dqdd/du])."""
    q, qd, u = (_arr(a, dtype) for a in (q, qd, u))
    qdd = np.zeros(7, dtype)
    out = np.zeros(147, dtype)
    _fn("orc_fd_grad", dtype)(_ptr(q), _ptr(qd), _ptr(u), _ptr(qdd), _ptr(out))
    return qdd, out.reshape(21, 7).T


def form_kkt(cfg, N, xu, ee_ref, xs, dtype=np.float64):
    xu, ee_ref, xs = (_arr(a, dtype) for a in (xu, ee_ref, xs))
    G = np.zeros(245 * N, dtype)
    Cm = np.zeros(294 * N, dtype)
    g = np.zeros(21 * N, dtype)
    c = np.zeros(14 * N, dtype)
    _fn("orc_form_kkt", dtype)(C.byref(cfg), N, _ptr(xu), _ptr(ee_ref), _ptr(xs), _ptr(G), _ptr(Cm),
                               _ptr(g), _ptr(c))
    return G[: 245 * N - 49], Cm[: 294 * (N - 1)], g[: 21 * N - 7], c


def _pad(a, n, dtype):
    out = np.zeros(n, dtype)
    a = np.asarray(a, dtype).ravel()
    out[: a.size] = a
    return out


def form_schur(N, G, Cm, g, c, rho, dtype=np.float64):
    G, Cm, g = _pad(G, 245 * N, dtype), _pad(Cm, 294 * N, dtype), _pad(g, 21 * N, dtype)
    c = _arr(c, dtype)
    S = np.zeros(588 * N, dtype)
    P = np.zeros(588 * N, dtype)
    gam = np.zeros(14 * N, dtype)
    Ginv = np.zeros(245 * N, dtype)
    fn = _fn("orc_form_schur", dtype)
    fn.restype = C.c_int
    rc = fn(N, _ptr(G), _ptr(Cm), _ptr(g), _ptr(c), (C.c_double if dtype == np.float64 else C.c_float)(rho),
            _ptr(S), _ptr(P), _ptr(gam), _ptr(Ginv))
    if rc != 0:
        raise ValueError("form_schur: a regularised block is not positive definite")
    return S, P, gam, Ginv[: 245 * N - 49]


def bd_mv(N, S, x, dtype=np.float64):
    S, x = _arr(S, dtype), _arr(x, dtype)
    y = np.zeros(14 * N, dtype)
    _fn("orc_bd_mv", dtype)(N, _ptr(S), _ptr(x), _ptr(y))
    return y


def pcg(N, S, P, gamma, lam0, max_iter, tol, dtype=np.float64, stair=False, trace=False):
    """The reference's PCG (see orc_pcg_ex): (lambda, iterations, capped);
    with trace=True also |r^T z| before the first iteration and after each
    one (iterations + 1 values)."""
    S, P, gamma = (_arr(a, dtype) for a in (S, P, gamma))
    lam = _arr(lam0, dtype).copy()
    it, mx = C.c_int(), C.c_int()
    tolc = (C.c_double if dtype == np.float64 else C.c_float)(tol)
    etas = np.zeros(max_iter + 1, dtype)
    _fn("orc_pcg_trace", dtype)(N, _ptr(S), _ptr(P), _ptr(gamma), _ptr(lam), max_iter, tolc, int(stair),
                                C.byref(it), C.byref(mx), _ptr(etas) if trace else None)
    if trace:
        return lam, it.value, mx.value, etas[:it.value + 1]
    return lam, it.value, mx.value


def compute_dz(N, Ginv, Cm, g, lam, dtype=np.float64):
    Ginv, Cm, g = _pad(Ginv, 245 * N, dtype), _pad(Cm, 294 * N, dtype), _pad(g, 21 * N, dtype)
    lam = _arr(lam, dtype)
    dz = np.zeros(21 * N, dtype)
    _fn("orc_compute_dz", dtype)(N, _ptr(Ginv), _ptr(Cm), _ptr(g), _ptr(lam), _ptr(dz))
    return dz[: 21 * N - 7]


def merit(cfg, N, xu, ee_ref, xs, dz=None, alpha=0.0, with_x0=False, dtype=np.float64):
    xu, ee_ref, xs = (_arr(a, dtype) for a in (xu, ee_ref, xs))
    fn = _fn("orc_merit", dtype)
    rt = C.c_double if dtype == np.float64 else C.c_float
    fn.restype = rt
    knots = np.zeros(N, dtype)
    dzp = _ptr(_arr(dz, dtype)) if dz is not None else None
    dzk = _arr(dz, dtype) if dz is not None else None
    if dzk is not None:
        dzp = _ptr(dzk)
    val = fn(C.byref(cfg), N, _ptr(xu), dzp, rt(alpha), _ptr(ee_ref), _ptr(xs), int(with_x0), _ptr(knots))
    return float(val), knots


def ls_merits(cfg, N, xu, ee_ref, xs, dz, dtype=np.float64):
    return np.array([merit(cfg, N, xu, ee_ref, xs, dz, -1.0 / (1 << p), True, dtype)[0] for p in range(8)])


def sqp_solve(cfg, N, ee_ref, lam, xu, rho, rho_reset=1e-3, use_direct=False, dtype=np.float64):
    ee_ref = _arr(ee_ref, dtype)
    lam = _arr(lam, dtype).copy()
    xu = _arr(xu, dtype).copy()
    rt = C.c_double if dtype == np.float64 else C.c_float
    rho_c = rt(rho)
    n = max(cfg.sqp_max_iter, 1)
    pit = np.zeros(n, np.int32)
    pmx = np.zeros(n, np.int32)
    lss = np.full(n, -2, np.int32)
    si, na = C.c_int(), C.c_int()
    fn = _fn("orc_sqp_solve", dtype)
    fn.restype = C.c_int
    rc = fn(C.byref(cfg), N, _ptr(ee_ref), _ptr(lam), _ptr(xu), C.byref(rho_c), rt(rho_reset), int(use_direct),
            C.byref(si), _ptr(pit), _ptr(pmx), _ptr(lss), C.byref(na))
    if rc != 0:
        raise RuntimeError(f"orc_sqp_solve failed ({rc})")
    k = si.value
    return dict(xu=xu, lam=lam, rho=float(rho_c.value), sqp_iters=k, pcg_iters=pit[:k].tolist(),
                pcg_maxiter=pmx[:k].tolist(), ls_steps=lss[:k].tolist(), not_aborted=bool(na.value))


# -- sparse LDL^T --------------------------------------------------------------------------------

def csc_pattern(N):
    n = 14 * N
    nnz = lib().orc_csc_nnz(N)
    Ap = np.zeros(n + 1, np.int32)
    Ai = np.zeros(nnz, np.int32)
    lib().orc_csc_pattern(N, _ptr(Ap), _ptr(Ai))
    return Ap, Ai


def csc_values(N, S, dtype=np.float64):
    S = _arr(S, dtype)
    Ax = np.zeros(lib().orc_csc_nnz(N), dtype)
    _fn("orc_csc_values", dtype)(N, _ptr(S), _ptr(Ax))
    return Ax


class LDL:
    """Sparse LDL^T of the Schur matrix in the reference's CSC pattern."""

    def __init__(self, N, dtype=np.float64):
        self.N, self.dtype = N, np.dtype(dtype)
        f = _fn("orc_ldl_create", dtype)
        f.restype = C.c_void_p
        self.h = C.c_void_p(f(N))

    def solve(self, S, gamma):
        S, gamma = _arr(S, self.dtype), _arr(gamma, self.dtype)
        lam = np.zeros(14 * self.N, self.dtype)
        f = _fn("orc_ldl_schur_solve", self.dtype)
        f.restype = C.c_int
        rc = f(self.h, self.N, _ptr(S), _ptr(gamma), _ptr(lam))
        if rc != 0:
            raise ValueError("LDL factorisation hit a zero pivot")
        return lam

    def close(self):
        if self.h:
            _fn("orc_ldl_destroy", self.dtype)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ldl_factor_raw(N, S, dtype=np.float64):
    """Returns (Lp, Li, Lx, D) of the factorisation, for reconstruction checks."""
    n = 14 * N
    Ap, Ai = csc_pattern(N)
    Ax = csc_values(N, S, dtype)
    Lnz = np.zeros(n, np.int32)
    et = np.zeros(n, np.int32)
    iw = np.zeros(3 * n, np.int32)
    lnz = lib().orc_ldl_etree(n, _ptr(Ap), _ptr(Ai), _ptr(iw), _ptr(Lnz), _ptr(et))
    Lp = np.zeros(n + 1, np.int32)
    Li = np.zeros(max(lnz, 1), np.int32)
    Lx = np.zeros(max(lnz, 1), dtype)
    D = np.zeros(n, dtype)
    Dinv = np.zeros(n, dtype)
    fw = np.zeros(n, dtype)
    f = _fn("orc_ldl_factor", dtype)
    f.restype = C.c_int
    rc = f(n, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(Lp), _ptr(Li), _ptr(Lx), _ptr(D), _ptr(Dinv), _ptr(Lnz), _ptr(et),
           _ptr(iw), _ptr(fw))
    return Lp, Li[:lnz], Lx[:lnz], D, rc


# -- dense views of the reference layouts (for numpy cross-checks) --------------------------------

def dense_S(N, S):
    """Dense 14N x 14N matrix from the 3-block-per-row layout."""
    S = np.asarray(S, np.float64)
    M = np.zeros((14 * N, 14 * N))
    for b in range(N):
        for col, bc in ((0, b - 1), (1, b), (2, b + 1)):
            if 0 <= bc < N:
                blk = S[588 * b + 196 * col: 588 * b + 196 * (col + 1)].reshape(14, 14).T
                M[14 * b:14 * b + 14, 14 * bc:14 * bc + 14] = blk
    return M


def dense_G(N, G):
    G = np.asarray(G, np.float64)
    n = 21 * N - 7
    M = np.zeros((n, n))
    for k in range(N):
        M[21 * k:21 * k + 14, 21 * k:21 * k + 14] = G[245 * k:245 * k + 196].reshape(14, 14).T
        if k < N - 1:
            M[21 * k + 14:21 * k + 21, 21 * k + 14:21 * k + 21] = G[245 * k + 196:245 * k + 245].reshape(7, 7).T
    return M


def dense_C(N, Cm):
    """Constraint Jacobian with row block 0 = [I 0 ...] and row block k+1 =
    [.. -A_k -B_k I ..] (C stores -A_k, -B_k)."""
    Cm = np.asarray(Cm, np.float64)
    M = np.zeros((14 * N, 21 * N - 7))
    M[0:14, 0:14] = np.eye(14)
    for k in range(N - 1):
        mA = Cm[294 * k:294 * k + 196].reshape(14, 14).T
        mB = Cm[294 * k + 196:294 * k + 294].reshape(7, 14).T
        M[14 * (k + 1):14 * (k + 2), 21 * k:21 * k + 14] = mA
        M[14 * (k + 1):14 * (k + 2), 21 * k + 14:21 * k + 21] = mB
        M[14 * (k + 1):14 * (k + 2), 21 * (k + 1):21 * (k + 1) + 14] = np.eye(14)
    return M


def ldl_bench(N, S, gamma, reps, dtype=np.float32):
    """Seconds for reps x (fill + factor + solve) of S lambda = gamma in C."""
    S, gamma = _arr(S, dtype), _arr(gamma, dtype)
    lam = np.zeros(14 * N, dtype)
    f = _fn("orc_ldl_bench", dtype)
    f.restype = C.c_double
    t = f(N, _ptr(S), _ptr(gamma), int(reps), _ptr(lam))
    return float(t), lam


# ---- differentiable QP layer (fp64 restatement of the GPU adjoint) --------

def _real(x, dtype):
    return (C.c_double if np.dtype(dtype) == np.float64 else C.c_float)(x)


def qp_layer(cfg, N, xu, ee_ref, xs, rho, dtype=np.float64):
    """x* = xu - dz of the QP formed at xu (direct LDL^T solve); returns (x*, dz, lambda)."""
    xu, ee_ref, xs = (_arr(a, dtype) for a in (xu, ee_ref, xs))
    L = 21 * N - 7
    xs_, dz, lam = np.zeros(L, dtype), np.zeros(L, dtype), np.zeros(14 * N, dtype)
    rc = _fn("orc_qp_layer", dtype)(C.byref(cfg), N, _ptr(xu), _ptr(ee_ref), _ptr(xs), _real(rho, dtype),
                                   _ptr(xs_), _ptr(dz), _ptr(lam))
    if rc:
        raise RuntimeError(f"orc_qp_layer failed ({rc})")
    return xs_, dz, lam


def qp_loss(N, xu, ee_ref, dtype=np.float64):
    """L = sum_k |ee(q_k) - ref_k|^2 with dL/dxu (21N-7) and dL/dref (6N)."""
    xu, ee_ref = _arr(xu, dtype), _arr(ee_ref, dtype)
    dx, dr = np.zeros(21 * N - 7, dtype), np.zeros(6 * N, dtype)
    fn = _fn("orc_qp_loss", dtype)
    fn.restype = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    L = fn(N, _ptr(xu), _ptr(ee_ref), _ptr(dx), _ptr(dr))
    return float(L), dx, dr


def qp_adjoint(cfg, N, xu, ee_ref, xs, rho, dz, a, dtype=np.float64):
    """Adjoint of qp_layer: weight gradient (14 wx + 7 wu) and the implicit part of dL/dref (6N)."""
    xu, ee_ref, xs, dz, a = (_arr(v, dtype) for v in (xu, ee_ref, xs, dz, a))
    gw, gref = np.zeros(21, dtype), np.zeros(6 * N, dtype)
    rc = _fn("orc_qp_adjoint", dtype)(C.byref(cfg), N, _ptr(xu), _ptr(ee_ref), _ptr(xs), _real(rho, dtype),
                                     _ptr(dz), _ptr(a), _ptr(gw), _ptr(gref))
    if rc:
        raise RuntimeError(f"orc_qp_adjoint failed ({rc})")
    return gw, gref


# -- closed loop -----------------------------------------------------------------------------------

def cl_noise(seed, step, dtype=np.float64):
    """The 14 standard normals of control step `step` (splitmix64 + Box-Muller)."""
    out = np.zeros(14, dtype)
    _fn("orc_cl_noise", dtype)(C.c_uint64(seed), C.c_uint32(step), _ptr(out))
    return out


def closed_loop(cfg, N, steps, x0, path, xu, lam, rho=1e-3, substeps=50, sim_dt=2e-4, sigma=1e-3, seed=12345,
                rho_reset=1e-3, traj=None, dtype=np.float64):
    """Closed-loop tracking (cfg.sqp_max_iter SQP iterations per control step).
    Returns dict(states (steps+1, 14), ee_err (steps, 2: Euclidean, L1), pcg_iters (steps, K), xu, lam, rho)."""
    x0, path = _arr(x0, dtype), _arr(path, dtype)
    traj = None if traj is None else _arr(traj, dtype)
    xu, lam = _arr(xu, dtype).copy(), _arr(lam, dtype).copy()
    K = max(cfg.sqp_max_iter, 1)
    states = np.zeros((steps + 1) * 14, dtype)
    err = np.zeros(2 * steps, dtype)
    pit = np.zeros(steps * K, np.int32)
    rt = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    rho_c = rt(rho)
    fn = _fn("orc_closed_loop", dtype)
    fn.restype = C.c_int
    rc = fn(C.byref(cfg), N, steps, substeps, rt(sim_dt), rt(sigma), C.c_uint64(seed), rt(rho_reset), _ptr(x0),
            _ptr(path), _ptr(traj) if traj is not None else None, _ptr(xu), _ptr(lam), C.byref(rho_c),
            _ptr(states), _ptr(err), _ptr(pit))
    if rc:
        raise RuntimeError(f"orc_closed_loop failed ({rc})")
    return dict(states=states.reshape(steps + 1, 14), ee_err=err.reshape(steps, 2), pcg_iters=pit.reshape(steps, K), xu=xu, lam=lam,
                rho=float(rho_c.value))

