"""Pins the oracle's KKT blocks, Schur complement, PCG, dz, merit, sparse
LDL^T and SQP driver (CPU only) against mathematics the reference fixes:

- KKT blocks: central differences of the Euler map x + dt [qd; FD(x, u)] and
  of the tracking cost; the terminal-knot cost reads x_{N-2} as in
  rbdfiles/iiwa_eepos_plant.cuh:399 unless the flag asks for x_{N-1}.
- Schur / gamma / dz: dense numpy assembly of the full KKT system
  [G~ C^T; C 0] with C = [I; -A -B I] rows (schur.cuh:391-524 layouts); the
  step solves it (with c_0 dropped, oldschur.cuh:639-654).
- PCG converges to numpy.linalg.solve; exit-rule semantics.
- LDL^T: L D L^T reconstructs S in the reference CSC pattern.
- merit: brute force from the pinned FK / FD; SQP: an independent Python
  statement of the line search and rho rule (toplevel.cuh:384-444).
- a planted sign error in the oracle's gamma is caught by the KKT check.
"""
import numpy as np
import pytest

import helpers


def euler_map(orc, cfg, x, u):
    qdd = orc.fd(x[:7], x[7:], u)
    return x + cfg.dt * np.r_[x[7:], qdd]


@pytest.fixture(scope="module")
def inst(orc):
    N = 6
    x0, ee, xu = helpers.perturbed_instance(orc, 77, N, scale=0.03)
    return N, xu.astype(np.float64), ee.astype(np.float64), xu[:14].astype(np.float64)


def test_kkt_dynamics_blocks_are_euler_jacobians(orc, inst):
    N, xu, ee, xs = inst
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    h = 1e-6
    for k in range(N - 1):
        x, u, xn = xu[21 * k:21 * k + 14], xu[21 * k + 14:21 * k + 21], xu[21 * (k + 1):21 * (k + 1) + 14]
        z = np.r_[x, u]
        f = lambda zz: euler_map(orc, cfg, zz[:14], zz[14:])
        J = np.stack([(f(z + h * e) - f(z - h * e)) / (2 * h) for e in np.eye(21)], 1)
        mA = Cm[294 * k:294 * k + 196].reshape(14, 14).T
        mB = Cm[294 * k + 196:294 * k + 294].reshape(7, 14).T
        assert np.abs(-mA - J[:, :14]).max() < 1e-6
        assert np.abs(-mB - J[:, 14:]).max() < 1e-6
        assert np.abs(c[14 * (k + 1):14 * (k + 2)] - (xn - f(z))).max() < 1e-12
    assert np.abs(c[:14] - (xu[:14] - xs)).max() == 0


def tracking_cost(orc, cfg, x, ref):
    e = orc.fk(x[:7])[:3] - ref[:3]
    return 0.5 * cfg.q_ee * e @ e + 0.5 * cfg.qd_cost * x[7:] @ x[7:]


@pytest.mark.parametrize("own", [0, 1])
def test_kkt_cost_blocks(orc, inst, own):
    N, xu, ee, xs = inst
    cfg = orc.make_cfg(last_knot_own_state=own)
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    h = 1e-6
    for k in range(N):
        # knot N-1's cost is evaluated at x_{N-2} (reference) or x_{N-1}
        kx = k if (k < N - 1 or own) else N - 2
        x, ref = xu[21 * kx:21 * kx + 14], ee[6 * k:6 * k + 6]
        grad = np.array([(tracking_cost(orc, cfg, x + h * e, ref) - tracking_cost(orc, cfg, x - h * e, ref)) / (2 * h)
                         for e in np.eye(14)])
        qk = g[21 * k:21 * k + 14]
        assert np.abs(qk - grad).max() < 1e-7
        Q = G[245 * k:245 * k + 196].reshape(14, 14).T
        gq = qk[:7]
        assert np.allclose(Q[:7, :7], np.outer(gq, gq)) and np.allclose(Q[7:, 7:], cfg.qd_cost * np.eye(7))
        assert not Q[:7, 7:].any()
        if k < N - 1:
            u = xu[21 * k + 14:21 * k + 21]
            assert np.allclose(g[21 * k + 14:21 * k + 21], cfg.r_cost * u)
            assert np.allclose(G[245 * k + 196:245 * k + 245].reshape(7, 7), cfg.r_cost * np.eye(7))


def _kkt_dense(orc, N, G, Cm, g, c, rho):
    Gd = orc.dense_G(N, G) + rho * np.eye(21 * N - 7)
    Cd = orc.dense_C(N, Cm)
    cp = c.copy()
    cp[:14] = 0.0
    return Gd, Cd, cp


@pytest.mark.parametrize("rho", [1e-3, 0.3])
def test_schur_gamma_and_dz_solve_the_kkt_system(orc, inst, rho):
    N, xu, ee, xs = inst
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    S, P, gam, Gi = orc.form_schur(N, G, Cm, g, c, rho)
    Gd, Cd, cp = _kkt_dense(orc, N, G, Cm, g, c, rho)
    Sd = orc.dense_S(N, S)
    assert np.abs(Sd + Cd @ np.linalg.solve(Gd, Cd.T)).max() < 1e-9 * np.abs(Sd).max()
    assert np.abs(gam - (cp - Cd @ np.linalg.solve(Gd, g))).max() < 1e-9 * np.abs(gam).max()
    # block-Jacobi: inverse of every diagonal block, off-diagonal blocks zero
    Pd = orc.dense_S(N, P)
    for b in range(N):
        blk = Sd[14 * b:14 * b + 14, 14 * b:14 * b + 14]
        assert np.allclose(Pd[14 * b:14 * b + 14, 14 * b:14 * b + 14] @ blk, np.eye(14), atol=1e-8)
    mask = np.kron(np.eye(N), np.ones((14, 14))) == 0
    assert not Pd[mask].any()
    # the full KKT system [G~ C^T; C 0][dz; lambda] = [g; c'] is solved
    lam = np.linalg.solve(Sd, gam)
    dz = orc.compute_dz(N, Gi, Cm, g, lam)
    K = np.block([[Gd, Cd.T], [Cd, np.zeros((14 * N, 14 * N))]])
    sol = np.linalg.solve(K, np.r_[g, cp])
    assert np.abs(dz - sol[:21 * N - 7]).max() < 1e-8 * np.abs(sol).max()
    assert np.abs(lam - sol[21 * N - 7:]).max() < 1e-8 * np.abs(sol).max()


def test_planted_gamma_sign_error_is_caught(orc, inst):
    N, xu, ee, xs = inst
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    S, P, gam, Gi = orc.form_schur(N, G, Cm, g, c, 1e-3)
    Gd, Cd, cp = _kkt_dense(orc, N, G, Cm, g, c, 1e-3)
    bad = gam.copy()
    bad[14:28] = -bad[14:28]  # planted: sign of one block row of gamma
    lam = np.linalg.solve(orc.dense_S(N, S), bad)
    dz = orc.compute_dz(N, Gi, Cm, g, lam)
    assert np.abs(Cd @ dz - cp).max() > 1e-3


def test_pcg_converges_and_exit_semantics(orc, inst):
    N, xu, ee, xs = inst
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    lam_d = np.linalg.solve(orc.dense_S(N, S), gam)
    lam, it, mx = orc.pcg(N, S, P, gam, np.zeros(14 * N), 5000, 1e-18)
    assert mx == 0 and helpers.rel_err(lam, lam_d) < 1e-9
    # exact arithmetic terminates in <= 14 N steps; fp64 stays close to that
    assert it <= 2 * 14 * N
    # tol 0: exactly max_iter iterations, flag set
    _, it0, mx0 = orc.pcg(N, S, P, gam, np.zeros(14 * N), 7, 0.0)
    assert (it0, mx0) == (7, 1)
    # warm start inside the tolerance: no iteration
    _, it1, mx1 = orc.pcg(N, S, P, gam, lam_d, 100, 1e-5)
    assert (it1, mx1) == (0, 0)
    # exit test is on |r^T Pinv r|: the returned residual satisfies it
    lam5, it5, _ = orc.pcg(N, S, P, gam, np.zeros(14 * N), 5000, 1e-5)
    r = gam - orc.dense_S(N, S) @ lam5
    assert abs(r @ (orc.dense_S(N, P) @ r)) < 1e-5 * (1 + 1e-6)


@pytest.mark.parametrize("wx_q", [0.0, 0.3])
def test_cost_hessian_closed_form_inverse(orc, inst, wx_q):
    """The batched split path inverts the regularised cost blocks in closed
    form (k_schur.hip k_form_split, k_compute_dz_lr): the oracle's Q block
    is g_q g_q^T + diag(wx) (iiwa_eepos_plant.cuh:357-375), so
    (Q + rho I)^-1 = D^-1 - u u^T / s (D = diag(wx + rho), u = D^-1 v,
    s = 1 + v^T u, v = (g_q, 0)), and R + rho I is diagonal.  Checked
    against the oracle's own block inverses (form_schur's G~^-1), with the
    default weights and with a nonzero position weight."""
    N, xu, ee, xs = inst
    wx = np.array([wx_q] * 7 + [1e-2] * 7)
    cfg = orc.make_cfg(wx=wx)
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    rho = 1e-3
    _, _, _, Ginv = orc.form_schur(N, G, Cm, g, c, rho)
    wu = np.array([1e-4] * 7)
    for k in range(N):
        Q = G[245 * k:245 * k + 196].reshape(14, 14).T
        # Q - diag(wx) = v v^T: one nonzero eigenvalue, in the position block
        ev, V = np.linalg.eigh(Q - np.diag(wx))
        v = np.sqrt(max(ev[-1], 0.0)) * V[:, -1]
        assert np.abs(ev[:-1]).max() <= 1e-12 * max(ev[-1], 1e-30) + 1e-15
        assert np.abs(v[7:]).max() <= 1e-12 * np.abs(v).max() + 1e-15
        assert np.allclose(Q, np.outer(v, v) + np.diag(wx), rtol=1e-12, atol=1e-15)
        d = 1.0 / (wx + rho)
        u = d * v
        Qi = np.diag(d) - np.outer(u, u) / (1.0 + v @ u)
        Qi_orc = Ginv[245 * k:245 * k + 196].reshape(14, 14).T
        assert np.abs(Qi - Qi_orc).max() <= 1e-9 * np.abs(Qi_orc).max()
        if k + 1 < N:
            Ri_orc = Ginv[245 * k + 196:245 * k + 245].reshape(7, 7).T
            assert np.allclose(Ri_orc, np.diag(1.0 / (wu + rho)), rtol=1e-12, atol=0)


def test_pcg_trace_and_exit_band(orc, inst):
    """The oracle PCG's |r^T Pinv r| trace: entry k is that of the iterate
    after k iterations (recomputed from lambda), the tolerance exit is the
    first entry below tol, and the exit band of helpers.pcg_exit_ok accepts
    that count and rejects exits far from the crossing."""
    N, xu, ee, xs = inst
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    Sd, Pd = orc.dense_S(N, S), orc.dense_S(N, P)
    cap = 60
    _, itc, mxc, e = orc.pcg(N, S, P, gam, np.zeros(14 * N), cap, 0.0, trace=True)
    assert (itc, mxc, len(e)) == (cap, 1, cap + 1)
    for k in (0, 1, 5, 17):
        lam_k, _, _ = orc.pcg(N, S, P, gam, np.zeros(14 * N), k, 0.0)
        r = gam - Sd @ lam_k
        assert e[k] == pytest.approx(abs(r @ (Pd @ r)), rel=1e-6, abs=1e-300)
    tol = float(e[12]) * 1.0001  # a tolerance the trace crosses
    _, it, mx = orc.pcg(N, S, P, gam, np.zeros(14 * N), cap, tol)
    assert mx == 0 and it == int(np.flatnonzero(e < tol)[0])
    assert helpers.pcg_exit_ok(it, e, tol, cap)
    assert not helpers.pcg_exit_ok(0, e, tol, cap)
    clear = int(np.flatnonzero(e < 0.75 * tol)[0])
    assert not helpers.pcg_exit_ok(clear + 1, e, tol, cap)
    assert not helpers.pcg_exit_ok(cap, e, tol, cap)


@pytest.mark.parametrize("N", [2, 5, 12])
def test_ldl_reconstructs_schur_matrix(orc, N):
    x0, ee, xu = helpers.perturbed_instance(orc, 90 + N, N)
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xu[:14].astype(np.float64))
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    Ap, Ai = orc.csc_pattern(N)
    assert Ap[-1] == orc.lib().orc_csc_nnz(N) == 196 * (N - 1) + 105 * N
    for j in range(14 * N):  # upper triangle, rows from the previous block
        rows = Ai[Ap[j]:Ap[j + 1]]
        assert rows[-1] == j and rows[0] == max(0, (j // 14 - 1) * 14) and np.all(np.diff(rows) == 1)
    Lp, Li, Lx, D, rc = orc.ldl_factor_raw(N, S)
    assert rc == 0  # no positive pivots: S is negative definite
    n = 14 * N
    L = np.eye(n)
    for j in range(n):
        L[Li[Lp[j]:Lp[j + 1]], j] = Lx[Lp[j]:Lp[j + 1]]
    Sd = orc.dense_S(N, S)
    assert np.abs(L @ np.diag(D) @ L.T - Sd).max() < 1e-9 * np.abs(Sd).max()
    lam = orc.LDL(N).solve(S, gam)
    assert helpers.rel_err(lam, np.linalg.solve(Sd, gam)) < 1e-10
    lam32 = orc.LDL(N, np.float32).solve(S, gam)
    assert helpers.rel_err(lam32, np.linalg.solve(Sd, gam)) < 1e-3


def test_merit_brute_force(orc, inst):
    N, xu, ee, xs = inst
    cfg = orc.make_cfg()
    rng = np.random.default_rng(0)
    dz = rng.normal(0, 0.02, xu.size)
    for alpha, with_x0 in ((0.0, False), (-0.25, True), (-1.0, True)):
        z = xu + alpha * dz
        tot = 0.0
        for k in range(N):
            x = z[21 * k:21 * k + 14]
            e = orc.fk(x[:7])[:3] - ee[6 * k:6 * k + 3]
            J = 0.5 * cfg.qd_cost * x[7:] @ x[7:] + 0.5 * cfg.q_ee * e @ e
            if k < N - 1:
                u = z[21 * k + 14:21 * k + 21]
                J += 0.5 * cfg.r_cost * u @ u
                ck = np.abs(z[21 * (k + 1):21 * (k + 1) + 14] - euler_map(orc, cfg, x, u)).sum()
            else:
                ck = np.abs(z[:14] - xs).sum() if with_x0 else 0.0
            tot += J + cfg.mu * ck
        m, _ = orc.merit(cfg, N, xu, ee, xs, dz, alpha, with_x0)
        assert abs(m - tot) < 1e-10 * max(1, abs(tot))


def test_sqp_driver_matches_python_statement(orc):
    """Line search: smallest index with merit strictly below the current one;
    failure raises rho by drho = max(drho f, f); success lowers it by
    drho = min(drho / f, 1 / f); rho >= rho_min; rho > rho_max aborts."""
    N = 8
    x0, ee, xu = helpers.make_instance(orc, 5, N)
    xu, ee = xu.astype(np.float64), ee.astype(np.float64)
    cfg = orc.make_cfg(sqp_max_iter=4)
    ref = orc.sqp_solve(cfg, N, ee, np.zeros(14 * N), xu, 1e-3)
    xs = xu[:14].copy()
    z, lam, rho, drho = xu.copy(), np.zeros(14 * N), 1e-3, 1.0
    m0, _ = orc.merit(cfg, N, z, ee, xs)
    steps = []
    for _ in range(cfg.sqp_max_iter):
        G, Cm, g, c = orc.form_kkt(cfg, N, z, ee, xs)
        S, P, gam, Gi = orc.form_schur(N, G, Cm, g, c, rho)
        lam, it, mx = orc.pcg(N, S, P, gam, lam, cfg.pcg_max_iter, cfg.pcg_tol)
        dz = orc.compute_dz(N, Gi, Cm, g, lam)
        ms = orc.ls_merits(cfg, N, z, ee, xs, dz)
        below = [i for i in range(8) if ms[i] < m0 and ms[i] == ms[ms < m0].min()] if (ms < m0).any() else []
        if not below:
            steps.append(-1)
            drho = max(drho * cfg.rho_factor, cfg.rho_factor)
            rho = max(rho * drho, cfg.rho_min)
            if rho > cfg.rho_max:
                break
            continue
        i = below[0]
        steps.append(i)
        drho = min(drho / cfg.rho_factor, 1 / cfg.rho_factor)
        rho = max(rho * drho, cfg.rho_min)
        z = z - dz / (1 << i)
        m0 = ms[i]
    assert steps == ref["ls_steps"]
    assert helpers.rel_err(ref["xu"], z) < 1e-12 and abs(ref["rho"] - rho) < 1e-15
