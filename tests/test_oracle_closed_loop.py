"""Pinning of the oracle's closed-loop harness (orc_closed_loop): the noise
definition, one control step against an independent restatement (the
oracle's own SQP for the plan, numpy Euler over orc_fd for the plant, the
shift rule), and convergence towards a stationary reachable target."""
import numpy as np
import pytest

import helpers

scipy_stats = pytest.importorskip("scipy.stats")


def _rest_instance(orc, N, steps):
    q0 = np.array([0.1, -0.2, 0.3, -0.4, 0.2, 0.1, -0.1])
    x0 = np.r_[q0, np.zeros(7)]
    qt = q0 + np.array([0.05, 0.04, -0.03, 0.05, 0.02, -0.04, 0.03])
    path = np.tile(orc.fk(qt), steps + N)
    xu = np.zeros(21 * N - 7)
    for k in range(N):
        xu[21 * k:21 * k + 14] = x0
    return x0, path, xu


def test_noise_is_standard_normal_and_deterministic(orc):
    draws = np.stack([orc.cl_noise(99, s) for s in range(3000)])
    assert np.array_equal(draws[17], orc.cl_noise(99, 17))
    assert not np.array_equal(orc.cl_noise(99, 17), orc.cl_noise(100, 17))
    flat = draws.ravel()
    assert abs(flat.mean()) < 4 / np.sqrt(flat.size)
    assert abs(flat.var() - 1) < 0.03
    assert scipy_stats.kstest(flat, "norm").pvalue > 1e-3
    # components and steps are uncorrelated
    c = np.corrcoef(draws.T)
    assert np.abs(c - np.eye(14)).max() < 0.08
    assert abs(np.corrcoef(draws[:-1, 3], draws[1:, 3])[0, 1]) < 0.08


def test_one_control_step_matches_restatement(orc):
    N, steps, sigma, seed = 8, 1, 1e-3, 7
    x0, path_q = helpers.synth_instance_spec(31, steps + N)
    path = np.concatenate([orc.fk(q) for q in path_q])
    traj = np.zeros(21 * (steps + N) - 7)
    for k in range(steps + N):
        traj[21 * k:21 * k + 7] = path_q[k]
        if k < steps + N - 1:
            traj[21 * k + 14:21 * k + 21] = 0.1 * (k + 1)
    xu = traj[:21 * N - 7].copy()
    cfg = orc.make_cfg()
    cfg.sqp_max_iter = 2
    lam0 = np.zeros(14 * N)
    r = orc.closed_loop(cfg, N, steps, x0, path, xu, lam0, sigma=sigma, seed=seed, traj=traj)
    # restatement: the plan from x0, Euler under u_0, noise, shift
    plan0 = xu.copy()
    plan0[:14] = x0
    s = orc.sqp_solve(cfg, N, path[:6 * N], lam0, plan0, 1e-3)
    plan, lam = s["xu"], s["lam"]
    x = x0.copy()
    for _ in range(50):
        qdd = orc.fd(x[:7], x[7:], plan[14:21])
        x = np.r_[x[:7] + 2e-4 * x[7:], x[7:] + 2e-4 * qdd]
    x = x + sigma * orc.cl_noise(seed, 0)
    np.testing.assert_allclose(r["states"][1], x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(r["states"][0], x0, rtol=0, atol=0)
    d = orc.fk(x[:7])[:3] - path[6:9]
    np.testing.assert_allclose(r["ee_err"][0], [np.linalg.norm(d), np.abs(d).sum()], rtol=1e-12)
    # shift: x_k <- x_{k+1}, u_k <- u_{k+1}; tail from the trajectory; x_0 <- plant
    want = np.r_[plan[21:21 * N - 28 + 21], traj[21 * (N - 1) + 14:21 * (N - 1) + 35]]
    want[:14] = x
    np.testing.assert_allclose(r["xu"], want, rtol=0, atol=1e-15)
    np.testing.assert_allclose(r["lam"], np.r_[lam[14:], lam[-14:]], rtol=0, atol=1e-15)
    assert r["rho"] == pytest.approx(s["rho"])


def test_shift_without_trajectory_zeroes_the_last_control(orc):
    N, steps = 6, 1
    x0, path, xu = _rest_instance(orc, N, steps)
    xu[14:] += 0.01
    cfg = orc.make_cfg()
    cfg.sqp_max_iter = 1
    r = orc.closed_loop(cfg, N, steps, x0, path, xu, np.zeros(14 * N), sigma=0.0)
    out = r["xu"]
    assert np.all(out[21 * (N - 2) + 14:21 * (N - 2) + 21] == 0)
    np.testing.assert_allclose(out[:14], r["states"][1], rtol=0, atol=0)


def test_stationary_target_is_approached(orc):
    """Reachable fixed target 1.6 cm away, no noise, converged linear
    solves: the end-effector error falls steadily (the line search accepts
    short steps, so the approach is slow but monotone)."""
    N, steps = 16, 200
    x0, path, xu = _rest_instance(orc, N, steps)
    cfg = orc.make_cfg()
    cfg.sqp_max_iter = 3
    cfg.pcg_tol = 1e-20
    cfg.pcg_max_iter = 3000
    r = orc.closed_loop(cfg, N, steps, x0, path, xu, np.zeros(14 * N), sigma=0.0)
    e = r["ee_err"][:, 0]
    e0 = np.linalg.norm(orc.fk(x0[:7])[:3] - path[:3])
    assert e[-1] < 0.5 * e0
    assert np.all(np.diff(e[::20]) < 0)
