"""Pinning of the oracle's differentiable QP layer and its adjoint (fp64).

The layer maps the cost weights (wx: 14, wu: 7) and the reference points to
L = sum_k |ee(q*_k) - ref_k|^2 with x* = xu - dz the full-step solution of
the SQP step's QP at a fixed iterate xu (direct LDL^T solve).  The adjoint
gradient must equal fp64 central differences with h = 1e-4, Richardson-
extrapolated with h/2 (the plain difference is O(h^2)-biased for the small
control weights, whose inverse enters the solution).
"""
import numpy as np
import pytest

import helpers

RHO = 1e-3


def _instance(orc, seed, N):
    x0, ee, xu = helpers.perturbed_instance(orc, seed, N)
    return xu.astype(np.float64), ee.astype(np.float64), xu[:14].astype(np.float64)


def _weights(seed):
    rng = np.random.default_rng(seed)
    wx, wu = (np.asarray(w) for w in __import__("oracle").default_weights())
    wx = wx + rng.uniform(0.0, 0.05, 14)  # also exercises the (otherwise zero) q weights
    wu = wu + rng.uniform(0.0, 2e-3, 7)
    return wx, wu


def _loss(orc, cfg, N, xu, ee, xs):
    xstar, _, _ = orc.qp_layer(cfg, N, xu, ee, xs, RHO)
    return orc.qp_loss(N, xstar, ee)[0]


def _richardson(f, x0, h=1e-4):
    def cd(hh):
        g = np.zeros_like(x0)
        for j in range(len(x0)):
            e = np.zeros_like(x0)
            e[j] = hh
            g[j] = (f(x0 + e) - f(x0 - e)) / (2 * hh)
        return g
    return (4 * cd(h / 2) - cd(h)) / 3


@pytest.mark.parametrize("N,seed,own", [(4, 1, 0), (9, 2, 0), (9, 3, 1)])
def test_oracle_adjoint_matches_finite_differences(orc, N, seed, own):
    xu, ee, xs = _instance(orc, 40 + seed, N)
    wx, wu = _weights(seed)
    cfg = orc.make_cfg(wx=wx, wu=wu, last_knot_own_state=own)
    xstar, dz, _ = orc.qp_layer(cfg, N, xu, ee, xs, RHO)
    L, a, gref_direct = orc.qp_loss(N, xstar, ee)
    gw, gref = orc.qp_adjoint(cfg, N, xu, ee, xs, RHO, dz, a)
    gref = gref + gref_direct

    theta0 = np.concatenate([wx, wu])
    fd_w = _richardson(lambda th: _loss(orc, orc.make_cfg(wx=th[:14], wu=th[14:], last_knot_own_state=own),
                                        N, xu, ee, xs), theta0)
    assert helpers.rel_err(gw, fd_w) < 1e-6

    idx = np.array([6 * k + b for k in range(N) for b in range(3)])

    def f_ref(r):
        e2 = ee.copy()
        e2[idx] = r
        return _loss(orc, cfg, N, xu, e2, xs)
    fd_r = _richardson(f_ref, ee[idx].copy())
    assert helpers.rel_err(gref[idx], fd_r) < 1e-6
    assert not gref[np.setdiff1d(np.arange(6 * N), idx)].any()


def test_layer_solves_the_kkt_system(orc):
    """x* = xu - dz with [G~ C^T; C 0][dz; lambda] = [g; c] (c_0 treated as 0),
    checked against a dense fp64 solve of the assembled KKT system."""
    N = 6
    xu, ee, xs = _instance(orc, 77, N)
    cfg = orc.make_cfg()
    xstar, dz, lam = orc.qp_layer(cfg, N, xu, ee, xs, RHO)
    G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xs)
    Gd, Cd = orc.dense_G(N, G) + RHO * np.eye(21 * N - 7), orc.dense_C(N, Cm)
    n, m = Gd.shape[0], Cd.shape[0]
    K = np.block([[Gd, Cd.T], [Cd, np.zeros((m, m))]])
    cz = np.asarray(c, np.float64).copy()
    cz[:14] = 0.0
    sol = np.linalg.solve(K, np.concatenate([g, cz]))
    assert helpers.rel_err(dz, sol[:n]) < 1e-9
    assert helpers.rel_err(lam, sol[n:]) < 1e-9
    assert np.allclose(xstar, xu - dz)
