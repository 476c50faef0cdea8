"""Pinning of the oracle's stair preconditioner (orc_pcg_ex, stair = 1):
iterate by iterate against a numpy PCG with the explicit dense
preconditioner D^-1 - D^-1 (L + L^T) D^-1 (the form the reference stores,
oldschur.cuh:334-494), convergence to the dense solve, and the halving of
the iteration count against block Jacobi that this preconditioner gives on
a block-tridiagonal system."""
import numpy as np
import pytest

import helpers


def _system(orc, seed, N, perturbed):
    x0, ee, xu = (helpers.perturbed_instance if perturbed else helpers.make_instance)(orc, seed, N)
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xu[:14].astype(np.float64))
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    return S, P, gam


def _dense_pcg(A, b, M, iters):
    x = np.zeros_like(b)
    r = b - A @ x
    z = M @ r
    p, eta = z.copy(), r @ z
    for _ in range(iters):
        y = A @ p
        a = eta / (p @ y)
        x += a * p
        r -= a * y
        z = M @ r
        en = r @ z
        p = z + (en / eta) * p
        eta = en
    return x


@pytest.mark.parametrize("N,perturbed", [(3, False), (8, True), (24, True)])
def test_stair_iterates_match_explicit_preconditioner(orc, N, perturbed):
    S, P, gam = _system(orc, 80 + N, N, perturbed)
    A = orc.dense_S(N, S)
    D = np.zeros_like(A)
    for k in range(N):
        D[14 * k:14 * k + 14, 14 * k:14 * k + 14] = A[14 * k:14 * k + 14, 14 * k:14 * k + 14]
    Di = np.linalg.inv(D)
    M = Di - Di @ (A - D) @ Di
    np.testing.assert_allclose(M, M.T, atol=1e-9 * np.abs(M).max())
    for k in (1, 2, 5, 12):
        lam, it, _ = orc.pcg(N, S, P, gam, np.zeros(14 * N), k, 0.0, stair=True)
        ref = _dense_pcg(A, gam, M, k)
        assert it == k
        assert helpers.rel_err(lam, ref) < 1e-9, k


@pytest.mark.parametrize("N", [16, 64])
def test_stair_converges_and_halves_iterations(orc, N):
    S, P, gam = _system(orc, 90 + N, N, True)
    lam_d = np.linalg.solve(orc.dense_S(N, S), gam)
    lam_s, it_s, mx_s = orc.pcg(N, S, P, gam, np.zeros(14 * N), 20000, 1e-22, stair=True)
    lam_j, it_j, mx_j = orc.pcg(N, S, P, gam, np.zeros(14 * N), 20000, 1e-22)
    print(f"N={N}: block Jacobi {it_j} iterations, stair {it_s}")
    assert mx_s == 0 and mx_j == 0
    assert helpers.rel_err(lam_s, lam_d) < 1e-7
    assert abs(2 * it_s - it_j) <= 0.1 * it_j
