"""CPU: the split-preconditioned form of block-Jacobi PCG that the batched
GPU path runs (k_pcg_split.hip), restated in numpy (helpers.split_cg, in
fp64 here) and checked against the oracle's PCG (the reference's
pcg<T,14,N>, toplevel.cuh:143-157).

With theta_b = -S(b,b) = U_b^T U_b (Cholesky) and W_b = U_b^-1, block-Jacobi
PCG on S lambda = gamma is CG on W^T S W lambda^ = W^T gamma with
lambda = W lambda^: the transformed diagonal blocks are -I, the
preconditioner is -I, and r^T Pinv r = r^ . (-r^).  In fp64 the two must
give the same iteration count (the same exit decisions) and the same
multipliers to rounding, from a zero and from a warm start; a dropped
transpose, a wrong neighbour block or a sign slip in the transform breaks
one of these."""
import numpy as np
import pytest

import helpers


def _blocks(N, S):
    B = S.reshape(N, 3, 14, 14)  # column-major blocks: element (r, c) at r + 14 c
    return np.transpose(B[:, 0], (0, 2, 1)), np.transpose(B[:, 1], (0, 2, 1))


def _system(orc, N, seed):
    cfg = orc.make_cfg()
    x0, ee, xu = helpers.perturbed_instance(orc, seed, N)
    G, Cm, g, c = orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), x0.astype(np.float64))
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    return S, P, gam


@pytest.mark.parametrize("N,seed", [(2, 1), (9, 2), (32, 3), (128, 4)])
def test_split_cg_is_block_jacobi_pcg(orc, N, seed):
    S, P, gam = _system(orc, N, seed)
    Lh, W, U, gh = helpers.split_system(N, S, gam)
    _, D = _blocks(N, S)
    # the transform: W^T theta W = I, and gamma^ . gamma^ = gamma^T (-Pinv) gamma
    assert np.allclose(np.einsum("bki,bkl,blj->bij", W, -D, W), np.eye(14)[None], atol=1e-9)
    Pd = np.transpose(P.reshape(N, 3, 14, 14)[:, 1], (0, 2, 1))
    gg = gam.reshape(N, 14)
    assert np.sum(gh * gh) == pytest.approx(-np.einsum("bi,bij,bj->", gg, Pd, gg), rel=1e-9)
    rng = np.random.default_rng(seed)
    for lam0 in (np.zeros(14 * N), 1e-2 * rng.standard_normal(14 * N)):
        for cap, tol in ((173, 1e-5), (40, 1e-12), (5000, 1e-9)):
            lo, it_o, _ = orc.pcg(N, S, P, gam, lam0, cap, tol)
            ls, it_s = helpers.split_cg(N, S, gam, lam0, cap, tol)
            assert it_s == it_o, (cap, tol)
            assert helpers.rel_err(ls, lo) < 1e-7, (cap, tol)  # fp64 rounding over up to 5000 iterations


def test_split_catches_a_transposed_neighbour_block(orc):
    """A plausible slip (W_{b-1} where W_b belongs) changes the solve."""
    N = 16
    S, P, gam = _system(orc, N, 7)
    lo, _, _ = orc.pcg(N, S, P, gam, np.zeros(14 * N), 5000, 1e-12)
    L, D = _blocks(N, S)
    U = np.transpose(np.linalg.cholesky(-D), (0, 2, 1))
    W = np.linalg.inv(U)
    Lbad = np.zeros_like(L)
    for b in range(1, N):
        Lbad[b] = W[b - 1].T @ L[b] @ W[b]
    Lgood, _, _, _ = helpers.split_system(N, S, gam)
    assert np.abs(Lbad - Lgood).max() > 1e-3 * np.abs(Lgood).max()
    ls, _ = helpers.split_cg(N, S, gam, np.zeros(14 * N), 5000, 1e-12)
    assert helpers.rel_err(ls, lo) < 1e-7
    # solving the system built from the slipped blocks gives another lambda
    Lh = Lbad
    Sh = -np.eye(14 * N)
    for b in range(1, N):
        Sh[14 * b:14 * b + 14, 14 * (b - 1):14 * b] = Lh[b]
        Sh[14 * (b - 1):14 * b, 14 * b:14 * b + 14] = Lh[b].T
    gh = np.einsum("bkc,bk->bc", W, gam.reshape(N, 14)).ravel()
    lbad = np.einsum("brc,bc->br", W, np.linalg.solve(Sh, gh).reshape(N, 14)).ravel()
    assert helpers.rel_err(lbad, lo) > 1e-3
