"""The entry point that allocates nothing (mpc_direct_solve_static: the
direct Schur solve with its factor workspace in a library-owned device
array) against the one it varies (mpc_direct_solve, caller workspace):
bitwise equal solutions on two calls in a row, the second with a smaller
system, for refined and unrefined solves, and the refined solution within
2e-5 of the oracle's fp64 sparse LDL^T (the bound of the direct-solve
parity test)."""
import numpy as np
import pytest

import helpers

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _schur(orc, seed, N):
    x0, ee, xu = helpers.perturbed_instance(orc, seed, N)
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xu[:14].astype(np.float64))
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    return np.asarray(S, np.float32), np.asarray(gam, np.float32)


def test_static_direct_solve_matches_workspace_entry(solver, orc, pkg):
    for N in (512, 200, 64):  # each call smaller than the one before
        S32, gam32 = _schur(orc, 900 + N, N)
        S, gam = torch.from_numpy(S32).cuda(), torch.from_numpy(gam32).cuda()
        for refine in (2, 0):
            lam_ref = torch.full((14 * N,), 7.0, device="cuda")  # overwritten by both solves
            lam_sta = torch.full((14 * N,), -3.0, device="cuda")
            solver.direct_solve(N, S, gam, lam_ref, refine=refine)
            pkg.direct_solve_static(N, S, gam, lam_sta, refine=refine)
            torch.cuda.synchronize()
            a, b = lam_ref.cpu().numpy(), lam_sta.cpu().numpy()
            print(f"N={N} refine={refine}: bitwise equal={np.array_equal(a, b)}")
            assert np.all(np.isfinite(b))
            assert np.array_equal(a, b)
            if refine:
                lam_d = orc.LDL(N).solve(S32.astype(np.float64), gam32.astype(np.float64))
                assert helpers.rel_err(b.astype(np.float64), lam_d) < 2e-5


def test_static_direct_solve_rejects_sizes_beyond_its_array(pkg):
    n = pkg.DIRECT_STATIC_MAX_KNOTS + 1  # buffers large enough: the library itself must refuse
    S = torch.zeros(588 * n, device="cuda")
    gam = torch.zeros(14 * n, device="cuda")
    lam = torch.full((14 * n,), 5.0, device="cuda")
    with pytest.raises(pkg.MpcError, match=r"bad argument \(1\)"):
        pkg.direct_solve_static(n, S, gam, lam)
    with pytest.raises(pkg.MpcError, match=r"bad argument \(1\)"):
        pkg.direct_solve_static(1, S, gam, lam)
    torch.cuda.synchronize()
    assert bool((lam == 5.0).all())  # nothing was launched
