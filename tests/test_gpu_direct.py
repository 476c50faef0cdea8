"""GPU direct solve of the Schur system (block LDL^T of -S, C-ABI
mpc_direct_solve and MPC_FLAG_DIRECT_SOLVE) against the oracle's sparse
LDL^T (fp64; its fp32 twin measures the fp32 floor), and the plain and
batched SQP in direct mode (Schur solve and KKT step both refined with fp64
residuals) against the oracle's fp64 direct SQP and each other.

Tolerances: a direct fp32 elimination of S (condition up to ~1e7 here) is
off by about cond x eps: the unrefined solve (refine = 0) must stay within
16x the fp32 LDL^T twin's error (both are fp32 eliminations, measured
0.1x..9x apart), and the refined solve (2 steps with an fp64 residual)
within 2e-5 of the fp64 LDL^T solution, with an fp64-evaluated residual
within 2x that of the fp64 solution rounded to fp32 (storing the exact
solution in fp32 alone leaves residuals up to ~1e-4 on these systems)."""
import numpy as np
import pytest

import helpers

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def _schur(orc, seed, N, perturbed=False):
    if perturbed:
        x0, ee, xu = helpers.perturbed_instance(orc, seed, N)
    else:
        x0, ee, xu = helpers.make_instance(orc, seed, N)
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xu[:14].astype(np.float64))
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    return np.asarray(S, np.float32), np.asarray(gam, np.float32)


@pytest.mark.parametrize("N", [2, 3, 16, 64, 128, 200, 256, 512])
def test_direct_solve_matches_ldl(solver, orc, N):
    for perturbed in (False, True):
        S32, gam32 = _schur(orc, 60 + N, N, perturbed)
        lam0, lam = torch.zeros(14 * N, device="cuda"), torch.zeros(14 * N, device="cuda")
        solver.direct_solve(N, dev(S32), dev(gam32), lam0, refine=0)
        solver.direct_solve(N, dev(S32), dev(gam32), lam)
        lam_u, lam_g = host(lam0), host(lam)
        lam_d = orc.LDL(N).solve(S32.astype(np.float64), gam32.astype(np.float64))
        lam_t = orc.LDL(N, np.float32).solve(S32, gam32).astype(np.float64)
        Sd, g64 = orc.dense_S(N, S32.astype(np.float64)), gam32.astype(np.float64)
        res = lambda l: np.linalg.norm(g64 - Sd @ l) / np.linalg.norm(g64)
        err, err_u, err_t = (helpers.rel_err(v, lam_d) for v in (lam_g, lam_u, lam_t))
        print(f"N={N} perturbed={perturbed}: rel(gpu, ldl64) unrefined={err_u:.2e} refined={err:.2e} "
              f"residual={res(lam_g):.2e} (fp32-rounded exact: {res(lam_d.astype(np.float32).astype(np.float64)):.2e}) | "
              f"fp32 LDL twin rel={err_t:.2e} residual={res(lam_t):.2e}")
        assert np.all(np.isfinite(lam_g))
        assert err_u <= max(1e-4, 16 * err_t)
        floor = res(lam_d.astype(np.float32).astype(np.float64))
        assert err < 2e-5 and res(lam_g) <= 2 * floor + 1e-7


@pytest.fixture
def direct_mode(solver, pkg):
    saved = solver.cfg.flags, solver.cfg.sqp_max_iter
    solver.set_config(flags=pkg.FLAG_DIRECT_SOLVE)
    yield
    solver.set_config(flags=saved[0], sqp_max_iter=saved[1])


@pytest.mark.parametrize("N", [8, 32, 128])
def test_plain_direct_sqp_vs_oracle(solver, orc, direct_mode, N):
    x0, ee, xu = helpers.make_instance(orc, 140 + N, N)
    cfg = orc.make_cfg(sqp_max_iter=3)
    r = orc.sqp_solve(cfg, N, ee.astype(np.float64), np.zeros(14 * N), xu.astype(np.float64), 1e-3, use_direct=True)
    solver.set_config(sqp_max_iter=3)
    dxu, dlam = dev(xu), torch.zeros(14 * N, device="cuda")
    pcg_iters, _, _, sqp_iters, not_aborted, _, rho, steps = solver.sqp_solve(N, dev(ee), dlam, dxu, 1e-3)
    print(f"N={N} gpu steps={steps} oracle steps={r['ls_steps']} rel(xu)={helpers.rel_err(host(dxu), r['xu']):.2e}")
    assert steps == r["ls_steps"] and sqp_iters == r["sqp_iters"]
    assert all(p == 0 for p in pcg_iters)
    # the refined direct steps are exact to fp32 storage (measured ~1e-6)
    assert helpers.rel_err(host(dxu), r["xu"]) < 1e-4


@pytest.mark.parametrize("horizons", [[32, 64, 17, 128, 2, 45], [256, 129, 200, 64, 512, 3]])
def test_batched_direct_matches_plain_direct(solver, orc, pkg, direct_mode, horizons):
    xus, ees = [], []
    for i, N in enumerate(horizons):
        _, ee, xu = helpers.make_instance(orc, 2000 + i, N)
        xus.append(xu)
        ees.append(ee)
    n, T = len(horizons), sum(horizons)

    def batch():
        xu, lam = dev(np.concatenate(xus)), torch.zeros(14 * T, device="cuda")
        rho, st = torch.full((n,), 1e-3, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda")
        solver.batch_setup(horizons)
        solver.batch_sqp(xu, lam, dev(np.concatenate(ees)), rho, 3, st)
        return host(xu), host(lam), host(st)

    xb, lb, st = batch()
    xb2, lb2, _ = batch()
    assert np.array_equal(xb, xb2) and np.array_equal(lb, lb2)
    assert (st.astype(np.int64) & ~np.int64(1)).max() == 0
    off = loff = 0
    solver.set_config(sqp_max_iter=3)
    for i, N in enumerate(horizons):
        dxu, dlam = dev(xus[i]), torch.zeros(14 * N, device="cuda")
        solver.sqp_solve(N, dev(ees[i]), dlam, dxu, 1e-3)
        L = 21 * N - 7
        assert helpers.rel_err(xb[off:off + L], host(dxu)) < 1e-5, i
        assert helpers.rel_err(lb[loff:loff + 14 * N], host(dlam)) < 1e-5, i
        off += L
        loff += 14 * N
