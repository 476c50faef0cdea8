"""GPU parity of the full solves: plain SQP (drop-in for sqpSolve) against
the oracle's SQP driver, and batched SQP against the plain solves."""
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


@pytest.mark.parametrize("N", [8, 32, 200, 300, 512])
def test_plain_sqp_vs_oracle(solver, orc, N):
    x0, ee, xu = helpers.make_instance(orc, 100 + N, N)
    cfg = orc.make_cfg(sqp_max_iter=3)
    r = orc.sqp_solve(cfg, N, ee.astype(np.float64), np.zeros(14 * N), xu.astype(np.float64), 1e-3)
    dxu, dlam = dev(xu), torch.zeros(14 * N, device="cuda")
    out = solver.sqp_solve(N, dev(ee), dlam, dxu, 1e-3)
    pcg_iters, linsys, sqp_us, sqp_iters, not_aborted, maxit, rho, steps = out
    print(f"N={N} gpu steps={steps} pcg={pcg_iters} oracle steps={r['ls_steps']} pcg={r['pcg_iters']}")
    assert sqp_iters == r["sqp_iters"] and not_aborted == r["not_aborted"]
    assert steps == r["ls_steps"]
    assert helpers.rel_err(host(dxu), r["xu"]) < 1e-3
    assert abs(rho - r["rho"]) <= 1e-6 * r["rho"]


def _batch(orc, pkg, horizons, seed0=1000):
    xus, ees = [], []
    for i, N in enumerate(horizons):
        x0, qref = pkg.synth_instance(seed0 + i, int(N))
        ee = np.concatenate([orc.fk(q.astype(np.float64)) for q in qref]).astype(np.float32)
        xu = np.zeros(21 * N - 7, np.float32)
        for k in range(N):
            xu[21 * k:21 * k + 14] = x0
        xus.append(xu)
        ees.append(ee)
    return xus, ees


@pytest.mark.parametrize("horizons", [[32, 64, 17, 128, 2, 45], [256, 129, 200, 64, 256, 3],
                                      [512, 300, 257, 64, 200, 400]])
def test_batch_matches_plain_and_is_deterministic(solver, orc, pkg, horizons):
    xus, ees = _batch(orc, pkg, horizons)
    n = len(horizons)
    T = sum(horizons)
    xu = dev(np.concatenate(xus))
    ee = dev(np.concatenate(ees))
    lam = torch.zeros(14 * T, device="cuda")
    rho = torch.full((n,), 1e-3, device="cuda")
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    pit = torch.zeros(n * 3, dtype=torch.int32, device="cuda")
    solver.batch_setup(horizons)
    # the plain driver's PCG is the reference form: compare with the same
    # form here (the split form, default where every N <= 256, is checked
    # against the oracle in test_gpu_split.py and test_gpu_batch_oracle.py)
    solver.batch_set_plan(split=0)
    try:
        _batch_vs_plain(solver, xus, ees, horizons, xu, lam, ee, rho, st, pit)
    finally:
        solver.batch_set_plan(split=1)


def _batch_vs_plain(solver, xus, ees, horizons, xu, lam, ee, rho, st, pit):
    n, T = len(horizons), sum(horizons)
    solver.batch_sqp(xu, lam, ee, rho, 3, st, pit)
    xb, lb = host(xu), host(lam)
    # determinism: a second run from the same inputs is bitwise identical
    xu2, lam2 = dev(np.concatenate(xus)), torch.zeros(14 * T, device="cuda")
    rho2 = torch.full((n,), 1e-3, device="cuda")
    solver.batch_sqp(xu2, lam2, ee, rho2, 3, st, pit)
    assert np.array_equal(host(xu2), xb) and np.array_equal(host(lam2), lb)
    # per instance: same as the plain one-instance solve
    off = 0
    loff = 0
    for i, N in enumerate(horizons):
        dxu, dlam = dev(xus[i]), torch.zeros(14 * N, device="cuda")
        solver.set_config(sqp_max_iter=3)
        solver.sqp_solve(N, dev(ees[i]), dlam, dxu, 1e-3)
        L = 21 * N - 7
        assert helpers.rel_err(xb[off:off + L], host(dxu)) < 1e-5, i
        assert helpers.rel_err(lb[loff:loff + 14 * N], host(dlam)) < 1e-5, i
        off += L
        loff += 14 * N
    assert (host(st).astype(np.int64) & ~np.int64(3)).max() == 0
