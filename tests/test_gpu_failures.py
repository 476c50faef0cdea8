"""Failure paths on the GPU, each forced from the host so that every kernel
still exits normally:

- the rho abort of the line search (toplevel.cuh:396-407) with rho_max just
  above rho_min, plain and batched, against the oracle's SQP driver;
- a non-finite reference point: MPC_ERR_NONFINITE from the plain driver,
  MPC_ST_NONFINITE for that instance only in a batch;
- a PCG team whose exchange words never arrive (the fault-injection hook
  makes one part post under a wrong epoch): every team solve ends with flag
  2, the plain driver returns MPC_ERR_PCG_FAULT, the batched solve stops the
  instance with MPC_ST_PCG_FAULT without taking its step, and the resident
  bins of the same batch are untouched; after the hook is cleared, the same
  solves are bitwise the normal ones;
- the SQP wall-clock budget (sqp_max_time_us, toplevel.cuh:223-231): a tiny
  budget ends the plain solve with time_exit and the batched one with
  MPC_ST_TIME_BUDGET before any step, a huge one changes nothing;
- the reference's per-horizon PCG caps (perform_experiment.sh:14-19);
- the per-instance PCG exit state and the direct re-solve of the last
  batched Schur systems."""
import numpy as np
import pytest

import helpers

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture
def restore(solver):
    saved = {k: getattr(solver.cfg, k) for k, _ in type(solver.cfg)._fields_}
    yield solver
    solver.set_fault_injection(0)
    solver.batch_set_graph(False)
    solver.set_config(**saved)


def _batch(orc, pkg, horizons, seed0=3000):
    xus, ees = [], []
    for i, N in enumerate(horizons):
        x0, qref = pkg.synth_instance(seed0 + i, int(N))
        ees.append(np.concatenate([orc.fk(q.astype(np.float64)) for q in qref]).astype(np.float32))
        xu = np.zeros(21 * N - 7, np.float32)
        for k in range(N):
            xu[21 * k:21 * k + 14] = x0
        xus.append(xu)
    return xus, ees


def _run_batch(solver, horizons, xus, ees, iters=3, graph=False):
    n, T = len(horizons), sum(horizons)
    xu, ee = dev(np.concatenate(xus)), dev(np.concatenate(ees))
    lam = torch.zeros(14 * T, device="cuda")
    rho = torch.full((n,), 1e-3, device="cuda")
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    pit = torch.zeros(n * iters, dtype=torch.int32, device="cuda")
    solver.batch_setup(horizons)
    solver.batch_set_graph(graph)
    solver.batch_sqp(xu, lam, ee, rho, iters, st, pit)
    return host(xu), host(lam), host(rho), host(st).astype(np.int64), host(pit).reshape(n, iters)


def _slices(horizons):
    off = loff = 0
    for N in horizons:
        yield slice(off, off + 21 * N - 7), slice(loff, loff + 14 * N)
        off += 21 * N - 7
        loff += 14 * N


def test_rho_abort_plain_and_batched_vs_oracle(restore, orc, pkg):
    solver = restore
    N = 16
    x0, ee, xu = helpers.make_instance(orc, 7, N)
    # the second SQP iteration's line search fails (rho 1e-3 -> 1.2e-3) and
    # aborts above rho_max = 1.1e-3
    cfg = orc.make_cfg(sqp_max_iter=3, rho_max=1.1e-3)
    r = orc.sqp_solve(cfg, N, ee.astype(np.float64), np.zeros(14 * N), xu.astype(np.float64), 1e-3)
    assert not r["not_aborted"]
    solver.set_config(sqp_max_iter=3, rho_max=1.1e-3, rho_reset=1e-3)
    dxu, dlam = dev(xu), torch.zeros(14 * N, device="cuda")
    pcg, _, _, sqp_iters, not_aborted, _, rho, steps = solver.sqp_solve(N, dev(ee), dlam, dxu, 1e-3)
    print(f"gpu steps={steps} sqp_iters={sqp_iters} rho={rho}; oracle steps={r['ls_steps']} "
          f"sqp_iters={r['sqp_iters']} rho={r['rho']}")
    assert not not_aborted and sqp_iters == r["sqp_iters"] and steps == r["ls_steps"]
    assert rho == pytest.approx(1e-3) and r["rho"] == pytest.approx(1e-3)
    assert helpers.rel_err(host(dxu), r["xu"]) < 1e-3
    # batched: the same instance beside two others; its status carries the
    # abort, its rho is reset, its iterate is the plain solve's
    horizons = [N, 32, 64]
    xus, ees = _batch(orc, pkg, horizons)
    xus[0], ees[0] = xu, ee
    xb, _, rb, st, _ = _run_batch(solver, horizons, xus, ees)
    sl = next(_slices(horizons))[0]
    assert st[0] & pkg.ST_RHO_ABORT and not st[0] & pkg.ST_ACTIVE
    assert rb[0] == pytest.approx(1e-3)
    assert helpers.rel_err(xb[sl], host(dxu)) < 1e-5


def test_nonfinite_reference(restore, orc, pkg):
    solver = restore
    solver.set_config(sqp_max_iter=3)
    N = 32
    x0, ee, xu = helpers.make_instance(orc, 11, N)
    bad = ee.copy()
    bad[6 * 5 + 1] = np.nan
    dxu = dev(xu)
    with pytest.raises(pkg.MpcError, match=r"\(4\)"):
        solver.sqp_solve(N, dev(bad), torch.zeros(14 * N, device="cuda"), dxu, 1e-3)
    assert np.array_equal(host(dxu), xu)  # no step taken
    horizons = [64, N, 128, 32]
    xus, ees = _batch(orc, pkg, horizons)
    xus[1], ees[1] = xu, ee
    good = _run_batch(solver, horizons, xus, ees)
    ees[1] = bad
    nan = _run_batch(solver, horizons, xus, ees)
    st = nan[3]
    print(f"status with the NaN reference: {st.tolist()}")
    assert st[1] & pkg.ST_NONFINITE
    assert all(not (s & pkg.ST_NONFINITE) for i, s in enumerate(st) if i != 1)
    for i, (sx, sl) in enumerate(_slices(horizons)):
        if i != 1:  # the other instances are bitwise unaffected
            assert np.array_equal(nan[0][sx], good[0][sx]) and np.array_equal(nan[1][sl], good[1][sl])
    assert np.array_equal(nan[0][next(_slices(horizons))[0].stop:][:21 * N - 7], xu)


@pytest.mark.parametrize("N", [200, 400])
def test_team_fault_plain(restore, orc, pkg, N):
    solver = restore
    solver.set_config(sqp_max_iter=2)
    x0, ee, xu = helpers.make_instance(orc, 40 + N, N)
    ref_xu, ref_lam = dev(xu), torch.zeros(14 * N, device="cuda")
    solver.sqp_solve(N, dev(ee), ref_lam, ref_xu, 1e-3)
    for part in range(1, 5 if N > 256 else 3):
        solver.set_fault_injection(part)
        dxu = dev(xu)
        with pytest.raises(pkg.MpcError, match=r"\(6\)"):
            solver.sqp_solve(N, dev(ee), torch.zeros(14 * N, device="cuda"), dxu, 1e-3)
        assert np.array_equal(host(dxu), xu)  # the abandoned solve's step was not taken
    solver.set_fault_injection(0)
    dxu, dlam = dev(xu), torch.zeros(14 * N, device="cuda")
    solver.sqp_solve(N, dev(ee), dlam, dxu, 1e-3)
    assert np.array_equal(host(dxu), host(ref_xu)) and np.array_equal(host(dlam), host(ref_lam))


@pytest.mark.parametrize("graph", [False, True])
def test_team_fault_batched(restore, orc, pkg, graph):
    solver = restore
    horizons = [200, 64, 400, 32, 130, 128]
    xus, ees = _batch(orc, pkg, horizons, seed0=3100)
    good = _run_batch(solver, horizons, xus, ees, graph=graph)
    solver.set_fault_injection(2)
    bad = _run_batch(solver, horizons, xus, ees, graph=graph)
    st = bad[3]
    print(f"graph={graph} status: {st.tolist()}")
    for i, (N, (sx, sl)) in enumerate(zip(horizons, _slices(horizons))):
        if N > 128:  # on a team: stopped at its first PCG phase, no step taken
            assert st[i] & pkg.ST_PCG_FAULT and not st[i] & pkg.ST_ACTIVE
            assert np.array_equal(bad[0][sx], xus[i])
        else:  # resident bins: bitwise the fault-free solve
            assert not st[i] & pkg.ST_PCG_FAULT
            assert np.array_equal(bad[0][sx], good[0][sx]) and np.array_equal(bad[1][sl], good[1][sl])
    solver.set_fault_injection(0)
    again = _run_batch(solver, horizons, xus, ees, graph=graph)
    assert np.array_equal(again[0], good[0]) and np.array_equal(again[1], good[1])


def test_time_budget(restore, orc, pkg):
    solver = restore
    N = 32
    x0, ee, xu = helpers.make_instance(orc, 21, N)
    solver.set_config(sqp_max_iter=3, sqp_max_time_us=1.0)
    stats = pkg.SqpStats()
    dxu = dev(xu)
    out = solver.sqp_solve(N, dev(ee), torch.zeros(14 * N, device="cuda"), dxu, 1e-3, stats=stats)
    assert stats.time_exit == 1 and out[3] == 0
    assert np.array_equal(host(dxu), xu)
    horizons = [N, 64, 200, 128]
    xus, ees = _batch(orc, pkg, horizons, seed0=3200)
    solver.set_config(sqp_max_time_us=0.0)
    none = _run_batch(solver, horizons, xus, ees)
    solver.set_config(sqp_max_time_us=1e9)
    huge = _run_batch(solver, horizons, xus, ees)
    assert np.array_equal(huge[0], none[0]) and np.array_equal(huge[1], none[1])
    assert np.array_equal(huge[3], none[3])
    solver.set_config(sqp_max_time_us=1.0)
    tiny = _run_batch(solver, horizons, xus, ees)
    assert all(s & pkg.ST_TIME_BUDGET and not s & pkg.ST_ACTIVE for s in tiny[3])
    assert np.array_equal(tiny[0], np.concatenate(xus))


def test_pcg_cap_by_horizon_and_exit_state(restore, orc, pkg):
    solver = restore
    horizons = [32, 64, 128, 200, 256, 300, 40]
    xus, ees = _batch(orc, pkg, horizons, seed0=3300)
    n = len(horizons)
    solver.set_config(flags=pkg.FLAG_PCG_CAP_BY_N, pcg_max_iter=173, sqp_max_iter=2)
    _, lam, _, st, pit = _run_batch(solver, horizons, xus, ees, iters=2)
    caps = [173 if N <= 32 else 167 if N <= 128 else 118 if N <= 256 else 67 for N in horizons]
    print(f"pcg iterations {pit.tolist()} caps {caps}")
    assert all(pit[i].max() <= caps[i] for i in range(n))
    assert any(pit[i].max() == caps[i] for i in range(n))  # the caps bind on this workload
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    fl = torch.zeros(n, dtype=torch.int32, device="cuda")
    eta = torch.zeros(n, device="cuda")
    solver.batch_pcg_exit(it, fl, eta)
    it, fl, eta = host(it), host(fl), host(eta)
    tol = solver.cfg.pcg_exit_tol
    assert np.array_equal(it, pit[:, -1])
    for i in range(n):
        assert (fl[i] == 0 and eta[i] < tol) or (fl[i] == 1 and it[i] == caps[i] and eta[i] >= tol), i


def test_direct_resolve_is_the_direct_mode_solve(restore, orc, pkg):
    solver = restore
    horizons = [32, 200, 64, 400, 128]
    xus, ees = _batch(orc, pkg, horizons, seed0=3400)
    T = sum(horizons)
    # one SQP iteration: both modes solve the same first Schur systems
    solver.set_config(sqp_max_iter=1)
    _run_batch(solver, horizons, xus, ees, iters=1)
    lam_re = torch.zeros(14 * T, device="cuda")
    solver.batch_direct_resolve(lam_re, refine=2)
    solver.set_config(flags=pkg.FLAG_DIRECT_SOLVE, direct_refine=2, direct_kkt_refine=0)
    _, lam_dir, _, _, _ = _run_batch(solver, horizons, xus, ees, iters=1)
    assert np.array_equal(host(lam_re), lam_dir)


def test_direct_resolve_after_split_solve(restore, orc, pkg):
    """A split-form solve leaves G and C in the compact form (g_q and the
    lower dynamics rows); the resolve expands them and must give, bitwise,
    the direct mode's solve of the same first Schur systems."""
    solver = restore
    horizons = [32, 200, 64, 256, 128]  # every horizon in reach of the split form
    xus, ees = _batch(orc, pkg, horizons, seed0=3500)
    T = sum(horizons)
    solver.set_config(sqp_max_iter=1)
    _run_batch(solver, horizons, xus, ees, iters=1)
    assert solver.batch_get_plan()["split"] == 1
    lam_re = torch.zeros(14 * T, device="cuda")
    solver.batch_direct_resolve(lam_re, refine=2)
    lam_re2 = torch.zeros(14 * T, device="cuda")
    solver.batch_direct_resolve(lam_re2, refine=2)  # the compact form is still there
    solver.set_config(flags=pkg.FLAG_DIRECT_SOLVE, direct_refine=2, direct_kkt_refine=0)
    _, lam_dir, _, _, _ = _run_batch(solver, horizons, xus, ees, iters=1)
    assert np.array_equal(host(lam_re), lam_dir)
    assert np.array_equal(host(lam_re2), lam_dir)
