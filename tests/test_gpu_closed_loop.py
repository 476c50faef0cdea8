"""GPU closed-loop harness (C-ABI mpc_closed_loop) against the oracle's
restatement (orc_closed_loop, itself pinned in test_oracle_closed_loop.py),
graph replay against stream launches, and the full closed-loop workload
(N = 128, 2000 control steps, PCG tolerance 5e-6, noise 1e-3).

Tolerances: the SQP's PCG exits at an absolute |r^T Pinv r| < 5e-6, so any
fp32 evaluation of the loop (the GPU's, or the oracle's fp32 twin) leaves
the fp64 one by the PCG solution error, which the plant states carry on
(mostly the joint velocities).  The GPU may differ from the fp64 oracle by
max(1e-3, 4 x the fp32 twin's own distance to it), states and errors alike;
graph replay is the same kernels on the same data, so it is bitwise
equal."""
import numpy as np
import pytest

import helpers

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def cl_solver(pkg):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    s = pkg.Solver(pkg.default_config(max_knots=128, max_instances=4, max_total_knots=512, pcg_exit_tol=5e-6))
    yield s
    s.close()


def _run(pkg, s, N, steps, seed, K=3, sigma=1e-3, graph=1, timing=0, with_traj=True):
    x0, ee, traj = s.track_inputs(seed, steps, N)
    xu = traj[:21 * N - 7].clone()
    lam = torch.zeros(14 * N, device="cuda")
    pcg = torch.zeros(steps * K, dtype=torch.int32, device="cuda")
    cl = pkg.default_closed_loop(steps=steps, sqp_iters=K, noise_sigma=sigma, use_graph=graph, timing=timing)
    states, err, _, stats = s.closed_loop(N, x0, ee, xu, lam, cl, traj if with_traj else None, pcg_iters=pcg)
    return dict(x0=host(x0), ee=host(ee), traj=host(traj), states=host(states), err=host(err), xu=host(xu),
                lam=host(lam), pcg=host(pcg).reshape(steps, K), stats=stats)


@pytest.mark.parametrize("N,with_traj", [(16, True), (32, True), (32, False)])
def test_closed_loop_matches_oracle(pkg, orc, cl_solver, N, with_traj):
    steps, K = 8, 3
    g = _run(pkg, cl_solver, N, steps, 300 + N, K=K, with_traj=with_traj)
    cfg = orc.make_cfg()
    cfg.sqp_max_iter = K
    cfg.pcg_tol = 5e-6
    xu0 = g["traj"][:21 * N - 7].copy()
    run = lambda dt: orc.closed_loop(cfg, N, steps, g["x0"], g["ee"], xu0, np.zeros(14 * N), sigma=1e-3, seed=12345,
                                     traj=g["traj"] if with_traj else None, dtype=dt)
    o, o32 = run(np.float64), run(np.float32)
    ds = np.abs(g["states"] - o["states"]).max()
    de = np.abs(g["err"] - o["ee_err"]).max()
    fs = np.abs(o32["states"] - o["states"]).max()
    fe = np.abs(o32["ee_err"] - o["ee_err"]).max()
    print(f"N={N} traj={with_traj}: GPU vs fp64 oracle: states {ds:.2e}, errors {de:.2e}; fp32 twin vs fp64: "
          f"states {fs:.2e}, errors {fe:.2e}; pcg gpu {g['pcg'][:, 0].tolist()} oracle {o['pcg_iters'][:, 0].tolist()}")
    assert np.all(np.isfinite(g["states"]))
    assert ds <= max(1e-3, 4 * fs) and de <= max(1e-3, 4 * fe)


def test_graph_replay_equals_stream_launches(pkg, cl_solver):
    a = _run(pkg, cl_solver, 32, 12, 77, graph=1)
    b = _run(pkg, cl_solver, 32, 12, 77, graph=0)
    assert a["stats"]["graph"] == 1 and b["stats"]["graph"] == 0
    for k in ("states", "err", "xu", "lam", "pcg"):
        assert np.array_equal(a[k], b[k]), k


def test_timing_mode_reports_phases(pkg, cl_solver):
    r = _run(pkg, cl_solver, 32, 5, 78, graph=0, timing=1)
    st = r["stats"]
    assert st["total_us"] > 0 and st["plant_us"] > 0 and st["solve_us"] > st["linsys_us"] > 0
    assert st["total_us"] >= 0.9 * (st["solve_us"] + st["plant_us"])


def test_bad_arguments_are_rejected(pkg, cl_solver):
    x0, ee, traj = cl_solver.track_inputs(5, 4, 16)
    xu, lam = traj[:21 * 16 - 7].clone(), torch.zeros(14 * 16, device="cuda")
    for kw in ({"sqp_iters": 0}, {"substeps": 0}, {"sim_dt": 0.0}, {"noise_sigma": -1.0},
               {"use_graph": 1, "timing": 1}, {"steps": 0}):
        cl = pkg.default_closed_loop(**{"steps": 4, **kw})
        with pytest.raises(pkg.MpcError):
            cl_solver.closed_loop(16, x0, ee, xu, lam, cl, traj)


def test_closed_loop_n128_2000_steps(pkg):
    """The closed-loop workload: finite throughout, deterministic, and the
    tracking error of the reference SQP with its PCG at tol 5e-6."""
    s = pkg.Solver(pkg.default_config(max_knots=128, max_instances=1, max_total_knots=128, pcg_exit_tol=5e-6))
    try:
        r = _run(pkg, s, 128, 2000, 2024)
        r2 = _run(pkg, s, 128, 50, 2024)
    finally:
        s.close()
    e = r["err"]
    rms, l1 = float(np.sqrt(np.mean(e[:, 0] ** 2))), float(e[:, 1].mean())
    print(f"N=128 x 2000 steps: {r['stats']['total_us'] / 2000:.1f} us per control step, RMS EE error {rms:.4f} m, "
          f"mean L1 {l1:.4f} m, PCG iterations mean {r['pcg'].mean():.1f}")
    assert np.all(np.isfinite(r["states"])) and np.all(np.isfinite(e))
    assert np.array_equal(r["states"][:51], r2["states"])
    assert rms < 0.5


@pytest.mark.parametrize("direct", [False, True])
def test_closed_loop_on_reference_trajectory(pkg, direct):
    """The reference's own closed-loop input (testfiles/0_0, dt = 1/64,
    QD = R = 1e-4, no noise) at N = 32, 3 SQP iterations per step: the arm
    follows the recorded trajectory (mean L1 end-effector error measured
    0.050 m with PCG and 0.050 m direct; the reference reports 0.049 m with
    PCG at N = 128)."""
    import importlib
    fx = importlib.import_module(pkg.__name__ + ".fixtures")
    g = np.load(f"{helpers.GOLDEN}/fk_0_0.npz")
    N, dt = 32, 1.0 / 64
    s = pkg.Solver(pkg.default_config(max_knots=N, max_instances=1, max_total_knots=N, dt=dt, qd_cost=1e-4,
                                      r_cost=1e-4, pcg_exit_tol=1e-5, flags=pkg.FLAG_DIRECT_SOLVE if direct else 0))
    try:
        x0, path, traj, steps = fx.closed_loop_inputs(g["xu"].astype(np.float32), g["ee"].astype(np.float32), N)
        xu, lam = traj[:21 * N - 7].clone(), torch.zeros(14 * N, device="cuda")
        cl = pkg.default_closed_loop(steps=steps, sqp_iters=3, substeps=78, sim_dt=dt / 78, noise_sigma=0.0)
        states, err, _, _ = s.closed_loop(N, x0, path, xu, lam, cl, traj)
        e, st = host(err), host(states)
    finally:
        s.close()
    l1 = float(e[:, 1].mean())
    print(f"0_0 trajectory, N=32, {'direct' if direct else 'PCG'}: mean L1 {l1:.4f} m over {steps} steps")
    assert steps == 665 and np.all(np.isfinite(st))
    assert l1 < 0.08
