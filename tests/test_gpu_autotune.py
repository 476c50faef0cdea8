"""GPU: launch plans of the batched PCG and the autotuner that picks one.
A plan only changes how instances are packed into workgroups (bin shape
512 or 256 lanes, instances per bin, bins in the last team launch), so on
bins and teams every plan gives bitwise the same solve; routing the
128 < N <= 256 horizons to the streaming kernel changes its arithmetic, so
there only the solve's health is checked.  The autotuner times every plan
that differs for the layout and keeps the fastest."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HORIZONS = [32, 64, 17, 128, 2, 45, 256, 129, 200, 100, 64, 300, 33, 48]


@pytest.fixture(scope="module")
def tsolver(pkg):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    s = pkg.Solver(pkg.default_config(max_knots=512, max_instances=len(HORIZONS), max_total_knots=sum(HORIZONS)))
    yield s
    s.close()


def _solve(pkg, orc, s, K=3):
    n, T = len(HORIZONS), sum(HORIZONS)
    xus, ees = [], []
    for i, N in enumerate(HORIZONS):
        x0, qref = pkg.synth_instance(700 + i, N)
        ees.append(np.concatenate([orc.fk(q.astype(np.float64)) for q in qref]).astype(np.float32))
        xu = np.zeros(21 * N - 7, np.float32)
        for k in range(N):
            xu[21 * k:21 * k + 14] = x0
        xus.append(xu)
    xu = torch.from_numpy(np.concatenate(xus)).cuda()
    ee = torch.from_numpy(np.concatenate(ees)).cuda()
    lam = torch.zeros(14 * T, device="cuda")
    rho = torch.full((n,), 1e-3, device="cuda")
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    pit = torch.zeros(n * K, dtype=torch.int32, device="cuda")
    s.batch_sqp(xu, lam, ee, rho, K, st, pit)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (xu, lam, rho, st, pit)]


DEFAULT = dict(small_nt=512, max_items=8, mixed=1, teams=1, split=1)  # split: no effect here (N = 300)


@pytest.mark.parametrize("plan", [dict(small_nt=256), dict(max_items=1), dict(max_items=2, small_nt=256),
                                  dict(mixed=0), dict(mixed=0, small_nt=256, max_items=4)])
def test_packing_plans_are_bitwise_equal(pkg, orc, tsolver, plan):
    s = tsolver
    s.batch_setup(HORIZONS)
    s.batch_set_plan(**DEFAULT)
    ref = _solve(pkg, orc, s)
    s.batch_set_plan(**dict(DEFAULT, **plan))
    assert s.batch_get_plan() == dict(DEFAULT, **plan)
    m = s.batch_pcg_launch_map()
    h = np.asarray(HORIZONS)
    if plan.get("small_nt") == 256:
        assert len(set(m[h <= 64])) == 1 and m[h <= 64][0] == m.max()  # the last launch: 256-lane bins
    got = _solve(pkg, orc, s)
    for name, a, b in zip(("xu", "lam", "rho", "status", "pcg_iters"), got, ref):
        assert np.array_equal(a, b), (plan, name)
    s.batch_set_plan(**DEFAULT)


def test_streamed_team_horizons_solve(pkg, orc, tsolver):
    s = tsolver
    s.batch_setup(HORIZONS)
    s.batch_set_plan(**dict(DEFAULT, teams=0))
    lay = s.batch_pcg_layout()
    assert lay["teams"] == 0 and lay["streamed"] == sum(1 for n in HORIZONS if n > 128)
    xu, lam, rho, st, pit = _solve(pkg, orc, s)
    assert np.all(np.isfinite(xu)) and (st & (8 | 16)).max() == 0 and (pit > 0).any()
    s.batch_set_plan(**DEFAULT)


def test_autotune_picks_the_fastest_plan(pkg, orc, tsolver):
    s = tsolver
    s.batch_setup(HORIZONS)
    s.batch_set_plan(**DEFAULT)
    res = s.batch_autotune(iters=8, reps=2)
    print("\n".join(str(r) for r in res))
    # bin shape x instances per bin x (pool or static launches) x (teams or streamed)
    assert len(res) == 2 * 3 * 4
    assert all(r["us"] > 0 for r in res)
    best = min(res, key=lambda r: r["us"])
    assert s.batch_get_plan() == {k: best[k] for k in DEFAULT}
    # the tuned plan solves like any other
    xu, lam, rho, st, pit = _solve(pkg, orc, s)
    assert np.all(np.isfinite(xu)) and (st & (8 | 16)).max() == 0
    with pytest.raises(pkg.MpcError):
        s.batch_set_plan(small_nt=128)
    s.batch_set_plan(**DEFAULT)
