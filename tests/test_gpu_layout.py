"""GPU: the PCG launch plan of a batch.  Long horizons run on cooperative
team launches (four-workgroup teams for 256 < N <= 512 first, then
two-workgroup teams for 128 < N <= 256); the last two-workgroup launch also
solves the first (longest) resident bins in its spare workgroups; the
remaining bins run on the resident kernel.  Every instance must be solved
exactly once, whichever launch holds it."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pool", [1, 0])
def test_launch_map_and_every_instance_solved(pkg, orc, pool):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    # 140 two-workgroup teams (more than one launch's 128), two four-workgroup
    # teams, bins of every packing (one 128, two 64s, several short instances)
    horizons = [200] * 140 + [64] * 20 + [128] * 10 + [300] * 2 + [3] * 5
    rng = np.random.default_rng(7)
    horizons = [int(h) for h in rng.permutation(horizons)]
    n, T = len(horizons), sum(horizons)
    cfg = pkg.default_config(max_knots=512, max_instances=n, max_total_knots=T)
    s = pkg.Solver(cfg)
    try:
        s.batch_setup(horizons)
        s.batch_set_plan(mixed=pool)
        lay = s.batch_pcg_layout()
        L = lay["team_launches"]
        m = s.batch_pcg_launch_map()
        h = np.asarray(horizons)
        print("layout", lay, "launch counts", np.bincount(m))
        team, quad, res = (h > 128) & (h <= 256), (h > 256) & (h <= 512), h <= 128
        assert lay["teams"] == team.sum() + quad.sum() and lay["streamed"] == 0
        # four-workgroup teams first
        assert (m[quad] == 0).all()
        if pool:
            # one persistent launch for the two-workgroup teams and the bins
            assert L == 2 and (m[team] == 1).all() and (m[res] == 1).all()
        else:
            # team launches in instance order, at most one launch's share each,
            # then one launch of the bins
            assert L == 3 and (m[team] >= 1).all() and (m[team] < L).all()
            per = np.bincount(m[team] - 1, minlength=L - 1)
            assert per[:-1].min() == per.max() and per[-1] <= per[0]
            assert np.all(np.diff(m[team].astype(np.int64)) >= 0)
            assert (m[res] == L).all()
        # a batched solve touches every instance: each one's PCG ran (or it
        # started converged) and its multipliers were written
        xu = torch.zeros(21 * T - 7 * n, device="cuda")
        ee = torch.zeros(6 * T, device="cuda")
        k0 = np.concatenate([[0], np.cumsum(h)])
        xo = np.concatenate([[0], np.cumsum(21 * h - 7)])
        for i, N in enumerate(horizons):
            x0, qref = pkg.synth_instance(500 + i, N)
            e = np.concatenate([orc.fk(q.astype(np.float64)) for q in qref]).astype(np.float32)
            ee[6 * k0[i]:6 * k0[i + 1]] = torch.from_numpy(e).cuda()
            xi = np.zeros(21 * N - 7, np.float32)
            for k in range(N):
                xi[21 * k:21 * k + 14] = x0
            xu[xo[i]:xo[i + 1]] = torch.from_numpy(xi).cuda()
        lam = torch.zeros(14 * T, device="cuda")
        rho = torch.full((n,), 1e-3, device="cuda")
        st = torch.zeros(n, dtype=torch.int32, device="cuda")
        pit = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        s.batch_sqp(xu, lam, ee, rho, 1, st, pit)
        torch.cuda.synchronize()
        pit, lam_h = pit.cpu().numpy(), lam.cpu().numpy()
        assert (pit > 0).all(), np.nonzero(pit <= 0)
        for i in range(n):
            assert np.abs(lam_h[14 * k0[i]:14 * k0[i + 1]]).max() > 0, (i, horizons[i], m[i])
        assert (st.cpu().numpy() & ~3).max() == 0
    finally:
        s.close()
