"""GPU: the batched SQP solve as a captured HIP graph (mpc_batch_set_graph).
A replay runs the same kernels on the same data as the eager launch
sequence, so its results are bitwise equal to it, on a layout with two-
and four-workgroup teams and resident bins in the team launch.  Replays are checked one by one with
eager work and synchronisations between them; a change of the iteration
count or of the buffers re-captures.  The closed loop with a team-sized
horizon (N = 200) also replays its control step as a graph."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HORIZONS = [256, 129, 200, 64, 256, 3, 300, 45, 128, 17]


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().copy()


@pytest.fixture(scope="module")
def gsolver(pkg):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    T = sum(HORIZONS)
    s = pkg.Solver(pkg.default_config(max_knots=512, max_instances=len(HORIZONS), max_total_knots=T))
    yield s
    s.close()


def _inputs(pkg, orc, horizons, seed0=900):
    xus, ees = [], []
    for i, N in enumerate(horizons):
        x0, qref = pkg.synth_instance(seed0 + i, N)
        ees.append(np.concatenate([orc.fk(q.astype(np.float64)) for q in qref]).astype(np.float32))
        xu = np.zeros(21 * N - 7, np.float32)
        for k in range(N):
            xu[21 * k:21 * k + 14] = x0
        xus.append(xu)
    return np.concatenate(xus), np.concatenate(ees)


class Bufs:
    def __init__(self, xu0, n, T, K):
        self.xu0 = torch.from_numpy(xu0).cuda()
        self.xu = self.xu0.clone()
        self.lam = torch.zeros(14 * T, device="cuda")
        self.rho = torch.full((n,), 1e-3, device="cuda")
        self.st = torch.zeros(n, dtype=torch.int32, device="cuda")
        self.pit = torch.zeros(n * K, dtype=torch.int32, device="cuda")

    def reset(self):
        # in place: the graph is bound to these buffers
        self.xu.copy_(self.xu0)
        self.lam.zero_()
        self.rho.fill_(1e-3)
        self.st.zero_()
        self.pit.zero_()

    def result(self):
        return [host(t) for t in (self.xu, self.lam, self.rho, self.st, self.pit)]


def test_batched_graph_replays_equal_eager(pkg, orc, gsolver):
    n, T = len(HORIZONS), sum(HORIZONS)
    xu0, ee_h = _inputs(pkg, orc, HORIZONS)
    ee = torch.from_numpy(ee_h).cuda()
    s = gsolver
    s.batch_setup(HORIZONS)
    lay = s.batch_pcg_layout()
    m = s.batch_pcg_launch_map()
    L = lay["team_launches"]
    print("layout", lay, "launch map", m.tolist())
    h = np.asarray(HORIZONS)
    assert lay["teams"] == ((h > 128) & (h <= 512)).sum() and (m[h <= 128] == L - 1).any()  # bins in the team launch
    for K in (3, 2):
        b = Bufs(xu0, n, T, K)
        s.batch_set_graph(False)
        s.batch_sqp(b.xu, b.lam, ee, b.rho, K, b.st, b.pit)
        eager = b.result()
        assert (eager[3] & (8 | 16)).max() == 0, eager[3]  # no non-finite step, no PCG fault
        s.batch_set_graph(True)
        for rep in range(3):
            b.reset()
            torch.cuda.synchronize()
            s.batch_sqp(b.xu, b.lam, ee, b.rho, K, b.st, b.pit)
            got = b.result()
            for name, x, y in zip(("xu", "lam", "rho", "status", "pcg_iters"), got, eager):
                assert np.array_equal(x, y), (K, rep, name)
            # eager work between replays
            _ = (torch.randn(1 << 16, device="cuda") * 2).sum().item()
        s.batch_set_graph(False)
    # new buffers: a fresh capture bound to them
    b2 = Bufs(xu0, n, T, 3)
    s.batch_set_graph(True)
    s.batch_sqp(b2.xu, b2.lam, ee, b2.rho, 3, b2.st, b2.pit)
    got = b2.result()
    s.batch_set_graph(False)
    b3 = Bufs(xu0, n, T, 3)
    s.batch_sqp(b3.xu, b3.lam, ee, b3.rho, 3, b3.st, b3.pit)
    for name, x, y in zip(("xu", "lam", "rho", "status", "pcg_iters"), got, b3.result()):
        assert np.array_equal(x, y), name


@pytest.mark.parametrize("split", [0, 1])
def test_closed_loop_graph_with_teams(pkg, split):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    N, steps, K = 200, 6, 2
    s = pkg.Solver(pkg.default_config(max_knots=N, max_instances=2, max_total_knots=2 * N, pcg_exit_tol=5e-6))
    try:
        s.batch_set_plan(split=split)  # N = 200: a two-workgroup team, or one split-form bin
        out = {}
        for graph in (1, 0):
            x0, ee, traj = s.track_inputs(41, steps, N)
            xu = traj[:21 * N - 7].clone()
            lam = torch.zeros(14 * N, device="cuda")
            pcg = torch.zeros(steps * K, dtype=torch.int32, device="cuda")
            cl = pkg.default_closed_loop(steps=steps, sqp_iters=K, use_graph=graph)
            states, err, _, stats = s.closed_loop(N, x0, ee, xu, lam, cl, traj, pcg_iters=pcg)
            assert stats["graph"] == graph
            lay = s.batch_pcg_layout()
            assert (lay["teams"], lay["bins"]) == ((0, 1) if split else (1, 0)), lay
            out[graph] = [host(t) for t in (states, err, xu, lam, pcg)]
        for name, a, b in zip(("states", "err", "xu", "lam", "pcg"), out[1], out[0]):
            assert np.array_equal(a, b), name
        assert np.all(np.isfinite(out[1][0])) and (out[1][4] > 0).any()
    finally:
        s.close()
