"""GPU: block-Jacobi PCG in split-preconditioned form (k_pcg_split.hip),
the batched path's PCG wherever every horizon is at most 256.

The kernel runs CG on W^T S W (W_b = chol(-S(b,b))^-1) and maps the
result back; in exact arithmetic its iterates, exit test and iteration
count are those of the reference's PCG (GPU-PCG pcg<T,14,N>, call site
toplevel.cuh:143-157; restated in tests/test_split_form.py).  Here:
- ragged layouts (N = 2 .. 256, up to 8 instances in a bin): after each of
  two SQP iterations, every instance's multipliers against the fp64 oracle
  (Schur formation and PCG) on the GPU's own KKT blocks (the stage entry
  point at the batch's iterate, the batch's rho, warm started from the
  batch's previous multipliers) within 1e-4, or 4 x the fp32 floor of that
  solve (the largest deviation of the oracle's fp32 formation and PCG on G
  and on seven one-ulp perturbations of G, the PCG in the reference form and
  in the split form);
  iteration counts where the fp64 oracle's |r^T z| crosses the tolerance,
  within the band of fp32 rounding (helpers.pcg_exit_ok); the exit state
  (flag, |r^T Pinv r|) as the exit rule defines it;
- every packing of the same layout (instances per bin 8, 4, 2): bitwise
  the same solve (an instance's arithmetic does not depend on its bin);
- an instance with a non-finite reference beside healthy ones: only its
  status changes, the others are bitwise the same;
- the autotuner times the split-form plans beside the others."""
import numpy as np
import pytest

import helpers

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CAP = 173


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


class Batch:
    def __init__(self, pkg, orc, horizons, seed0):
        self.h = [int(N) for N in horizons]
        self.n, self.T = len(self.h), sum(self.h)
        self.K0 = np.r_[0, np.cumsum(self.h)]
        self.X0 = np.r_[0, np.cumsum([21 * N - 7 for N in self.h])]
        xus, ees = [], []
        for i, N in enumerate(self.h):
            x0, qref = pkg.synth_instance(seed0 + i, N)
            ees.append(np.concatenate([orc.fk(q.astype(np.float64)) for q in qref]).astype(np.float32))
            xu = np.zeros(21 * N - 7, np.float32)
            for k in range(N):
                xu[21 * k:21 * k + 14] = x0
            xus.append(xu)
        self.xu0, self.ee = np.concatenate(xus), np.concatenate(ees)

    def run(self, solver, iters, ee=None):
        xu = dev(self.xu0)
        lam = torch.zeros(14 * self.T, device="cuda")
        rho = torch.full((self.n,), 1e-3, device="cuda")
        st = torch.zeros(self.n, dtype=torch.int32, device="cuda")
        pit = torch.zeros(self.n * iters, dtype=torch.int32, device="cuda")
        solver.batch_sqp(xu, lam, dev(self.ee if ee is None else ee), rho, iters, st, pit)
        return host(xu), host(lam), host(rho), host(st).astype(np.int64), host(pit).reshape(self.n, iters)

    def xs(self, i):
        return slice(self.X0[i], self.X0[i + 1])

    def ls(self, i):
        return slice(14 * self.K0[i], 14 * self.K0[i + 1])


@pytest.fixture
def split_solver(pkg):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    s = pkg.Solver(pkg.default_config(max_knots=256, max_instances=32, max_total_knots=4096, sqp_max_iter=2))
    yield s
    s.close()


LAYOUTS = {
    "ragged": [2, 256, 3, 17, 129, 45, 255, 64, 100, 200, 5],
    "full_bins": [32] * 8 + [64] * 4 + [128, 128, 256],
}


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_split_layout_vs_oracle_pcg(split_solver, pkg, orc, name):
    s = split_solver
    B = Batch(pkg, orc, LAYOUTS[name], 5100)
    s.batch_setup(B.h)
    assert s.batch_get_plan()["split"] == 1
    lay = s.batch_pcg_layout()
    assert lay["teams"] == 0 and lay["streamed"] == 0 and lay["team_launches"] == 0
    print("layout", lay)
    states = [(B.xu0, np.zeros(14 * B.T, np.float32), np.full(B.n, 1e-3, np.float32))]
    pits = []
    for k in (1, 2):
        xu, lam, rho, st, pit = B.run(s, k)
        assert (st & ~np.int64(pkg.ST_ACTIVE | pkg.ST_PCG_MAXITER)).max() == 0, st
        states.append((xu, lam, rho))
        pits.append(pit)
    # exit state of the last solve (SQP iteration 2)
    it = torch.zeros(B.n, dtype=torch.int32, device="cuda")
    fl = torch.zeros(B.n, dtype=torch.int32, device="cuda")
    eta = torch.zeros(B.n, device="cuda")
    s.batch_pcg_exit(it, fl, eta)
    it, fl, eta = host(it), host(fl), host(eta)
    tol = s.cfg.pcg_exit_tol
    assert np.array_equal(it, pits[1][:, 1])
    for i in range(B.n):
        assert (fl[i] == 0 and eta[i] < tol) or (fl[i] == 1 and it[i] == CAP and eta[i] >= tol), i
    worst = 0.0
    for i, N in enumerate(B.h):
        G = torch.zeros(245 * N - 49, device="cuda")
        Cm = torch.zeros(294 * (N - 1), device="cuda")
        g = torch.zeros(21 * N - 7, device="cuda")
        c = torch.zeros(14 * N, device="cuda")
        ee = dev(B.ee[6 * B.K0[i]:6 * B.K0[i + 1]])
        xs = dev(B.xu0[B.xs(i)][:14])
        for k in range(2):
            xu_k, lam_k, rho_k = states[k]
            s.form_kkt(N, dev(xu_k[B.xs(i)]), ee, xs, G, Cm, g, c)
            kkt = [host(t).astype(np.float64) for t in (G, Cm, g, c)]
            rho_i = float(rho_k[i])
            S64, P64, g64, _ = orc.form_schur(N, *kkt, rho_i)
            warm = lam_k[B.ls(i)].astype(np.float64)
            lo, it_o, _ = orc.pcg(N, S64, P64, g64, warm, CAP, 1e-5)
            floor = helpers.schur_pcg_fp32_floor(orc, N, *kkt, rho_i, warm, CAP, 1e-5, lo, split=True)
            _, _, _, e64 = orc.pcg(N, S64, P64, g64, warm, CAP, 0.0, trace=True)
            err = helpers.rel_err(states[k + 1][1][B.ls(i)], lo)
            it_g = int(pits[1][i, k])
            worst = max(worst, err)
            print(f"inst {i} N={N} SQP iteration {k}: PCG gpu={it_g} oracle={it_o} rel(lambda)={err:.2e} "
                  f"fp32 floor={floor:.2e} fp64 |eta| at gpu exit={e64[min(it_g, CAP)]:.3e}")
            assert helpers.pcg_exit_ok(it_g, e64, 1e-5, CAP), (i, k, it_g, it_o)
            assert err < max(1e-4, 4 * floor), (i, k)
    print(f"worst rel(lambda) {worst:.2e}")


def test_split_packing_is_bitwise_invariant(split_solver, pkg, orc):
    s = split_solver
    B = Batch(pkg, orc, LAYOUTS["full_bins"] + [17, 3, 200], 5200)
    s.batch_setup(B.h)
    ref = B.run(s, 2)
    bins = s.batch_pcg_layout()["bins"]
    try:
        for mi in (4, 2):
            s.batch_set_plan(max_items=mi)
            assert s.batch_pcg_layout()["bins"] >= bins
            got = B.run(s, 2)
            for name, a, b in zip(("xu", "lam", "rho", "status", "pcg_iters"), got, ref):
                assert np.array_equal(a, b), (mi, name)
    finally:
        s.batch_set_plan(max_items=8)


def test_split_nonfinite_instance_leaves_the_others(split_solver, pkg, orc):
    s = split_solver
    B = Batch(pkg, orc, [32, 64, 32, 128, 17], 5300)
    s.batch_setup(B.h)
    good = B.run(s, 2)
    bad = B.ee.copy()
    bad[6 * B.K0[2] + 7] = np.nan
    nan = B.run(s, 2, ee=bad)
    assert nan[3][2] & pkg.ST_NONFINITE
    for i in range(B.n):
        if i != 2:
            assert not nan[3][i] & pkg.ST_NONFINITE
            assert np.array_equal(nan[0][B.xs(i)], good[0][B.xs(i)]), i
            assert np.array_equal(nan[1][B.ls(i)], good[1][B.ls(i)]), i
    assert np.array_equal(nan[0][B.xs(2)], B.xu0[B.xs(2)])  # no step taken (its multipliers are not finite)


def test_split_autotune(split_solver, pkg, orc):
    s = split_solver
    B = Batch(pkg, orc, LAYOUTS["ragged"], 5400)
    s.batch_setup(B.h)
    res = s.batch_autotune(iters=8, reps=2)
    print("\n".join(str(r) for r in res))
    split = [r for r in res if r["split"]]
    assert len(split) == 3 and all(r["us"] > 0 for r in res)
    best = min(res, key=lambda r: r["us"])
    assert s.batch_get_plan() == {k: best[k] for k in ("small_nt", "max_items", "mixed", "teams", "split")}
    _, lam, _, st, _ = B.run(s, 2)
    assert np.all(np.isfinite(lam)) and (st & (8 | 16)).max() == 0
    s.batch_set_plan(split=1, max_items=8, small_nt=512, mixed=1, teams=1)
