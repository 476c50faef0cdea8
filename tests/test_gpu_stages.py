"""GPU parity of every stage entry point against the CPU oracle (fp64).

Each test calls the HIP library through the C-ABI (ctypes) and compares with
the oracle on the same seeded inputs.  Tolerances are fp32-vs-fp64 and are
written per test."""
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


def test_fk_golden_and_oracle(solver, orc):
    g = np.load(f"{helpers.GOLDEN}/fk_0_0.npz")
    q = g["q"].astype(np.float32)
    ee = host(solver.fk(dev(q)))
    # golden file has 6 significant digits
    assert np.abs(ee[:, :3] - g["ee"][:, :3]).max() < 2e-5
    ref = np.stack([orc.fk(qq.astype(np.float64)) for qq in q])
    assert np.abs(ee - ref).max() < 2e-5


def test_dynamics_gradient(solver, orc):
    rng = np.random.default_rng(3)
    pts = np.concatenate([rng.uniform(-1.5, 1.5, (37, 7)), rng.normal(0, 1.0, (37, 7)),
                          rng.normal(0, 5.0, (37, 7))], 1).astype(np.float32)
    qdd, df = solver.dynamics_gradient(dev(pts))
    qdd, df = host(qdd), host(df).reshape(-1, 21, 7).transpose(0, 2, 1)
    for i, p in enumerate(pts.astype(np.float64)):
        rq, rdf = orc.fd_grad(p[:7], p[7:14], p[14:])
        assert helpers.rel_err(qdd[i], rq) < 2e-4, i
        assert helpers.rel_err(df[i], rdf) < 2e-4, i


def _kkt_inputs(orc, seed, N):
    x0, ee, xu = helpers.perturbed_instance(orc, seed, N)
    return xu, ee, xu[:14].copy()


@pytest.mark.parametrize("N,flag", [(2, 0), (5, 0), (32, 0), (32, 1)])
def test_form_kkt(solver, orc, pkg, N, flag):
    solver.set_config(flags=flag)
    try:
        xu, ee, xs = _kkt_inputs(orc, 10 + N, N)
        G = torch.zeros(245 * N - 49, device="cuda")
        Cm = torch.zeros(294 * (N - 1), device="cuda")
        g = torch.zeros(21 * N - 7, device="cuda")
        c = torch.zeros(14 * N, device="cuda")
        solver.form_kkt(N, dev(xu), dev(ee), dev(xs), G, Cm, g, c)
        cfg = orc.make_cfg(last_knot_own_state=flag)
        rG, rC, rg, rc = orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xs.astype(np.float64))
        assert helpers.rel_err(host(G), rG) < 1e-4
        assert helpers.rel_err(host(Cm), rC) < 1e-4
        assert helpers.rel_err(host(g), rg) < 1e-4
        assert np.abs(host(c) - rc).max() < 1e-4 * max(1.0, np.abs(rc).max())
    finally:
        solver.set_config(flags=0)


def _schur_inputs(orc, seed, N):
    xu, ee, xs = _kkt_inputs(orc, seed, N)
    cfg = orc.make_cfg()
    return orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xs.astype(np.float64))


@pytest.mark.parametrize("N,rho", [(2, 1e-3), (7, 1e-3), (32, 1e-3), (32, 0.5)])
def test_form_schur(solver, orc, N, rho):
    G, Cm, g, c = _schur_inputs(orc, 20 + N, N)
    dG = dev(G)
    S = torch.full((588 * N,), 7.0, device="cuda")
    P = torch.full((588 * N,), 7.0, device="cuda")
    gam = torch.zeros(14 * N, device="cuda")
    solver.form_schur(N, dG, dev(Cm), dev(g), dev(c), S, P, gam, rho)
    G32, C32, g32, c32 = (np.asarray(a, np.float32).astype(np.float64) for a in (G, Cm, g, c))
    rS, rP, rgam, rGi = orc.form_schur(N, G32, C32, g32, c32, rho)
    assert helpers.rel_err(host(S), rS) < 5e-4
    assert helpers.rel_err(host(P), rP) < 5e-4
    assert helpers.rel_err(host(gam), rgam) < 5e-4
    assert helpers.rel_err(host(dG), rGi) < 5e-4


def test_form_schur_unaligned_outputs(solver, orc):
    """The Schur blocks leave the kernel as float4 rows when S and Pinv are
    16-byte aligned and as plain stores otherwise; S and Pinv one float off
    alignment must give the same bits."""
    N = 32
    G, Cm, g, c = _schur_inputs(orc, 52, N)
    outs = []
    for off in (0, 1):
        dG = dev(G)
        Sb = torch.full((588 * N + 4,), 7.0, device="cuda")
        Pb = torch.full((588 * N + 4,), 7.0, device="cuda")
        S, P = Sb[off:off + 588 * N], Pb[off:off + 588 * N]
        assert (S.data_ptr() % 16 == 0) == (off == 0) and (P.data_ptr() % 16 == 0) == (off == 0)
        gam = torch.zeros(14 * N, device="cuda")
        solver.form_schur(N, dG, dev(Cm), dev(g), dev(c), S, P, gam, 1e-3)
        outs.append([host(t) for t in (S, P, gam, dG)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def _natural_schur(orc, seed, N):
    """Schur system of a synthetic instance at its initial iterate."""
    x0, ee, xu = helpers.make_instance(orc, seed, N)
    cfg = orc.make_cfg()
    G, Cm, g, c = orc.form_kkt(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xu[:14].astype(np.float64))
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    return (np.asarray(a, np.float32) for a in (S, P, gam))


def _gpu_pcg(solver, N, S32, P32, gam32, max_iter, tol):
    lam = torch.zeros(14 * N, device="cuda")
    it = torch.zeros(1, dtype=torch.int32, device="cuda")
    fl = torch.zeros(1, dtype=torch.int32, device="cuda")
    solver.pcg(N, dev(S32), dev(P32), dev(gam32), lam, max_iter, tol, it, fl)
    return host(lam), int(it.item()), int(fl.item())


def _fp32_floor(orc, N, S32, P32, gam32, tol, ref):
    """Rounding floor of an fp32 PCG for this system: the deviation of the
    fp32 oracle twin (same algorithm, sequential sums) from ref.  Any fp32
    implementation, the reference's own included, scatters at this level."""
    lt, _, _ = orc.pcg(N, S32, P32, gam32, np.zeros(14 * N), 5000, tol, dtype=np.float32)
    return helpers.rel_err(lt, ref)


# Same exit rule on both sides (|r^T Pinv r| < 1e-5): fp32 GPU vs the fp64
# oracle PCG.  Bound: 1e-4, or 4 x the fp32 rounding floor where that floor
# (the oracle's fp32 twin against its fp64 run) is above 2.5e-5: past ~250
# iterations any fp32 PCG scatters beyond 1e-4 (measured on MI355X: 1.3e-6
# .. 3.3e-5 up to N = 128, 1.3e-4 at N = 256 after 851 iterations, where the
# twin's own floor is 3.7e-4).
@pytest.mark.parametrize("N", [2, 16, 32, 64, 128, 256])
def test_pcg_same_rule_parity(solver, orc, N):
    S32, P32, gam32 = _natural_schur(orc, 30 + N, N)
    lam_g, it_g, fl_g = _gpu_pcg(solver, N, S32, P32, gam32, 5000, 1e-5)
    lam_o, it_o, fl_o = orc.pcg(N, S32, P32, gam32, np.zeros(14 * N), 5000, 1e-5)
    lam_d = np.linalg.solve(orc.dense_S(N, S32), gam32.astype(np.float64))
    floor = _fp32_floor(orc, N, S32, P32, gam32, 1e-5, lam_o)
    err = helpers.rel_err(lam_g, lam_o)
    print(f"N={N} iters gpu={it_g} oracle={it_o} rel(gpu,oracle)={err:.2e} fp32-floor={floor:.2e} "
          f"rel(gpu,direct)={helpers.rel_err(lam_g, lam_d):.2e} rel(oracle,direct)={helpers.rel_err(lam_o, lam_d):.2e}")
    assert fl_g == 0 and fl_o == 0
    assert abs(it_g - it_o) <= max(1, it_o // 50)
    assert err < max(1e-4, 4 * floor)


# Convergence to the direct solve: run to |r^T Pinv r| < 1e-9 (hundreds to
# ~1100 iterations).  At that depth every fp32 PCG scatters by rounding
# alone: on these systems the GPU kernels and the fp32 oracle twin each land
# 5e-6 .. 2e-4 from the direct solution, in no fixed order.  So the checks
# are (a) the true residual |gamma - S lambda| / |gamma| of the GPU solution,
# evaluated in fp64, is within 2x the twin's -- an error in the arithmetic
# shows up there first -- and (b) the error stays below 2e-4 (measured on
# MI355X: 5.6e-6 .. 1.0e-4).  At the SQP's tolerance 1e-5 (absolute) no PCG
# is within 1e-4 of the direct solution (the fp64 oracle PCG itself is 2e-4
# .. 6e-3 away): that comparison is made at the same rule, above.
@pytest.mark.parametrize("N", [2, 16, 32, 64, 128, 200, 256])
def test_pcg_converges_to_direct_solve(solver, orc, N):
    S32, P32, gam32 = _natural_schur(orc, 30 + N, N)
    lam_g, it_g, fl_g = _gpu_pcg(solver, N, S32, P32, gam32, 5000, 1e-9)
    ldl = orc.LDL(N)
    lam_d = ldl.solve(S32.astype(np.float64), gam32.astype(np.float64))
    lam_t, _, _ = orc.pcg(N, S32, P32, gam32, np.zeros(14 * N), 5000, 1e-9, dtype=np.float32)
    Sd = orc.dense_S(N, S32.astype(np.float64))
    g64 = gam32.astype(np.float64)

    def res(lam):
        return np.linalg.norm(g64 - Sd @ lam) / np.linalg.norm(g64)

    err, err_t = helpers.rel_err(lam_g, lam_d), helpers.rel_err(lam_t, lam_d)
    print(f"N={N} iters={it_g} rel(gpu,direct)={err:.2e} residual={res(lam_g):.2e} | "
          f"fp32 twin: rel={err_t:.2e} residual={res(lam_t):.2e}")
    assert fl_g == 0
    assert res(lam_g) <= 2 * res(lam_t) + 1e-7
    assert err < 2e-4


# The on-chip PCG kernels (resident for N <= 128, teams of two workgroups
# for N <= 256 and of four for N <= 512) and the streaming kernel follow the
# oracle's iterations (exactly on one workgroup, within 2 % across the
# teams' split reductions) and agree with it to the fp32 floor on several
# seeds.
@pytest.mark.parametrize("N", [16, 64, 128, 200, 400, 512])
def test_pcg_kernels_agree(solver, orc, pkg, N):
    for seed in range(3):
        S32, P32, gam32 = _natural_schur(orc, 500 + 11 * seed + N, N)
        lam_o, it_o, _ = orc.pcg(N, S32, P32, gam32, np.zeros(14 * N), 5000, 1e-5)
        floor = _fp32_floor(orc, N, S32, P32, gam32, 1e-5, lam_o)
        out = []
        for flag in (0, pkg.FLAG_PCG_STREAM):
            solver.set_config(flags=flag)
            try:
                out.append(_gpu_pcg(solver, N, S32, P32, gam32, 5000, 1e-5))
            finally:
                solver.set_config(flags=0)
        for lam_g, it_g, fl_g in out:
            assert fl_g == 0 and (it_g == it_o if N <= 128 else abs(it_g - it_o) <= max(2, it_o // 50))
            assert helpers.rel_err(lam_g, lam_o) < max(1e-4, 4 * floor)


# Ill-conditioned stress: perturbed trajectories with large controls need
# ~10x more iterations; GPU and oracle still follow the same iteration, and
# agree to 1e-4 or 4 x the fp32 floor (measured 6.5e-6 at N = 16, 2.7e-4 at
# N = 128 after 1325 iterations).
@pytest.mark.parametrize("N", [16, 128])
def test_pcg_stress_same_rule(solver, orc, N):
    G, Cm, g, c = _schur_inputs(orc, 30 + N, N)
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    S32, P32, gam32 = (np.asarray(a, np.float32) for a in (S, P, gam))
    lam_g, it_g, fl_g = _gpu_pcg(solver, N, S32, P32, gam32, 5000, 1e-5)
    lam_o, it_o, fl_o = orc.pcg(N, S32, P32, gam32, np.zeros(14 * N), 5000, 1e-5)
    floor = _fp32_floor(orc, N, S32, P32, gam32, 1e-5, lam_o)
    err = helpers.rel_err(lam_g, lam_o)
    print(f"stress N={N} iters gpu={it_g} oracle={it_o} rel={err:.2e} fp32-floor={floor:.2e}")
    assert fl_g == fl_o == 0 and abs(it_g - it_o) <= max(2, it_o // 50)
    assert err < max(1e-4, 4 * floor)


def test_pcg_maxiter_and_warm_start(solver, orc):
    N = 16
    G, Cm, g, c = _schur_inputs(orc, 41, N)
    S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3)
    S32, P32, gam32 = (np.asarray(a, np.float32) for a in (S, P, gam))
    lam = torch.zeros(14 * N, device="cuda")
    it = torch.zeros(1, dtype=torch.int32, device="cuda")
    fl = torch.zeros(1, dtype=torch.int32, device="cuda")
    # tol 0 can never be met: exactly max_iter iterations, flag set
    solver.pcg(N, dev(S32), dev(P32), dev(gam32), lam, 3, 0.0, it, fl)
    assert int(it.item()) == 3 and int(fl.item()) == 1
    # warm start at a converged solution: zero iterations
    solver.pcg(N, dev(S32), dev(P32), dev(gam32), lam, 2000, 1e-5, it, fl)
    lam2 = lam.clone()
    solver.pcg(N, dev(S32), dev(P32), dev(gam32), lam2, 2000, 1e-5, it, fl)
    assert int(it.item()) == 0 and int(fl.item()) == 0
    assert torch.equal(lam, lam2)


@pytest.mark.parametrize("N", [2, 9, 32])
def test_compute_dz(solver, orc, N):
    G, Cm, g, c = _schur_inputs(orc, 50 + N, N)
    S, P, gam, Gi = orc.form_schur(N, G, Cm, g, c, 1e-3)
    lam = np.linalg.solve(orc.dense_S(N, S), gam)
    Gi32, C32, g32, l32 = (np.asarray(a, np.float32) for a in (Gi, Cm, g, lam))
    dz = torch.zeros(21 * N - 7, device="cuda")
    solver.compute_dz(N, dev(Gi32), dev(C32), dev(g32), dev(l32), dz)
    ref = orc.compute_dz(N, *(a.astype(np.float64) for a in (Gi32, C32, g32, l32)))
    # elementwise bound: 64 fp32 ulps of the absolute-value evaluation
    Ga = orc.dense_G(N, np.abs(Gi32.astype(np.float64)))
    Ca = orc.dense_C(N, np.abs(C32.astype(np.float64)))
    bound = Ga @ (np.abs(g32.astype(np.float64)) + np.abs(Ca).T @ np.abs(l32.astype(np.float64)))
    assert np.all(np.abs(host(dz) - ref) <= 64 * 2.0 ** -24 * bound + 1e-30)


@pytest.mark.parametrize("N", [2, 8, 32])
def test_merits(solver, orc, N):
    xu, ee, xs = _kkt_inputs(orc, 60 + N, N)
    rng = np.random.default_rng(N)
    dz = rng.normal(0, 0.05, 21 * N - 7).astype(np.float32)
    cfg = orc.make_cfg()
    out = torch.zeros(8, device="cuda")
    solver.ls_compute_merit(N, dev(xs), dev(xu), dev(ee), dev(dz), out)
    ref = orc.ls_merits(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xs.astype(np.float64),
                        dz.astype(np.float64))
    assert helpers.rel_err(host(out), ref) < 1e-4
    m0 = torch.zeros(1, device="cuda")
    solver.compute_merit(N, dev(xu), dev(ee), m0)
    r0, _ = orc.merit(cfg, N, xu.astype(np.float64), ee.astype(np.float64), xs.astype(np.float64))
    assert abs(host(m0)[0] - r0) < 1e-4 * abs(r0)
