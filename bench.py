"""Benchmark of the batched SQP tracking-MPC inner loop on one MI355X.

Workload (the batched configuration): 1024 independent KUKA iiwa
end-effector tracking instances, horizons drawn uniformly from
{32, 64, 128, 256}, dt = 0.01 s, Q_ee = 1, QD = 1e-2, R = 1e-4, fp32,
block-Jacobi PCG with exit tolerance 1e-5, 3 SQP iterations per solve.
Synthetic seeded inputs (splitmix64): q0 ~ U(-0.5, 0.5), qd0 ~ N(0, 0.1^2),
reference = FK of joint sinusoids (amplitude 0.3 rad, f ~ U(0.2, 1) Hz,
random phase); start trajectory x_k = x0, u_k = 0, lambda = 0, rho = 1e-3.

One step = one batched SQP solve (3 SQP iterations, every phase on the
device) over all instances, starting from the same initial iterate; each
SQP iteration of an instance is one KKT solve (linearisation, Schur +
block-Jacobi, PCG, dz, 8-alpha line search, update).  The solve is
captured once as a HIP graph and replayed; value = KKT solves / s over the
whole step.  The same launch sequence issued without capture is timed
beside it ("graph"), with a bitwise comparison of the two results.  Also reported: the plain
one-instance-per-launch solve (reference call pattern) on a sample, the
per-phase split, the roofline of the dominant kernel and a CPU baseline
(the sparse LDL^T solve of the Schur system, fp32: timed on one core, then
on up to --cpu-threads host threads at once, one instance per thread).

Also reported ("closed_loop"): the closed-loop tracking workload, one
instance at N = 128 for 2000 control steps (PCG tolerance 5e-6, state
noise 1e-3, plant = 50 Euler substeps of 2e-4 s per step), run on the
device with one captured control step replayed as a HIP graph: time per
control step, RMS / mean-L1 end-effector error, and from a separate
event-timed run the per-SQP-iteration linear-solve time.

Also reported ("differentiation"): the adjoint / forward cost of the
differentiable QP layer on 256 instances at N = 64 (same cost and inputs):
forward = the QP layer at the final SQP iterate plus the tracking loss and
its gradient, adjoint = recomputed linearisation, adjoint KKT solve and the
parameter gradients (21 weights + all reference points); both with two
steps of iterative refinement.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "gpu-pcg-tracking_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))  # the CPU baseline leg only

HBM_PEAK_GBS = 8000.0
PHASES = ("linearize", "schur", "pcg_team", "pcg_rest", "dz", "merit", "update")
# algorithmic bytes / flops of one PCG iteration per knot (fp32): S diagonal
# upper triangle 105 + one off-diagonal block 196 + Jacobi upper triangle 105
# floats plus 12 x 14 vector floats; 1708 flops (S p 1176, Pinv r 392, two
# dots 56, three axpys 84)
PCG_BYTES_PER_KNOT_ITER = 4 * (105 + 196 + 105 + 12 * 14)
PCG_FLOPS_PER_KNOT_ITER = 1708
ENGINE_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
FP32_PEAK_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_batched_1024.json")
# per-kernel counters of the batched workload (rocprofv3 --pmc: LDS and VALU
# instruction counts per launch, the LDS bytes per instruction of the PCG
# loop body from the kernel's own ISA); see the .md beside it
COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters_batched_1024.json")
# algorithmic HBM bytes and flops per knot of the other phases of an SQP
# iteration (fp32), as the batched split path runs them:
# linearize: reads x_k, u_k, x_{k+1}, ref_k (21 + 14 + 6 floats), writes
#   the lower block rows of C (147: the upper ones are fixed by the Euler
#   step, G follows from g_q and the weights), g (21), c (14), g_q (7);
#   ~2e4 flops (dynamics, its gradient, FK)
# schur (split form, k_form_split): reads the lower rows of C of the knot
#   before (147), g_q of both knots (14), q, r of the knot before and q of
#   its own (35), c (14); writes L^ and W (392) and gamma^ (14); ~2.5e4
#   flops (theta, its whitening factor, the two block products of L^)
# dz (k_compute_dz_lr): reads the lower rows of C (147), g_q (7), g (21),
#   lambda of the knot and the next (28), writes dz (21); 1.1e3 flops
# merit: reads xu, dz, ref (21 + 21 + 6 floats) per knot for all 8 step
#   lengths; flops not counted (priced by bytes only)
# pcg: per solve L^ and W (392), gamma^ (14), lambda in and out (28) from
#   HBM; per iteration the PCG_BYTES_PER_KNOT_ITER / PCG_FLOPS_PER_KNOT_ITER
#   above, on chip (LDS)
# The flops of linearize, schur, dz and merit are the counted ones (fp32
# VALU instructions x 64 lanes) when the counter file has them.
PHASE_MODEL = {
    "linearize": (4 * (21 + 14 + 6 + 147 + 21 + 14 + 7), 2.0e4),
    "schur": (4 * (147 + 14 + 35 + 14 + 392 + 14), 2.5e4),
    "dz": (4 * (147 + 7 + 21 + 28 + 21), 1.1e3),
    "merit": (4 * (21 + 21 + 6), 0.0),
}
PCG_HBM_BYTES_PER_KNOT = 4 * (392 + 14 + 28)


def load_package():
    """The package (ctypes front end of libmpcpcg.so) from its directory."""
    import importlib.util
    name = "gpu_pcg_tracking_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def host_info():
    """Cores and CPU model of the host the CPU baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "model": model}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--sqp-iters", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline's parallel leg (0: every core this process may use)")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-closed-loop", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="issue every solve eagerly (no HIP graph replay), e.g. under counter profiling")
    ap.add_argument("--no-modes", action="store_true",
                    help="skip the stair-preconditioner and direct-solve legs (e.g. under tracing)")
    ap.add_argument("--no-stress", action="store_true",
                    help="skip the stress configuration (4096 instances, N in {128, 256, 512}, Q/R swept)")
    return ap.parse_args()


def build_workload(pkg, solver, torch, n, seed, dt):
    horizons = pkg.synth_horizons(seed, n).astype(np.int64)
    xus, qrefs, offs = [], [], []
    for i, N in enumerate(horizons):
        x0, qref = pkg.synth_instance(seed * 7919 + i, int(N), dt)
        xu = np.zeros(21 * N - 7, np.float32)
        for k in range(N):
            xu[21 * k:21 * k + 14] = x0
        xus.append(xu)
        qrefs.append(qref)
    q_all = torch.from_numpy(np.concatenate(qrefs)).cuda()
    ee = solver.fk(q_all).reshape(-1).contiguous()  # device FK of the reference path
    xu = torch.from_numpy(np.concatenate(xus)).cuda()
    return horizons, xu, ee


def pcg_accuracy(solver, torch, horizons, lam, tol, cap):
    """Accuracy of the last batched PCG solves of `solver` (their
    multipliers in lam): the per-instance exit state (mpc_batch_pcg_exit)
    and the relative error max|lam - lam_d| / max|lam_d| per instance
    against the same Schur systems solved directly (block LDL^T with two
    fp64-residual refinements, mpc_batch_direct_resolve; overwrites the
    preconditioner workspace)."""
    n = len(horizons)
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    fl = torch.zeros(n, dtype=torch.int32, device="cuda")
    eta = torch.zeros(n, device="cuda")
    solver.batch_pcg_exit(it, fl, eta)
    lam_d = torch.empty_like(lam)
    solver.batch_direct_resolve(lam_d, refine=2)
    it, fl, eta = it.cpu().numpy(), fl.cpu().numpy(), eta.cpu().numpy().astype(np.float64)
    a, b = lam.cpu().numpy().astype(np.float64), lam_d.cpu().numpy().astype(np.float64)
    k0 = np.concatenate([[0], np.cumsum(horizons)]) * 14
    err = np.array([np.abs(a[k0[i]:k0[i + 1]] - b[k0[i]:k0[i + 1]]).max() /
                    max(np.abs(b[k0[i]:k0[i + 1]]).max(), 1e-30) for i in range(n)])
    capped = fl == 1
    q = lambda v, p: round(float(np.quantile(v, p)), 8) if len(v) else None
    return {"pcg_exit_tol": tol, "pcg_max_iter": cap, "instances": n,
            "capped_frac": round(float(capped.mean()), 4), "faulted": int((fl == 2).sum()),
            "iters_mean": round(float(it.mean()), 1),
            "eta_at_exit": {"median": q(eta, 0.5), "max": round(float(eta.max()), 8),
                            "median_capped": q(eta[capped], 0.5), "median_converged": q(eta[~capped], 0.5)},
            "rel_err_vs_direct": {"median": q(err, 0.5), "p90": q(err, 0.9), "max": round(float(err.max()), 8),
                                  "median_capped": q(err[capped], 0.5), "median_converged": q(err[~capped], 0.5)},
            "note": "eta = |r^T Pinv r| of each instance's last exit test (the exit rule's quantity); the error "
                    "is of the PCG multipliers against a direct solve of the same Schur system"}


def config1(pkg, torch, seed, n=64, N=32, reps=3):
    """The plain single SQP step at N = 32 (one instance per launch
    sequence, one SQP iteration, PCG tol 1e-5, cap 173), over n seeded
    instances: whole-step wall time and the PCG kernel time (device events
    around the PCG launch, the reference's 'linsys' without its two small
    copies)."""
    s = pkg.Solver(pkg.default_config(max_knots=N, max_instances=1, max_total_knots=N, sqp_max_iter=1))
    step_us, lin_us, iters = [], [], []
    for i in range(n):
        x0, qref = pkg.synth_instance(seed * 7919 + 300000 + i, N)
        ee = s.fk(torch.from_numpy(qref).cuda()).reshape(-1).contiguous()
        xu0 = np.zeros(21 * N - 7, np.float32)
        for k in range(N):
            xu0[21 * k:21 * k + 14] = x0
        d_xu0 = torch.from_numpy(xu0).cuda()
        for r in range(reps + 1):
            xu, lam = d_xu0.clone(), torch.zeros(14 * N, device="cuda")
            torch.cuda.synchronize()
            t = time.perf_counter()
            out = s.sqp_solve(N, ee, lam, xu, 1e-3)
            torch.cuda.synchronize()
            if r > 0:  # the first solve of an instance is a warm-up
                step_us.append((time.perf_counter() - t) * 1e6)
                lin_us.append(out[1][0])
                iters.append(out[0][0])
    s.close()
    it = np.asarray(iters)
    return {"workload": f"plain single SQP step, N={N}, {n} seeded instances x {reps} solves, PCG tol 1e-5, cap 173",
            "sqp_step_us": {"mean": round(float(np.mean(step_us)), 2), "median": round(float(np.median(step_us)), 2)},
            "pcg_linsys_us": {"mean": round(float(np.mean(lin_us)), 2), "median": round(float(np.median(lin_us)), 2)},
            "pcg_iters": {"mean": round(float(it.mean()), 1), "max": int(it.max()),
                          "hit_cap_frac": round(float((it >= 173).mean()), 3)},
            "pcg_linsys_us_per_iteration": round(float(np.sum(lin_us) / max(int(it.sum()), 1)), 3),
            "reference_rtx4090": {"pcg_linsys_us_tol1e-5": 51.76, "pcg_linsys_us_173_iters": 742.23,
                                  "source": "old-test-results/32-knot-point-100-iters-5-configs-test.txt:86, :91-96"}}


def differentiation(pkg, torch, seed, reps=10, n=256, N=64, sqp_iters=3):
    """Adjoint / forward cost of the QP layer on n instances at horizon N.
    The PCG cap is 1000 (the layer needs converged solves; the SQP setting's
    173 cuts some of them short)."""
    dt = 0.01
    cfg = pkg.default_config(max_knots=N, max_instances=n, max_total_knots=n * N, pcg_max_iter=1000)
    solver = pkg.Solver(cfg)
    xus, qrefs = [], []
    for i in range(n):
        x0, qref = pkg.synth_instance(seed * 7919 + 100000 + i, N, dt)
        xu = np.zeros(21 * N - 7, np.float32)
        for k in range(N):
            xu[21 * k:21 * k + 14] = x0
        xus.append(xu)
        qrefs.append(qref)
    ee = solver.fk(torch.from_numpy(np.concatenate(qrefs)).cuda()).reshape(-1).contiguous()
    xu = torch.from_numpy(np.concatenate(xus)).cuda()
    solver.batch_setup([N] * n)
    T = n * N
    lam, rho = torch.zeros(14 * T, device="cuda"), torch.full((n,), 1e-3, device="cuda")
    status = torch.zeros(n, dtype=torch.int32, device="cuda")
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    solver.batch_sqp(xu, lam, ee, rho, sqp_iters, status)
    e[1].record()
    torch.cuda.synchronize()
    sqp_ms = e[0].elapsed_time(e[1])
    lam0 = lam.clone()
    dz, xstar = torch.empty_like(xu), torch.empty_like(xu)
    loss = torch.empty(n, device="cuda")
    dLdx, dLdref = torch.empty_like(xu), torch.empty_like(ee)
    gw = torch.empty(21 * n, device="cuda")
    gr = torch.empty_like(ee)

    def forward():
        lam.copy_(lam0)
        solver.qp_layer(xu, ee, rho, lam, dz, xstar)
        solver.qp_loss(xstar, ee, loss, dLdx, dLdref)

    def adjoint():
        gr.copy_(dLdref)
        solver.qp_adjoint(xu, ee, rho, dz, dLdx, gw, gr)

    for _ in range(2):
        forward()
        adjoint()
    torch.cuda.synchronize()
    e[0].record()
    for _ in range(reps):
        forward()
    e[1].record()
    for _ in range(reps):
        adjoint()
    e[2].record()
    torch.cuda.synchronize()
    fwd_ms, adj_ms = e[0].elapsed_time(e[1]) / reps, e[1].elapsed_time(e[2]) / reps
    finite = bool(torch.isfinite(gw).all().item() and torch.isfinite(gr).all().item())
    solver.close()
    return {"workload": f"{n} instances x N={N}, Q_ee=1, QD=1e-2, R=1e-4, {sqp_iters} SQP iterations, then the "
                        f"QP layer (PCG tol 1e-5, 2 refinement steps) and its adjoint; 21 weight + {3 * N} "
                        f"reference-point gradients per instance",
            "sqp_forward_ms": round(sqp_ms, 3), "layer_forward_ms": round(fwd_ms, 3),
            "adjoint_ms": round(adj_ms, 3), "adjoint_over_layer_forward": round(adj_ms / fwd_ms, 3),
            "adjoint_over_sqp_plus_layer": round(adj_ms / (sqp_ms + fwd_ms), 3),
            "gradients_per_sec": round(n / (adj_ms * 1e-3), 1), "gradients_finite": finite}


def closed_loop(pkg, torch, seed, N=128, steps=2000, timed_steps=200, K=3):
    """The closed-loop workload; see the module docstring."""
    solver = pkg.Solver(pkg.default_config(max_knots=N, max_instances=1, max_total_knots=N, pcg_exit_tol=5e-6))
    x0, ee, traj = solver.track_inputs(seed, steps, N)

    def run(n, graph, timing):
        xu = traj[:21 * N - 7].clone()
        lam = torch.zeros(14 * N, device="cuda")
        pcg = torch.zeros(n * K, dtype=torch.int32, device="cuda")
        cl = pkg.default_closed_loop(steps=n, sqp_iters=K, use_graph=graph, timing=timing)
        _, err, _, st = solver.closed_loop(N, x0, ee, xu, lam, cl, traj, pcg_iters=pcg)
        return err.cpu().numpy(), pcg.cpu().numpy(), st, lam

    def last_solve_accuracy(lam):
        # the last control step's last PCG solve against a direct solve of
        # its Schur system; the loop has since shifted lambda by one knot,
        # so its knots 1 .. N-1 are lam[0 .. N-2]
        it = torch.zeros(1, dtype=torch.int32, device="cuda")
        fl = torch.zeros(1, dtype=torch.int32, device="cuda")
        eta = torch.zeros(1, device="cuda")
        solver.batch_pcg_exit(it, fl, eta)
        lam_d = torch.empty_like(lam)
        solver.batch_direct_resolve(lam_d, refine=2)
        a = lam[:14 * (N - 1)].cpu().numpy().astype(np.float64)
        b = lam_d[14:].cpu().numpy().astype(np.float64)
        return (int(it.item()), int(fl.item()), float(eta.item()),
                float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)))

    run(20, 1, 0)  # warm-up (first launches, graph instantiation paths)
    samples = []
    for n_s in (50, 200, 800):  # accuracy of the last solve after n_s control steps
        samples.append(last_solve_accuracy(run(n_s, 1, 0)[3]))
    err, pcg, st, lam_end = run(steps, 1, 0)
    samples.append(last_solve_accuracy(lam_end))
    _, pcg_t, st_t, _ = run(timed_steps, 0, 1)
    solver.close()
    s_eta = [s[2] for s in samples]
    s_err = [s[3] for s in samples]
    accuracy = {"pcg_exit_tol": 5e-6, "pcg_max_iter": 173, "samples": "the last PCG solve after 50, 200, 800 and 2000 "
                                                                        "control steps",
                "capped_frac": round(float(np.mean([s[1] == 1 for s in samples])), 3),
                "eta_at_exit": {"median": round(float(np.median(s_eta)), 8), "max": round(float(max(s_eta)), 8)},
                "rel_err_vs_direct": {"median": round(float(np.median(s_err)), 8),
                                      "max": round(float(max(s_err)), 8)},
                "note": "knots 1..N-1 of the multipliers (the loop shifts them by one knot after each solve)"}
    n_it = timed_steps * K
    ref_traj = reference_trajectory(pkg, torch, N, K)
    return {"workload": f"1 instance, N={N}, {steps} control steps, {K} SQP iterations per step, PCG tol 5e-6, "
                        f"state noise 1e-3, plant 50 x 2e-4 s Euler substeps, precomputed-trajectory warm start",
            "us_per_control_step": round(st["total_us"] / steps, 1), "graph": bool(st["graph"]),
            "rms_ee_error_m": round(float(np.sqrt(np.mean(err[:, 0] ** 2))), 4),
            "mean_l1_ee_error_m": round(float(err[:, 1].mean()), 4),
            "pcg_iters_mean": round(float(pcg.mean()), 1), "pcg_hit_cap_frac": round(float((pcg >= 173).mean()), 3),
            "pcg_accuracy": accuracy,
            "timed_run": {"steps": timed_steps, "linsys_us_per_sqp_iter": round(st_t["linsys_us"] / n_it, 1),
                          "sqp_us_per_sqp_iter": round(st_t["solve_us"] / n_it, 1),
                          "plant_us_per_step": round(st_t["plant_us"] / timed_steps, 1)},
            "reference_pcg_linsys_us": {"N128_tol1e-5_rtx4090": 448.79},
            "reference_trajectory_0_0": ref_traj}


def reference_trajectory(pkg, torch, N, K):
    """The reference's own closed-loop input (testfiles/0_0_traj.csv and
    0_0_eepos.traj, kept in tests/golden/fk_0_0.npz) with its settings:
    dt = 1/64 s, QD = R = 1e-4, PCG tol 1e-5 with its cap 167 at N = 128, no
    noise; one control step per knot (78 plant substeps) over the whole
    trajectory; PCG and direct solver.  The reference reports a mean L1
    end-effector error of 0.0489 m (PCG, tol 1e-5) and 0.0318 m (QDLDL)."""
    import importlib
    fx = importlib.import_module(pkg.__name__ + ".fixtures")
    g = np.load(os.path.join(ROOT, "tests", "golden", "fk_0_0.npz"))
    xu_rows, ee_rows = g["xu"].astype(np.float32), g["ee"].astype(np.float32)
    dt = 1.0 / 64
    out = {"settings": f"N={N}, {K} SQP iterations per step, dt=1/64, QD=R=1e-4, PCG tol 1e-5 cap 167, no noise",
           "reference_mean_l1_m": {"pcg_tol1e-5": 0.0489, "qdldl": 0.0318}}
    for mode, flags in (("pcg", 0), ("direct", pkg.FLAG_DIRECT_SOLVE)):
        s = pkg.Solver(pkg.default_config(max_knots=N, max_instances=1, max_total_knots=N, dt=dt, qd_cost=1e-4,
                                          r_cost=1e-4, pcg_exit_tol=1e-5, pcg_max_iter=167, flags=flags))
        x0, path, traj, steps = fx.closed_loop_inputs(xu_rows, ee_rows, N)
        xu = traj[:21 * N - 7].clone()
        lam = torch.zeros(14 * N, device="cuda")
        cl = pkg.default_closed_loop(steps=steps, sqp_iters=K, substeps=78, sim_dt=dt / 78, noise_sigma=0.0)
        _, err, _, st = s.closed_loop(N, x0, path, xu, lam, cl, traj)
        e = err.cpu().numpy()
        s.close()
        out[mode] = {"steps": steps, "mean_l1_ee_error_m": round(float(e[:, 1].mean()), 4),
                     "rms_ee_error_m": round(float(np.sqrt(np.mean(e[:, 0] ** 2))), 4),
                     "us_per_control_step": round(st["total_us"] / steps, 1)}
    return out


def stress(pkg, torch, seed, n=4096, ratios=(1e2, 1e4, 1e6), K=3):
    """Stress configuration: n instances with horizons drawn from {128, 256,
    512}; per Q/R ratio (R = Q_ee / ratio, Q_ee = 1) one batched solve of K
    SQP iterations each with block-Jacobi PCG, stair-preconditioned PCG and
    the direct solver, then the QP layer and its adjoint at the final
    iterate."""
    horizons = pkg.synth_horizons(seed + 5, n, (128, 256, 512)).astype(np.int64)
    T = int(horizons.sum())
    cfg = pkg.default_config(max_knots=512, max_instances=n, max_total_knots=T, sqp_max_iter=K)
    solver = pkg.Solver(cfg)
    xus, qrefs = [], []
    for i, N in enumerate(horizons):
        x0, qref = pkg.synth_instance(seed * 7919 + 200000 + i, int(N), 0.01)
        xu = np.zeros(21 * N - 7, np.float32)  # x_k = x0, u_k = 0
        for k in range(N):
            xu[21 * k:21 * k + 14] = x0
        xus.append(xu)
        qrefs.append(qref)
    ee = solver.fk(torch.from_numpy(np.concatenate(qrefs)).cuda()).reshape(-1).contiguous()
    xu0 = torch.from_numpy(np.concatenate(xus)).cuda()
    del xus, qrefs
    solver.batch_setup(horizons.tolist())
    xu, lam = torch.empty_like(xu0), torch.zeros(14 * T, device="cuda")
    rho, status = torch.empty(n, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda")
    log = torch.zeros(n * K, dtype=torch.int32, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    out = {"workload": f"{n} instances, N ~ U{{128, 256, 512}} (total {T} knots), {K} SQP iterations, PCG tol 1e-5, "
                       f"max_iter 173; Q_ee = 1, QD = 1e-2, R = 1 / ratio",
           "pcg_layout": solver.batch_pcg_layout(), "ratios": {}}
    for ratio in ratios:
        r_cost = 1.0 / ratio
        solver.set_config(r_cost=r_cost, flags=0)  # also resets the weights to (0, QD) / R
        row = {}
        for mode, flags in (("pcg", 0), ("pcg_stair", pkg.FLAG_PCG_STAIR), ("direct", pkg.FLAG_DIRECT_SOLVE)):
            solver.set_config(r_cost=r_cost, flags=flags)
            xu.copy_(xu0)
            lam.zero_()
            rho.fill_(1e-3)
            log.fill_(-1)
            torch.cuda.synchronize()
            ev[0].record()
            solver.batch_sqp(xu, lam, ee, rho, K, status, log)
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1])
            lg, st = log.cpu().numpy().reshape(n, K), status.cpu().numpy()
            ex = lg >= 0
            row[mode] = {"ms": round(ms, 2), "kkt_solves_per_sec": round(int(ex.sum()) / (ms * 1e-3), 1),
                         "aborted": int((st & 4 != 0).sum()), "nonfinite": int((st & 8 != 0).sum())}
            if mode != "direct":
                its = lg[ex]
                row[mode]["pcg_iters"] = {"mean": round(float(its.mean()), 1), "max": int(its.max()),
                                          "hit_cap_frac": round(float((its >= 173).mean()), 3)}
                by_n = {}
                for N in (128, 256, 512):
                    m = (horizons == N)[:, None] & ex
                    by_n[str(N)] = round(float(lg[m].mean()), 1)
                row[mode]["pcg_iters_mean_by_N"] = by_n
                row[mode]["accuracy"] = pcg_accuracy(solver, torch, horizons, lam, 1e-5, 173)
        # forward (QP layer + loss) and adjoint at the final (direct) iterate, PCG mode
        solver.set_config(r_cost=r_cost, flags=0)
        dz, xs = torch.empty_like(xu), torch.empty_like(xu)
        loss = torch.empty(n, device="cuda")
        dLdx, dLdref = torch.empty_like(xu), torch.empty_like(ee)
        gw, gr = torch.empty(21 * n, device="cuda"), torch.empty_like(ee)
        lam.zero_()
        ev[0].record()
        solver.qp_layer(xu, ee, rho, lam, dz, xs)
        solver.qp_loss(xs, ee, loss, dLdx, dLdref)
        ev[1].record()
        gr.copy_(dLdref)
        solver.qp_adjoint(xu, ee, rho, dz, dLdx, gw, gr)
        ev[2].record()
        torch.cuda.synchronize()
        fwd, adj = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
        row["qp_layer"] = {"forward_ms": round(fwd, 2), "adjoint_ms": round(adj, 2),
                           "adjoint_over_forward": round(adj / fwd, 3),
                           "gradients_finite": bool(torch.isfinite(gw).all().item() and torch.isfinite(gr).all().item())}
        out["ratios"][f"{ratio:.0e}"] = row
    solver.close()
    return out


def cpu_instance(orc, pkg, seed, N):
    """A workload instance for the CPU legs: the library's seeded generator,
    the oracle's FK for the reference path, x_k = x0 and u_k = 0."""
    x0, qref = pkg.synth_instance(seed, N)
    ee = np.concatenate([orc.fk(q.astype(np.float64)) for q in qref])
    xu = np.zeros(21 * N - 7, np.float64)
    for k in range(N):
        xu[21 * k:21 * k + 14] = x0
    return ee, xu


def _rate(fn, budget):
    """Seconds per call of fn(reps) -> elapsed, doubling reps up to ~budget s."""
    reps, el = 1, 0.0
    while el < budget:
        el = fn(reps)
        if el < budget:
            reps = max(reps * 2, int(reps * budget / max(el, 1e-6)))
    return el / reps


def cpu_baseline(orc, pkg, torch, horizons, seconds, seed):
    """The reference's QDLDL linsys restated on the host (the oracle's fp32
    sparse LDL^T factor + solve of the Schur system, one core), one instance
    per horizon of the workload at its initial iterate, weighted by the
    horizon mix; also the host Schur formation (the oracle's fp32
    form_schur) and the reference's copies around its solve
    (toplevel.cuh:314-335: S values and gamma to the host, lambda back;
    pageable host buffers as there), timed on this GPU.  Returns the rates,
    their description and the per-horizon inputs."""
    cfg = orc.make_cfg()
    per_n, kkt = {}, {}
    for N in sorted(set(int(h) for h in horizons)):
        ee, xu = cpu_instance(orc, pkg, seed + N, N)
        G, Cm, g, c = orc.form_kkt(cfg, N, xu, ee, xu[:14], dtype=np.float32)
        S, P, gam, _ = orc.form_schur(N, G, Cm, g, c, 1e-3, dtype=np.float32)
        per_n[N] = (S, gam)
        kkt[N] = (G, Cm, g, c)
    budget = seconds / len(per_n)
    t_per, t_schur, t_copy = {}, {}, {}
    for N, (S, gam) in per_n.items():
        t_per[N] = _rate(lambda reps: orc.ldl_bench(N, S, gam, reps)[0], budget)
        G, Cm, g, c = kkt[N]

        def schur(reps):
            t = time.perf_counter()
            for _ in range(reps):
                orc.form_schur(N, G, Cm, g, c, 1e-3, dtype=np.float32)
            return time.perf_counter() - t
        t_schur[N] = _rate(schur, budget / 4)
        if torch is None or not torch.cuda.is_available():
            continue
        nnz = len(orc.csc_values(N, S, dtype=np.float32))
        d_val, d_gam, d_lam = (torch.zeros(m, device="cuda") for m in (nnz, 14 * N, 14 * N))
        h_val, h_gam, h_lam = (torch.zeros(m) for m in (nnz, 14 * N, 14 * N))

        def copies(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                h_val.copy_(d_val)
                h_gam.copy_(d_gam)
                d_lam.copy_(h_lam)
                torch.cuda.synchronize()
            return time.perf_counter() - t
        t_copy[N] = _rate(copies, budget / 8)
    mix = [int(h) for h in horizons]
    mean = lambda t: float(np.mean([t[N] for N in mix]))
    desc = "; ".join(f"N={N}: {t_per[N] * 1e6:.0f} us factor+solve, "
                     + (f"{t_copy[N] * 1e6:.0f} us copies, " if t_copy else "")
                     + f"{t_schur[N] * 1e6:.0f} us Schur formation" for N in sorted(t_per))
    rates = {"ldl": 1.0 / mean(t_per), "ldl_with_copies": 1.0 / (mean(t_per) + mean(t_copy)) if t_copy else None,
             "schur": 1.0 / mean(t_schur)}
    return rates, desc, t_per, per_n


def cpu_baseline_threads(orc, per_n, t_per, horizons, seconds, threads):
    """The same solves on `threads` host threads at once, one instance per
    thread at a time (the oracle's LDL^T keeps its workspace per call and
    ctypes releases the GIL): each thread runs the workload's horizon mix
    for about `seconds` at the one-core rates.  Returns KKT solves / s."""
    from concurrent.futures import ThreadPoolExecutor
    h = np.asarray(horizons)
    mix = {N: int((h == N).sum()) for N in per_n}
    t_pass = sum(mix[N] * t_per[N] for N in per_n)
    reps = {N: max(1, int(round(mix[N] * seconds / t_pass))) for N in per_n}

    def work(_):
        for N, (S, gam) in per_n.items():
            orc.ldl_bench(N, S, gam, reps[N])

    with ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(work, range(threads)))
        wall = time.perf_counter() - t0
    return threads * sum(reps.values()) / wall


def measure_bandwidth(pkg, torch):
    """HBM and LDS bandwidth of this device from the library's own probes
    (mpc_bw_copy: a 1 GiB float4 copy, 2 GiB moved per launch; mpc_bw_lds:
    4096 workgroups x 256 lanes x 2048 x 64 B of 128-bit LDS reads), timed
    with events on the launch stream, best of 5 launches after 2 warm-ups."""
    def best(fn):
        for _ in range(2):
            fn()
        ts = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return min(ts)

    n = 1 << 28
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    hbm = 8.0 * n / best(lambda: pkg.bw_copy(src, dst)) / 1e9
    del src, dst
    n_wg, reps = 4096, 2048
    out = torch.empty(256 * n_wg, dtype=torch.int32, device="cuda")
    lds = 64.0 * 256 * reps * n_wg / best(lambda: pkg.bw_lds(out, n_wg, reps)) / 1e9
    return {"hbm_gbs": round(hbm, 1), "lds_gbs": round(lds, 1),
            "method": "mpc_bw_copy (1 GiB float4 copy, read + write bytes) and mpc_bw_lds (128-bit LDS reads, "
                      "4096 x 256 lanes x 2048 x 64 B), best of 5 event-timed launches"}


def phase_roofline(phases, knots, pcg_units, bw, counters=None):
    """Per phase of one SQP iteration: roofline time = max(B_LDS / BW_LDS,
    B_HBM / BW_HBM, F / fp32 peak) over the measured bandwidths, and the
    fraction roofline time / measured time.  The flops per knot of the
    linearisation, the split formation, dz and the merits are counted from
    their kernels when the counter file has them (fp32 VALU instructions x
    64 lanes), else the estimate."""
    rows = {}
    model = dict(PHASE_MODEL)
    flop_source = {}
    for name in ("linearize", "schur", "dz", "merit"):
        per = (counters or {}).get("flops_per_knot", {}).get(name)
        if per:
            model[name] = (model[name][0], float(per))
            flop_source[name] = "counted"
        else:
            flop_source[name] = "not counted" if name == "merit" else "estimate"

    def row(name, us, b_hbm, b_lds, flops):
        t = {"hbm": b_hbm / (bw["hbm_gbs"] * 1e9), "lds": b_lds / (bw["lds_gbs"] * 1e9),
             "fp32": flops / (FP32_PEAK_TFLOPS * 1e12)}
        bound = max(t, key=t.get)
        roof = t[bound] * 1e6
        rows[name] = {"us": round(us, 2), "hbm_bytes": int(b_hbm), "lds_bytes": int(b_lds), "flops": int(flops),
                      "roofline_us": round(roof, 2), "bound": bound, "frac": round(roof / us, 4) if us > 0 else None}

    for name in ("linearize", "schur"):
        b, f = model[name]
        row(name, phases[name], b * knots, 0.0, f * knots)
    row("pcg", phases["pcg_team"] + phases["pcg_rest"], PCG_HBM_BYTES_PER_KNOT * knots,
        PCG_BYTES_PER_KNOT_ITER * pcg_units, PCG_FLOPS_PER_KNOT_ITER * pcg_units)
    for name in ("dz", "merit"):
        b, f = model[name]
        row(name, phases[name], b * knots, 0.0, f * knots)
    return {"bandwidth": bw, "fp32_peak_tflops": FP32_PEAK_TFLOPS, "knots_per_sqp_iter": int(knots),
            "pcg_knot_iterations_per_sqp_iter": int(pcg_units), "phases": rows, "flops_source": flop_source,
            "note": "per SQP iteration of the headline workload; bytes are the algorithmic ones per knot of the "
                    "split path as it runs (PCG: per knot and iteration the reference algorithm's S, Pinv and "
                    "vectors, priced at the LDS rate); flops: PCG from the operation count of its algorithm, the "
                    "other phases counted from their kernels' fp32 VALU instructions x 64 lanes (an upper bound: "
                    "masked lanes count); fp32 peak is the published packed rate (this build issues unpacked "
                    "fp32: half of it)"}


def other_modes(pkg, solver, torch, timed, step, K):
    """The same workload with the stair preconditioner (half the PCG
    iterations for one more block product each) and with the direct Schur
    solve (block LDL^T on the GPU, one fp64-residual refinement of each KKT
    step) instead of block-Jacobi PCG."""
    solver.set_config(flags=pkg.FLAG_PCG_STAIR)
    ms_stair, _, r = timed(True)
    lg_s, st_s = r[4], r[3]
    its_s = lg_s[lg_s >= 0]
    stair = {"kkt_solves_per_sec": round(int((lg_s >= 0).sum()) / (ms_stair * 1e-3), 1),
             "ms_per_step": round(ms_stair, 4),
             "pcg_iters": {"mean": round(float(its_s.mean()), 1), "max": int(its_s.max()),
                           "hit_max_iter": int((st_s & 2 != 0).sum())}}

    solver.set_config(flags=pkg.FLAG_DIRECT_SOLVE)
    ms_direct, _, r = timed(True)
    st_d = r[3]
    kkt_d = int((r[4] >= 0).sum())
    solver.batch_set_timing(True)
    step()
    torch.cuda.synchronize()
    ph_d = solver.batch_phase_times()
    solver.batch_set_timing(False)
    direct = {"kkt_solves_per_sec": round(kkt_d / (ms_direct * 1e-3), 1), "ms_per_step": round(ms_direct, 4),
              "phase_us_per_sqp_iter": {name: round(v / K, 2) for name, v in zip(PHASES, ph_d)},
              "status": {"aborted": int((st_d & 4 != 0).sum()), "nonfinite": int((st_d & 8 != 0).sum())},
              "note": "phase 'pcg_team' holds the direct solve, 'dz' the dz recovery and its KKT refinement"}
    solver.set_config(flags=0)
    return stair, direct


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the instances are independent and would be split across "
                                   "GPUs by instance, which is not built here; only --gpus 1 is measured"}))
        sys.exit(2)
    import torch
    pkg = load_package()

    dt = 0.01
    n = args.instances
    cfg = pkg.default_config(max_knots=256, max_instances=n, max_total_knots=256 * n, sqp_max_iter=args.sqp_iters)
    solver = pkg.Solver(cfg)
    horizons, xu0, ee = build_workload(pkg, solver, torch, n, args.seed, dt)
    T = int(horizons.sum())
    solver.batch_setup(horizons.tolist())
    # launch plan of the batched PCG: every packing that differs for this
    # layout timed on the device, the fastest kept (mpc_batch_autotune)
    t_tune = time.perf_counter()
    tune = solver.batch_autotune(iters=16, reps=3)
    t_tune = time.perf_counter() - t_tune
    xu = torch.empty_like(xu0)
    lam = torch.zeros(14 * T, device="cuda")
    rho = torch.empty(n, device="cuda")
    status = torch.zeros(n, dtype=torch.int32, device="cuda")
    pcg_log = torch.zeros(n * args.sqp_iters, dtype=torch.int32, device="cuda")
    K = args.sqp_iters

    def step():
        xu.copy_(xu0)
        lam.zero_()
        rho.fill_(1e-3)
        pcg_log.fill_(-1)
        solver.batch_sqp(xu, lam, ee, rho, K, status, pcg_log)

    def timed(graph):
        # graph: the batched solve is captured once (in the first warm-up
        # step) and replayed as a HIP graph
        graph = graph and not args.no_graph
        solver.batch_set_graph(graph)
        for _ in range(max(args.warmup, 1 if graph else 0)):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record()
        for _ in range(args.steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        t = time.perf_counter() - t
        solver.batch_set_graph(False)
        res = [a.cpu().numpy().copy() for a in (xu, lam, rho, status, pcg_log)]
        return e0.elapsed_time(e1) / args.steps, t, res

    # headline: the solve as a captured HIP graph; beside it the same launch
    # sequence issued eagerly (cooperative team launches)
    ms_step, t_wall, res_graph = timed(True)
    ms_eager, _, res_eager = timed(False)
    graph_equal = all(np.array_equal(a, b) for a, b in zip(res_graph, res_eager))
    log = res_graph[4].reshape(n, K)
    st = res_graph[3]
    executed = log >= 0
    kkt_per_step = int(executed.sum())
    value = kkt_per_step / (ms_step * 1e-3)

    # per-phase split and roofline of the dominant kernel, the PCG pool (the
    # PCG of the horizons above 128 on teams of workgroups and of the
    # resident bins, one persistent launch), timed with HIP events on the
    # launch stream
    solver.batch_set_timing(True)
    step()
    torch.cuda.synchronize()
    ph = solver.batch_phase_times()
    solver.batch_set_timing(False)
    log_t = pcg_log.cpu().numpy().reshape(n, K)
    # the PCG phase before its first non-cooperative launch is `launches`
    # cooperative launches (one workgroup per CU: the pool of teams and
    # bins, or with the static plan the team launches); per launch: the
    # knot-iterations of every instance those launches solve and the phase
    # time, divided by that count
    layout = solver.batch_pcg_layout()
    plan = solver.batch_get_plan()
    # block Jacobi with every N <= 256: one split-form launch solves every
    # instance (its phase ends at the same event)
    split_on = bool(plan["split"]) and int(horizons.max()) <= 256
    if split_on:
        kernel, launches = "k_pcg_split", 1
        in_team_phase = np.ones(n, bool)
    else:
        kernel = "k_pcg_pool" if plan["mixed"] else "k_pcg_team"
        launches = max(layout["team_launches"], 1)
        in_team_phase = solver.batch_pcg_launch_map() < layout["team_launches"]
    units = float(sum(int(horizons[i]) * int(log_t[i, j]) for i in range(n) for j in range(K)
                      if in_team_phase[i] and log_t[i, j] >= 0)) / K
    t_phase = ph[2] / K
    t_launch = t_phase / launches
    units_launch = units / launches
    achieved = PCG_BYTES_PER_KNOT_ITER * units / (t_phase * 1e-6) / 1e9
    tflops = PCG_FLOPS_PER_KNOT_ITER * units / (t_phase * 1e-6) / 1e12
    # the PMC figures of the same kernel (profiles/), if they are of it
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")
    counters = None
    if os.path.exists(COUNTERS_FILE):
        with open(COUNTERS_FILE) as f:
            counters = json.load(f)
    # accuracy of the last SQP iteration's PCG solves (this timing step's)
    accuracy = pcg_accuracy(solver, torch, horizons, lam, cfg.pcg_exit_tol, cfg.pcg_max_iter)
    phases = {name: round(v / K, 2) for name, v in zip(PHASES, ph)}
    kkt_phase_us = (ph[1] + ph[2] + ph[3] + ph[4]) / K
    # per-phase roofline against this device's measured HBM and LDS bandwidth
    ran = log_t >= 0
    knots_it = float((horizons[:, None] * ran).sum()) / K
    pcg_it = float((horizons[:, None] * np.where(ran, log_t, 0)).sum()) / K
    proof = phase_roofline(phases, knots_it, pcg_it, measure_bandwidth(pkg, torch), counters)
    lds_peak = proof["bandwidth"]["lds_gbs"]
    lds_counted = None
    valu = None
    if counters and counters.get("pcg_kernel", {}).get("kernel") == kernel:
        tp = counters["pcg_kernel"]
        # VALU issue: a wave64 instruction occupies a 16-lane SIMD for 4 cycles
        simds = 4 * torch.cuda.get_device_properties(0).multi_processor_count
        peak_rate = simds * ENGINE_CLOCK_HZ / 4.0
        rate = tp["valu_wave_instructions_per_launch"] / (t_launch * 1e-6)
        valu = {"achieved": round(rate / 1e9, 2), "peak": round(peak_rate / 1e9, 2), "unit": "G wave-instructions/s",
                "frac": round(rate / peak_rate, 4),
                "source": "SQ_INSTS_VALU per launch (PMC) from " + os.path.relpath(COUNTERS_FILE, ROOT) +
                          f"; peak = {simds} SIMDs x {ENGINE_CLOCK_HZ / 1e9:.1f} GHz / 4 cycles per wave64 instruction"}
        b_launch = tp["lds_wave_instructions_per_launch"] * tp["lds_bytes_per_wave_instruction"]
        lds_counted = {"bytes_per_launch": int(b_launch), "achieved": round(b_launch / (t_launch * 1e-6) / 1e9, 1),
                       "peak": lds_peak, "unit": "GB/s", "frac": round(b_launch / (t_launch * 1e-6) / 1e9 / lds_peak, 4),
                       "source": "SQ_INSTS_LDS per launch (PMC) x LDS bytes per wave instruction of the loop body "
                                 "(the kernel's ISA), from " + os.path.relpath(COUNTERS_FILE, ROOT)}

    # plain: one instance per launch sequence, host decisions (sqpSolve pattern)
    rng = np.random.default_rng(args.seed)
    sample = rng.choice(n, size=min(args.plain_sample, n), replace=False)
    offs_xu = np.concatenate([[0], np.cumsum(21 * horizons - 7)])
    offs_k = np.concatenate([[0], np.cumsum(horizons)])
    plain_iters, plain_s = 0, 0.0
    for i in sample:
        N = int(horizons[i])
        pxu = xu0[offs_xu[i]:offs_xu[i + 1]].clone()
        pee = ee[6 * offs_k[i]:6 * offs_k[i + 1]].clone()
        plam = torch.zeros(14 * N, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = solver.sqp_solve(N, pee, plam, pxu, 1e-3)
        torch.cuda.synchronize()
        plain_s += time.perf_counter() - t
        plain_iters += out[3]
    plain_value = plain_iters / plain_s

    cpu = None
    if not args.no_cpu:
        import oracle as orc
        host = host_info()
        rates, desc, t_per, per_n = cpu_baseline(orc, pkg, torch, horizons, args.cpu_seconds, args.seed)
        threads = args.cpu_threads if args.cpu_threads > 0 else host["usable_cores"]
        cval = rates["ldl"]
        tval = cpu_baseline_threads(orc, per_n, t_per, horizons, args.cpu_seconds, threads) if threads > 1 else cval
        cpu = {"value": round(tval, 1), "unit": "KKT solves/s", "cores": threads, "kind": "port",
               "host": host,
               "single_core_value": round(cval, 1),
               "single_core_with_copies_value": round(rates["ldl_with_copies"], 1)
               if rates["ldl_with_copies"] else None,
               "single_core_schur_formation_per_sec": round(rates["schur"], 1),
               "sample": f"fp32 sparse LDL^T factor+solve of the Schur system (the oracle's restatement of the "
                         f"reference's QDLDL path, its CSC pattern), one instance per horizon at its initial "
                         f"iterate, weighted by the workload's horizon mix; one core: {desc}, ~{args.cpu_seconds:.0f} "
                         f"s; then {threads} threads (every core this process may use), one instance per thread "
                         f"at a time, ~{args.cpu_seconds:.0f} s each.  'with copies' adds the reference's "
                         f"device-host copies around its solve (S values and gamma down, lambda up, pageable, "
                         f"timed on this GPU); the Schur formation line is the oracle's fp32 form_schur"}

    stair = direct = None
    if not args.no_modes:
        stair, direct = other_modes(pkg, solver, torch, timed, step, K)

    c1 = config1(pkg, torch, args.seed)
    diff = None if args.no_diff else differentiation(pkg, torch, args.seed)
    closed = None if args.no_closed_loop else closed_loop(pkg, torch, args.seed)
    stress_out = None if args.no_stress else stress(pkg, torch, args.seed)

    its = log[executed]
    out = {
        "metric": "batched KKT solves/sec & SQP iters/sec vs plain at tol 1e-5; adjoint/fwd cost",
        "value": round(value, 1),
        "unit": "KKT solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 start states and sinusoid EE references)",
        "config": {"workload": "batched SQP: 1024 KUKA iiwa EE-tracking instances, N ~ U{32,64,128,256}, "
                               "3 SQP iterations, block-Jacobi PCG tol 1e-5"
                               + ("" if args.no_graph else ", replayed as a captured HIP graph"),
                   "instances": n, "total_knots": T, "sqp_iters": K, "pcg_exit_tol": cfg.pcg_exit_tol,
                   "pcg_max_iter": cfg.pcg_max_iter},
        "sqp_iters_per_sec": round(value, 1),
        "kkt_phase_solves_per_sec": round(kkt_per_step / (kkt_phase_us * 1e-6), 1),
        "plain_kkt_solves_per_sec": round(plain_value, 1),
        "speedup_vs_plain": round(value / plain_value, 2),
        "phase_us_per_sqp_iter": phases,
        "pcg_accuracy": accuracy,
        "pcg_iters": {"mean": round(float(its.mean()), 1), "max": int(its.max()), "min": int(its.min()),
                      "hit_max_iter": int((st & 2 != 0).sum())},
        "status": {"aborted": int((st & 4 != 0).sum()), "nonfinite": int((st & 8 != 0).sum()),
                   "pcg_fault": int((st & 16 != 0).sum())},
        "autotune": {"chosen": solver.batch_get_plan(), "iters": 16, "reps": 3, "seconds": round(t_tune, 3),
                     "candidates_us": sorted(tune, key=lambda r: r["us"])},
        "graph": {"headline_is_graph_replay": not args.no_graph, "bitwise_equal_to_eager": bool(graph_equal),
                  "eager_ms_per_step": round(ms_eager, 4),
                  "eager_kkt_solves_per_sec": round(kkt_per_step / (ms_eager * 1e-3), 1),
                  "note": "eager = the same launch sequence without capture, PCG teams as cooperative launches; "
                          "in the graph they are ordinary kernel nodes (bounded waits, fault status)"},
        "roofline": {"kernel": kernel,
                     "bound": "lds", "achieved": round(achieved, 1),
                     "peak": lds_peak, "unit": "GB/s", "frac": round(achieved / lds_peak, 4),
                     "peak_source": "LDS bandwidth measured on this device in this run (mpc_bw_lds, 128-bit reads)",
                     "traffic": traffic, "algorithmic_bytes_per_launch": int(PCG_BYTES_PER_KNOT_ITER * units_launch),
                     "knot_iterations_per_launch": int(units_launch), "avg_launch_us": round(t_launch, 2),
                     "launches_per_sqp_iter": launches, "pcg_layout": layout,
                     "note": "algorithmic bytes = 2296 B per knot and PCG iteration (S, Pinv and vectors) x the "
                             "knot-iterations of one launch; the system stays on chip (LDS and registers), so the "
                             "roof is the LDS rate; traffic = HBM bytes per launch (PMC), the one-time load; "
                             "k_pcg_split: block-Jacobi PCG as CG on W^T S W (one 14x14 block per knot on chip, "
                             "two block products per iteration: fewer bytes and flops than the algorithmic count), "
                             "one launch over bins of 512-lane workgroups; k_pcg_pool / k_pcg_team: the reference "
                             "form (S, Pinv on chip) on teams and bins",
                     "hbm": None if traffic is None else
                     {"achieved": round(traffic / (t_launch * 1e-6) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(traffic / (t_launch * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
                     "lds_counted": lds_counted,
                     "valu_issue": valu,
                     "fp32": {"achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": round(tflops / FP32_PEAK_TFLOPS, 4)}},
        "phase_roofline": proof,
        "cpu_baseline": cpu,
        "plain_single_step_n32": c1,
        "pcg_stair": stair,
        "direct_solver": direct,
        "differentiation": diff,
        "closed_loop": closed,
        "stress": stress_out,
        "wall_s_timed": round(t_wall, 3),
    }
    print(json.dumps(out))
    solver.close()


if __name__ == "__main__":
    main()
