"""Shared helpers of the test suite: package loading, the synthetic
workload specification (a Python statement of the generator the library
implements in C++) and small layout utilities."""
import importlib.util
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "gpu-pcg-tracking_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_pkg():
    name = "gpu_pcg_tracking_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


# -- workload specification ------------------------------------------------------------------
MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53

    def normal(self):
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def synth_instance_spec(seed, N, dt=0.01):
    r = SplitMix64(seed)
    q0 = np.array([-0.5 + r.uniform() for _ in range(7)])
    qd0 = np.array([0.1 * r.normal() for _ in range(7)])
    f = np.array([0.2 + 0.8 * r.uniform() for _ in range(7)])
    ph = np.array([2 * math.pi * r.uniform() for _ in range(7)])
    k = np.arange(N)[:, None]
    qref = q0[None, :] + 0.3 * np.sin(2 * math.pi * f[None, :] * k * dt + ph[None, :])
    return np.r_[q0, qd0], qref


def synth_horizons_spec(seed, n, choices=(32, 64, 128, 256)):
    r = SplitMix64(seed)
    return np.array([choices[min(int(r.uniform() * len(choices)), len(choices) - 1)] for _ in range(n)])


def make_instance(orc, seed, N, dt=0.01, dtype=np.float32):
    """Oracle-side instance: x0, ee_ref (6N) via the fp64 oracle FK, xu with
    x_k = x0 and u_k = 0 (all rounded to fp32 values)."""
    x0, qref = synth_instance_spec(seed, N, dt)
    x0 = x0.astype(np.float32).astype(np.float64)
    qref = qref.astype(np.float32).astype(np.float64)
    ee = np.concatenate([orc.fk(qref[k]) for k in range(N)]).astype(np.float32)
    xu = np.zeros(21 * N - 7, np.float32)
    for k in range(N):
        xu[21 * k:21 * k + 14] = x0
    return x0.astype(np.float32), ee, xu


def perturbed_instance(orc, seed, N, dt=0.01, scale=0.05):
    """Like make_instance but with a non-trivial trajectory (perturbed states
    and controls) so every KKT term is exercised."""
    x0, ee, xu = make_instance(orc, seed, N, dt)
    rng = np.random.default_rng(seed)
    xu = xu.astype(np.float64)
    for k in range(1, N):
        xu[21 * k:21 * k + 14] += rng.normal(0, scale, 14)
    for k in range(N - 1):
        xu[21 * k + 14:21 * k + 21] = rng.normal(0, 2.0, 7)
    return x0, ee, xu.astype(np.float32)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


# -- fp32 rounding floors ---------------------------------------------------------------------
def ulp_perturbed(a, seed):
    """a with every element scaled by 1 +- 2^-24 (random signs): the same
    values to fp32 precision, a different rounding history downstream."""
    rng = np.random.default_rng(seed)
    a = np.asarray(a, np.float64)
    return a * (1.0 + 2.0 ** -24 * rng.choice([-1.0, 1.0], a.shape))


def split_system(N, S, gamma, dtype=np.float64):
    """The split-preconditioned form of block-Jacobi PCG (k_pcg_split.hip):
    with theta_b = -S(b,b) = U_b^T U_b (Cholesky) and W_b = U_b^-1,
    L^_b = W_b^T S(b,b-1) W_{b-1} and gamma^_b = W_b^T gamma_b.  Returns
    L^, W, U, gamma^ (blocks row-major), computed in dtype."""
    f = np.dtype(dtype).type
    B = np.asarray(S, np.float64).reshape(N, 3, 14, 14).astype(f)  # column-major blocks: (r, c) at r + 14 c
    L, D = np.transpose(B[:, 0], (0, 2, 1)), np.transpose(B[:, 1], (0, 2, 1))
    U = np.transpose(np.linalg.cholesky(-D), (0, 2, 1)).astype(f)  # theta = U^T U, U upper
    W = np.linalg.inv(U).astype(f)
    Lh = np.zeros_like(L)
    for b in range(1, N):
        Lh[b] = (W[b].T @ L[b] @ W[b - 1]).astype(f)
    gh = np.einsum("bkc,bk->bc", W, np.asarray(gamma).reshape(N, 14).astype(f)).astype(f)
    return Lh, W, U, gh


def split_cg(N, S, gamma, lam0, cap, tol, dtype=np.float64):
    """Block-Jacobi PCG on S lambda = gamma as CG on W^T S W lambda^ =
    W^T gamma (lambda = W lambda^), in dtype: (lambda, iterations).  The
    diagonal blocks of W^T S W are -I, the preconditioner is -I and
    r^T Pinv r = r^ . (-r^) (pinned against the oracle's PCG in
    test_split_form.py)."""
    f = np.dtype(dtype).type
    Lh, W, U, gh = split_system(N, S, gamma, dtype)

    def mv(x):
        y = -x.copy()
        y[1:] += np.einsum("brc,bc->br", Lh[1:], x[:-1]).astype(f)
        y[:-1] += np.einsum("bcr,bc->br", Lh[1:], x[1:]).astype(f)
        return y.astype(f)

    lh = np.einsum("brc,bc->br", U, np.asarray(lam0).reshape(N, 14).astype(f)).astype(f)  # lambda^ = U lambda
    r = (gh - mv(lh)).astype(f)
    p = -r
    eta = f(-np.sum(r * r, dtype=f))
    it = 0
    if abs(eta) >= tol:
        while it < cap:
            y = mv(p)
            a = f(eta / f(np.sum(p * y, dtype=f)))
            lh = (lh + a * p).astype(f)
            r = (r - a * y).astype(f)
            en = f(-np.sum(r * r, dtype=f))
            it += 1
            if abs(en) < tol:
                break
            p = (-r + f(en / eta) * p).astype(f)
            eta = en
    return np.einsum("brc,bc->br", W, lh).ravel().astype(np.float64), it


def schur_pcg_fp32_floor(orc, N, G, Cm, g, c, rho, lam0, cap, tol, ref, samples=7, split=False):
    """The fp32 floor of a Schur formation and PCG solve from the KKT blocks
    (G, Cm, g, c): the largest relative distance from the fp64 result ref of
    the oracle's fp32 pipeline (form_schur, then PCG), run on G as given and
    on `samples` one-ulp perturbations of G.  It covers the rounding of the
    formation as well: the block inverses of the regularised cost Hessian,
    whose condition number grows as 1 / rho, and the products with them.
    A single fp32 run is one draw of this floor; after a capped solve (173
    iterations) the draws are heavy-tailed (measured on a capped N = 64
    solve: 16 draws from 1e-4 to 2.5e-3, median 4.6e-4), so a bound built
    on one draw rejects correct fp32 implementations at random.  split=True
    adds draws of the split form (split_cg in fp32 on the same fp32 S): the
    same iteration in exact arithmetic, with its own rounding, which on
    ill-conditioned capped solves (rho = 1e-3) scatters several times wider
    than the reference form's."""
    worst = 0.0
    for k in range(samples + 1):
        Gk = G if k == 0 else ulp_perturbed(G, k)
        S, P, gam, _ = orc.form_schur(N, Gk, Cm, g, c, rho, dtype=np.float32)
        lt, _, _ = orc.pcg(N, S, P, gam, lam0, cap, tol, dtype=np.float32)
        worst = max(worst, rel_err(lt, ref))
        if split:
            ls, _ = split_cg(N, S, gam, lam0, cap, tol, dtype=np.float32)
            worst = max(worst, rel_err(ls, ref))
    return worst


def pcg_exit_ok(it_gpu, etas64, tol, cap, band=0.25):
    """Whether an fp32 PCG may exit after it_gpu iterations, given the fp64
    oracle's |r^T z| after each iteration (etas64: before the first and
    after each of `cap` iterations, run with tol = 0).  Near the exit the
    fp32 runs' |r^T z| sit up to ~15 % from the fp64 values (measured on
    this workload for the oracle's own fp32 PCG), and |r^T z| is not
    monotone in CG, so a crossing within that band can move by several
    iterations.  Accepted: an exit no later than the first iteration where
    the fp64 value is clearly below tol (< (1 - band) tol), at an iteration
    where it is below (1 + band) tol, or at the cap when there is no clear
    crossing before it."""
    e = np.abs(np.asarray(etas64, np.float64))
    assert len(e) == cap + 1
    below = np.flatnonzero(e < (1.0 - band) * tol)
    clear = int(below[0]) if len(below) else cap
    if it_gpu > clear:
        return False
    return it_gpu == cap or e[it_gpu] < (1.0 + band) * tol
