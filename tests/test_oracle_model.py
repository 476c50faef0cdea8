"""Pins the oracle's kinematics and dynamics (CPU only).

- FK against the reference's own golden file (testfiles/0_0_eepos.traj,
  6 significant digits) and fp64 central differences for the gradient.
- Dynamics against the reference's spatial inertia table (fixture written
  from rbdfiles/iiwa_eepos_grid.cuh) through an independent numpy
  kinetic-energy construction of M, the Lagrangian identity for the bias
  forces, FD(ID(qdd)) = qdd, and central differences of FD.
- A planted error (a perturbed link inertia) is caught by these checks.
"""
import numpy as np
import pytest

import helpers

GOLD = f"{helpers.GOLDEN}/fk_0_0.npz"


def test_fk_matches_reference_golden(orc):
    g = np.load(GOLD)
    err = max(np.abs(orc.fk(q)[:3] - e[:3]).max() for q, e in zip(g["q"], g["ee"]))
    assert err < 1e-5  # 6 significant digits in the golden file
    # rpy agrees away from the +-pi wrap
    errs = []
    for q, e in zip(g["q"], g["ee"]):
        d = (orc.fk(q)[3:] - e[3:] + np.pi) % (2 * np.pi) - np.pi
        errs.append(np.abs(d).max())
    assert max(errs) < 1e-4


def test_fk_grad_central_differences(orc):
    rng = np.random.default_rng(0)
    for _ in range(5):
        q = rng.uniform(-2, 2, 7)
        J = orc.fk_grad(q)
        h = 1e-6
        Jfd = np.stack([(orc.fk(q + h * e)[:3] - orc.fk(q - h * e)[:3]) / (2 * h) for e in np.eye(7)], 1)
        assert np.abs(J - Jfd).max() < 1e-8


# --- independent numpy kinematics for the dynamics checks -----------------------------------
RA = np.array([[-1, 0, 0], [0, 0, 1], [0, 1, 0.]])
RB = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0.]])
RF = [np.eye(3), RA, RA, RB, RA, RB, RA]
PF = [(0, 0, 0.1575), (0, 0, 0.2025), (0, 0.2045, 0), (0, 0, 0.2155), (0, 0.1845, 0), (0, 0, 0.2155),
      (0, 0.081, 0)]


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def link_frames(q):
    """World rotation and origin of each link frame (after joint i)."""
    R, p, out = np.eye(3), np.zeros(3), []
    for i in range(7):
        c, s = np.cos(q[i]), np.sin(q[i])
        Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
        p = p + R @ np.array(PF[i])
        R = R @ RF[i] @ Rz
        out.append((R.copy(), p.copy()))
    return out


def mass_matrix_ke(q, I6):
    """M = sum_i J_i^T I_i J_i with J_i the body Jacobian of link i in link
    coordinates ([w; v] of the link origin), from the numpy frames."""
    fr = link_frames(q)
    M = np.zeros((7, 7))
    for i in range(7):
        Ri, pi = fr[i]
        J = np.zeros((6, 7))
        for j in range(i + 1):
            Rj, pj = fr[j]
            z = Rj[:, 2]
            w = Ri.T @ z
            v = Ri.T @ np.cross(z, pi - pj)
            J[:, j] = np.r_[w, v]
        M += J.T @ I6[i] @ J
    return M


def test_fk_numpy_frames_agree_with_golden():
    g = np.load(GOLD)
    for q, e in list(zip(g["q"], g["ee"]))[::37]:
        assert np.abs(link_frames(q)[-1][1] - e[:3]).max() < 1e-5


def test_mass_matrix_matches_reference_inertia_table(orc):
    I6 = np.load(f"{helpers.GOLDEN}/iiwa_spatial_inertia.npy")
    rng = np.random.default_rng(1)
    for _ in range(6):
        q = rng.uniform(-2.5, 2.5, 7)
        M = orc.mass_matrix(q)
        assert np.abs(M - M.T).max() < 1e-12
        assert np.linalg.eigvalsh(M).min() > 0
        assert np.abs(M - mass_matrix_ke(q, I6)).max() < 1e-10 * np.abs(M).max()


def test_planted_inertia_error_is_caught(orc):
    I6 = np.load(f"{helpers.GOLDEN}/iiwa_spatial_inertia.npy").copy()
    I6[3, 3, 3] *= 1.01  # 1 % mass error on link 4
    q = np.array([0.3, -0.7, 0.2, 1.1, -0.4, 0.9, 0.1])
    assert np.abs(orc.mass_matrix(q) - mass_matrix_ke(q, I6)).max() > 1e-6


def test_bias_forces_lagrangian_identity(orc):
    """c(q, qd) = Mdot qd - 1/2 d/dq (qd^T M qd) (gravity 0)."""
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(4):
        q, qd = rng.uniform(-2, 2, 7), rng.normal(0, 1.5, 7)
        c = orc.rnea(q, qd, np.zeros(7))
        dM = [(orc.mass_matrix(q + h * e) - orc.mass_matrix(q - h * e)) / (2 * h) for e in np.eye(7)]
        Mdot = sum(dM[j] * qd[j] for j in range(7))
        dke = np.array([qd @ dM[j] @ qd for j in range(7)])
        assert np.abs(c - (Mdot @ qd - 0.5 * dke)).max() < 1e-7 * max(1, np.abs(c).max())


def test_inverse_forward_dynamics_roundtrip(orc):
    rng = np.random.default_rng(3)
    for _ in range(5):
        q, qd, u = rng.uniform(-2, 2, 7), rng.normal(0, 1, 7), rng.normal(0, 10, 7)
        qdd = orc.fd(q, qd, u)
        assert np.abs(orc.rnea(q, qd, qdd) - u).max() < 1e-10
        M = orc.mass_matrix(q)
        assert np.abs(orc.minv(q) @ M - np.eye(7)).max() < 1e-9


def test_fd_gradient_central_differences(orc):
    rng = np.random.default_rng(4)
    h = 1e-6
    for _ in range(4):
        x = np.r_[rng.uniform(-2, 2, 7), rng.normal(0, 1, 7), rng.normal(0, 10, 7)]
        qdd, df = orc.fd_grad(x[:7], x[7:14], x[14:])
        f = lambda z: orc.fd(z[:7], z[7:14], z[14:])
        dfd = np.stack([(f(x + h * e) - f(x - h * e)) / (2 * h) for e in np.eye(21)], 1)
        assert np.abs(df - dfd).max() < 1e-6 * max(1, np.abs(dfd).max())


def test_rnea_gradient_central_differences(orc):
    rng = np.random.default_rng(5)
    h = 1e-6
    q, qd, qdd = rng.uniform(-2, 2, 7), rng.normal(0, 1, 7), rng.normal(0, 3, 7)
    d = orc.rnea_grad(q, qd, qdd)
    x = np.r_[q, qd]
    f = lambda z: orc.rnea(z[:7], z[7:], qdd)
    dfd = np.stack([(f(x + h * e) - f(x - h * e)) / (2 * h) for e in np.eye(14)], 1)
    assert np.abs(d - dfd).max() < 1e-6 * max(1, np.abs(dfd).max())


@pytest.mark.parametrize("fn", ["fk", "fd"])
def test_fp32_twin_agrees(orc, fn):
    rng = np.random.default_rng(6)
    q, qd, u = rng.uniform(-2, 2, 7), rng.normal(0, 1, 7), rng.normal(0, 10, 7)
    if fn == "fk":
        a, b = orc.fk(q), orc.fk(q, dtype=np.float32)
    else:
        a, b = orc.fd(q, qd, u), orc.fd(q, qd, u, dtype=np.float32)
    assert np.abs(a - b).max() < 1e-4 * max(1, np.abs(a).max())
