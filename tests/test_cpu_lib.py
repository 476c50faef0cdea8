"""CPU checks of the C-ABI library: it loads without a GPU, exports every
entry point the header declares, validates arguments, and its host-side
workload generator follows the written specification."""
import os
import re

import numpy as np
import pytest

import helpers

HEADER = os.path.join(helpers.ROOT, "include", "mpcpcg.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpc_\w+)\s*\(", src)))


def test_header_declares_the_expected_surface(pkg):
    assert declared_functions() == sorted(pkg.EXPORTS)


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_default_config_matches_workload(pkg):
    c = pkg.default_config()
    assert (c.dt, c.q_ee, c.mu, c.rho_factor, c.rho_max, c.pcg_max_iter, c.sqp_max_iter) == (
        pytest.approx(0.01), 1.0, 10.0, pytest.approx(1.2), 10.0, 173, 3)
    assert c.qd_cost == pytest.approx(1e-2) and c.r_cost == pytest.approx(1e-4)
    assert c.rho_min == pytest.approx(1e-3) and c.pcg_exit_tol == pytest.approx(1e-5)


def test_create_rejects_bad_arguments_before_touching_a_device(pkg):
    import ctypes as C
    lib = pkg.lib()
    h = C.c_void_p()
    for kw in ({"max_knots": 1}, {"max_knots": 600}, {"max_instances": 0}, {"dt": -1.0},
               {"rho_factor": 1.0}, {"sqp_max_iter": 1000}):
        cfg = pkg.default_config(**kw)
        assert lib.mpc_create(C.byref(cfg), C.byref(h)) == pkg.ERR_ARG
    assert lib.mpc_create(None, C.byref(h)) == pkg.ERR_ARG
    assert lib.mpc_status_string(pkg.ERR_ARG) == b"bad argument"
    assert lib.mpc_destroy(None) == pkg.ERR_ARG


def test_synth_generator_matches_spec(pkg):
    for seed, N in ((0, 5), (12345, 33), (2**63 + 11, 3)):
        x0, q = pkg.synth_instance(seed, N)
        sx0, sq = helpers.synth_instance_spec(seed, N)
        assert np.allclose(x0, sx0, atol=1e-6) and np.allclose(q, sq, atol=1e-6)
    h = pkg.synth_horizons(7, 100)
    assert np.array_equal(h, helpers.synth_horizons_spec(7, 100))
    assert set(np.unique(h)) <= {32, 64, 128, 256}


def test_synth_distribution_moments(pkg):
    x0s = np.stack([pkg.synth_instance(s, 2)[0] for s in range(2000)])
    q0, qd0 = x0s[:, :7], x0s[:, 7:]
    assert np.all(np.abs(q0) <= 0.5) and abs(q0.mean()) < 0.02 and abs(q0.var() - 1 / 12) < 0.01
    assert abs(qd0.mean()) < 0.01 and abs(qd0.std() - 0.1) < 0.01


def test_static_direct_solve_rejects_arguments_without_a_device(pkg):
    lib = pkg.lib()
    for N in (0, 1, pkg.DIRECT_STATIC_MAX_KNOTS + 1):
        assert lib.mpc_direct_solve_static(N, 8, 8, 8, 2, None) == pkg.ERR_ARG
    assert lib.mpc_direct_solve_static(8, None, 8, 8, 2, None) == pkg.ERR_ARG
    assert lib.mpc_direct_solve_static(8, 8, 8, 8, 17, None) == pkg.ERR_ARG
