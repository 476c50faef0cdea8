"""CPU: the host-side arithmetic of bench.py that turns measurements into
the reported figures: the per-phase roofline (time = max over the HBM,
LDS and fp32 limits, fraction = that time over the measured one) and the
threaded CPU baseline leg.  Checked against values worked out by hand."""
import importlib.util
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_phase_roofline_bounds_and_fractions(bench):
    bw = {"hbm_gbs": 5000.0, "lds_gbs": 100000.0, "method": "test"}
    phases = {"linearize": 200.0, "schur": 400.0, "pcg_team": 1500.0, "pcg_rest": 500.0, "dz": 60.0,
              "merit": 80.0, "update": 10.0}
    knots, pcg_units = 1.0e5, 1.5e7
    out = bench.phase_roofline(phases, knots, pcg_units, bw)
    rows = out["phases"]
    assert set(rows) == {"linearize", "schur", "pcg", "dz", "merit"}
    # schur (split form): 2464 B / knot (the lower rows of C of the knot
    # before, g_q, q, r, c in; L^, W and gamma^ out) over 5 TB/s against
    # 2.5e4 flops / knot over 157.3 TFLOP/s
    t_hbm = 4 * (147 + 14 + 35 + 14 + 392 + 14) * knots / 5000e9
    t_f = 2.5e4 * knots / 157.3e12
    assert rows["schur"]["bound"] == ("hbm" if t_hbm > t_f else "fp32")
    assert rows["schur"]["roofline_us"] == pytest.approx(max(t_hbm, t_f) * 1e6, abs=0.01)
    assert rows["schur"]["frac"] == pytest.approx(max(t_hbm, t_f) * 1e6 / 400.0, abs=1e-4)  # 4 decimals
    # pcg: both PCG phases together, 2296 B per knot-iteration at the LDS rate
    assert rows["pcg"]["us"] == pytest.approx(2000.0)
    t_lds = 2296 * pcg_units / 100000e9
    t_pf = 1708 * pcg_units / 157.3e12
    t_ph = 4 * (392 + 14 + 28) * knots / 5000e9
    assert rows["pcg"]["roofline_us"] == pytest.approx(max(t_lds, t_pf, t_ph) * 1e6, abs=0.01)
    assert rows["pcg"]["lds_bytes"] == int(2296 * pcg_units)
    # merit has no flops: priced by bytes
    assert rows["merit"]["flops"] == 0 and rows["merit"]["bound"] == "hbm"
    for r in rows.values():
        assert 0 < r["frac"] <= 1.5


def test_phase_roofline_picks_the_largest_limit(bench):
    # an LDS rate so low that every phase with LDS bytes is LDS-bound
    bw = {"hbm_gbs": 1e9, "lds_gbs": 1.0, "method": "test"}
    phases = dict.fromkeys(("linearize", "schur", "pcg_team", "pcg_rest", "dz", "merit", "update"), 100.0)
    rows = bench.phase_roofline(phases, 10.0, 10.0, bw)["phases"]
    assert rows["pcg"]["bound"] == "lds"
    assert rows["dz"]["bound"] == "fp32"  # no LDS bytes, HBM nearly free


def test_phase_roofline_uses_counted_flops(bench):
    bw = {"hbm_gbs": 5000.0, "lds_gbs": 100000.0, "method": "test"}
    phases = dict.fromkeys(("linearize", "schur", "pcg_team", "pcg_rest", "dz", "merit", "update"), 100.0)
    counters = {"flops_per_knot": {"linearize": 6.0e4, "merit": 2.0e4}}
    out = bench.phase_roofline(phases, 1.0e5, 1.0e6, bw, counters)
    rows = out["phases"]
    assert rows["linearize"]["flops"] == int(6.0e4 * 1.0e5) and rows["merit"]["flops"] == int(2.0e4 * 1.0e5)
    assert out["flops_source"] == {"linearize": "counted", "schur": "estimate", "dz": "estimate", "merit": "counted"}
    assert bench.phase_roofline(phases, 1.0e5, 1.0e6, bw)["flops_source"]["linearize"] == "estimate"


def test_cpu_baseline_threads_counts_solves(bench, orc, pkg):
    horizons = np.array([8, 16] * 4)
    rates, desc, t_per, per_n = bench.cpu_baseline(orc, pkg, None, horizons, 0.2, 5)
    assert rates["ldl"] > 0 and rates["schur"] > 0 and rates["ldl_with_copies"] is None
    assert set(t_per) == {8, 16} and "Schur formation" in desc
    two = bench.cpu_baseline_threads(orc, per_n, t_per, horizons, 0.2, 2)
    assert two > 0


def test_host_info_reports_cores_and_model(bench):
    h = bench.host_info()
    assert h["nproc"] >= 1 and 1 <= h["usable_cores"] <= h["nproc"]
    assert h["model"] is None or isinstance(h["model"], str)
