"""Fixture and report formats (gpu-pcg-tracking_amd/fixtures.py) against
the reference's own fixture data (testfiles/0_0_traj.csv and
0_0_eepos.traj, kept as tests/golden/fk_0_0.npz) and hand-computed
statistics in the conventions of printStats / getStatsString."""
import importlib
import os

import numpy as np
import pytest

import helpers


@pytest.fixture(scope="module")
def fx(pkg):
    return importlib.import_module(pkg.__name__ + ".fixtures")


@pytest.fixture(scope="module")
def golden():
    g = np.load(os.path.join(helpers.GOLDEN, "fk_0_0.npz"))
    return g["xu"], g["ee"]


def test_trajectory_files_round_trip(fx, golden, tmp_path):
    xu, ee = golden
    p1, p2 = tmp_path / "0_0_traj.csv", tmp_path / "0_0_eepos.traj"
    fx.write_rows(p1, xu.astype(np.float32))
    fx.write_rows(p2, ee.astype(np.float32))
    a, b = fx.read_xu_traj(p1), fx.read_eepos_traj(p2)
    assert a.shape == (666, 21) and b.shape == (666, 6) and a.dtype == np.float32
    assert np.array_equal(a, xu.astype(np.float32)) and np.array_equal(b, ee.astype(np.float32))
    # the first row of the reference file, as it reads
    assert a[0, 0] == np.float32(0.4550000131) and a[0, 13] == np.float32(-0.009999999776)
    with pytest.raises(ValueError):
        fx.read_xu_traj(p2)


def test_trailing_commas_and_tracking_path_format(fx, tmp_path):
    p = tmp_path / "path.result"
    fx.dump_tracking_data(str(tmp_path / "run"), 3, tracking_path=[[1.0, 2.5], [0.125, -3.0]],
                          pcg_iters=np.array([5, 7], np.int32), pcg_exits=np.array([True, False]),
                          tracking_errors=np.array([0.5, 0.25]))
    assert (tmp_path / "run_3_pcg_iters.result").read_text() == "5\n7\n"
    assert (tmp_path / "run_3_pcg_exits.result").read_text() == "1\n0\n"
    assert (tmp_path / "run_3_tracking_errors.result").read_text() == "0.5\n0.25\n"
    text = (tmp_path / "run_3_tracking_path.result").read_text()
    assert text == "1,2.5,\n0.125,-3,\n"
    p.write_text(text)
    assert np.array_equal(fx.read_rows(p), np.array([[1, 2.5], [0.125, -3]], np.float32))


def test_flatten_and_goal_extension(fx, golden):
    xu = golden[0].astype(np.float32)
    flat = fx.flatten_xu(xu[:5])
    assert flat.shape == (21 * 5 - 7,) and np.array_equal(flat[21:42], xu[1])
    x2, e2 = fx.extend_to_goal(xu, golden[1], 3)
    assert x2.shape == (669, 21) and e2.shape == (669, 6)
    assert np.array_equal(x2[-1, :7], xu[-1, :7]) and np.all(x2[-3:, 7:] == 0)
    assert np.array_equal(e2[-1], golden[1][-1].astype(np.float32))


def test_stats_conventions(fx):
    st = fx.stats([4.0, 1.0, 3.0, 2.0])  # even n: means of the middle pairs
    assert st == pytest.approx({"Average": 2.5, "Std Dev": np.sqrt(1.25), "Min": 1, "Max": 4, "Median": 2.5,
                                "Q1": 1.5, "Q3": 3.5})
    st = fx.stats([5.0, 1.0, 3.0])  # odd n: elements n/2, n/4, 3n/4 of the sorted data
    assert (st["Median"], st["Q1"], st["Q3"]) == (3.0, 1.0, 5.0)
    st = fx.stats(np.arange(1.0, 9.0))  # n = 8: Q1 = (s[1] + s[2]) / 2, Q3 = (s[5] + s[6]) / 2
    assert (st["Q1"], st["Median"], st["Q3"]) == (2.5, 4.5, 6.5)
    line = fx.stats_string([4.0, 1.0, 3.0, 2.0])
    assert line == "Average[2.5] Std Dev [1.11803] Min [1] Max [4] Median [2.5] Q1 [1.5] Q3 [3.5]"
    assert fx.stats_csv(line) == "2.5,1.11803,1,4,2.5,1.5,3.5"


def test_overall_stats_csv(fx, tmp_path):
    p = tmp_path / "x_overall_stats.csv"
    fx.write_overall_stats(p, [fx.stats_string([1.0, 2.0, 3.0]), fx.stats_string([10.0, 20.0, 30.0, 40.0])])
    lines = p.read_text().splitlines()
    assert lines[0] == "Average,Std Dev, Min, Max, Median, Q1, Q3"
    assert lines[1] == "2,0.816497,1,3,2,1,3"
    assert lines[2] == "25,11.1803,10,40,25,15,35"
