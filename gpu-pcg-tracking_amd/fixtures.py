"""Fixture and report formats of the reference's experiment driver.

Inputs: the precomputed trajectories under testfiles/: ``<s>_<g>_traj.csv``
(one knot per line, 21 comma-separated values: q, qd, u) and
``<s>_<g>_eepos.traj`` (6 values per line: end-effector xyz and rpy), both
read by readCSVToVecVec (runme.cu:15-40) and flattened row after row
(runme.cu:150-165).  Reports: the statistics line of printStats
(experiment_helpers.cuh:112-164), its CSV form (getStatsString,
experiment_helpers.cuh:93-110) and the overall-stats CSV (runme.cu:215-233),
and the per-run result files of dump_tracking_data (testutils.cuh:44-90).

closed_loop_inputs turns a trajectory pair into the device inputs of
Solver.closed_loop, extended past its end by the goal (last position, zero
velocity and control), as track() does near the end of the trajectory
(toplevel.cuh:680-704).
"""
from __future__ import annotations

import os

import numpy as np

XU_COLS, EE_COLS = 21, 6


def read_rows(path: str, ncols: int | None = None) -> np.ndarray:
    """Comma-separated rows as float32 (each field parsed as one float, like
    std::stof); empty fields (a trailing comma) are skipped."""
    rows = []
    with open(path) as f:
        for line in f:
            vals = [np.float32(float(v)) for v in line.strip().split(",") if v.strip() != ""]
            if vals:
                rows.append(vals)
    out = np.asarray(rows, dtype=np.float32)
    if ncols is not None and (out.ndim != 2 or out.shape[1] != ncols):
        raise ValueError(f"{path}: expected {ncols} columns, got shape {out.shape}")
    return out


def read_xu_traj(path: str) -> np.ndarray:
    """(K, 21) knots [q, qd, u] of a testfiles/*_traj.csv file."""
    return read_rows(path, XU_COLS)


def read_eepos_traj(path: str) -> np.ndarray:
    """(K, 6) end-effector poses of a testfiles/*_eepos.traj file."""
    return read_rows(path, EE_COLS)


def write_rows(path: str, rows, digits: int = 10) -> None:
    """Rows in the testfiles format (``digits`` significant digits, no trailing comma)."""
    with open(path, "w") as f:
        for r in np.asarray(rows):
            f.write(",".join(f"{float(v):.{digits}g}" for v in r) + "\n")


def flatten_xu(rows: np.ndarray) -> np.ndarray:
    """The xu layout (21 K - 7 floats) of K knot rows: the rows back to back,
    without the last knot's control."""
    rows = np.asarray(rows, dtype=np.float32)
    return rows.reshape(-1)[: XU_COLS * rows.shape[0] - 7].copy()


def extend_to_goal(xu_rows: np.ndarray, ee_rows: np.ndarray, extra: int):
    """Append ``extra`` knots holding the goal: the last position with zero
    velocity and control, and the last end-effector pose."""
    xu_rows, ee_rows = np.asarray(xu_rows, np.float32), np.asarray(ee_rows, np.float32)
    goal = np.zeros(XU_COLS, np.float32)
    goal[:7] = xu_rows[-1, :7]
    xu = np.concatenate([xu_rows, np.tile(goal, (extra, 1))])
    ee = np.concatenate([ee_rows, np.tile(ee_rows[-1], (extra, 1))])
    return xu, ee


# ---- statistics (printStats / getStatsString) -----------------------------------------------

def stats(values) -> dict:
    """Average, population standard deviation, min, max, median and the
    quartiles of printStats: for even n the mean of the two middle elements
    (and of the elements around n/4 and 3n/4), for odd n the elements at
    n/2, n/4 and 3n/4 of the sorted data."""
    v = np.asarray(values, dtype=np.float64)
    n = v.size
    if n == 0:
        raise ValueError("no data")
    s = np.sort(v)
    mean = float(v.mean())
    if n % 2 == 0:
        if n < 4:
            raise ValueError("printStats' quartiles need n >= 4 for even n")
        median = (s[n // 2 - 1] + s[n // 2]) / 2
        q1 = (s[n // 4 - 1] + s[n // 4]) / 2
        q3 = (s[3 * n // 4 - 1] + s[3 * n // 4]) / 2
    else:
        median, q1, q3 = s[n // 2], s[n // 4], s[3 * n // 4]
    return {"Average": mean, "Std Dev": float(np.sqrt(np.mean((v - mean) ** 2))), "Min": float(s[0]),
            "Max": float(s[-1]), "Median": float(median), "Q1": float(q1), "Q3": float(q3)}


def _g(x: float) -> str:
    # std::ostream's default formatting of a float: 6 significant digits, %g style
    return f"{x:g}"


def stats_string(values) -> str:
    """The line printStats prints and returns."""
    st = stats(values)
    return (f"Average[{_g(st['Average'])}] Std Dev [{_g(st['Std Dev'])}] Min [{_g(st['Min'])}] "
            f"Max [{_g(st['Max'])}] Median [{_g(st['Median'])}] Q1 [{_g(st['Q1'])}] Q3 [{_g(st['Q3'])}]")


def stats_csv(line: str) -> str:
    """getStatsString: the bracketed values of a stats line, comma-separated."""
    out, i = [], 0
    while True:
        a = line.find("[", i)
        if a < 0:
            break
        b = line.find("]", a + 1)
        if b < 0:
            break
        out.append(line[a + 1:b])
        i = b + 1
    return ",".join(out)


def write_overall_stats(path: str, lines) -> None:
    """The overall-stats CSV of runme.cu:215-233 (one row per stats line)."""
    with open(path, "w") as f:
        f.write("Average,Std Dev, Min, Max, Median, Q1, Q3\n")
        for line in lines:
            f.write(stats_csv(line) + "\n")


def dump_tracking_data(prefix: str, test_iter: int, tracking_path=None, **series) -> list[str]:
    """dump_tracking_data's files: ``<prefix>_<iter>_<name>.result`` with one
    value per line for each series (pcg_iters, linsys_times, sqp_times,
    sqp_iters, sqp_exits, tracking_errors, pcg_exits ...), and the
    tracking_path file with one comma-terminated state per line."""
    written = []
    for name, data in series.items():
        fn = f"{prefix}_{test_iter}_{name}.result"
        with open(fn, "w") as f:
            for v in np.asarray(data).reshape(-1):
                f.write((str(int(v)) if np.issubdtype(np.asarray(data).dtype, np.integer) or
                         np.asarray(data).dtype == bool else _g(float(v))) + "\n")
        written.append(fn)
    if tracking_path is not None:
        fn = f"{prefix}_{test_iter}_tracking_path.result"
        with open(fn, "w") as f:
            for row in np.asarray(tracking_path):
                f.write("".join(_g(float(x)) + "," for x in row) + "\n")
        written.append(fn)
    return written


# ---- device inputs of a closed-loop run ------------------------------------------------------

def closed_loop_inputs(xu_rows: np.ndarray, ee_rows: np.ndarray, N: int, steps: int | None = None):
    """(x0, ee path, trajectory, steps) as CUDA tensors for Solver.closed_loop:
    the run starts at the trajectory's first state and, by default, takes one
    control step per knot until the trajectory's last knot is the current
    target; the path and the trajectory are extended by the goal for the
    horizon beyond the end."""
    import torch
    K = xu_rows.shape[0]
    steps = K - 1 if steps is None else steps
    extra = max(steps + N - K, 0)
    xu, ee = extend_to_goal(xu_rows, ee_rows, extra)
    x0 = torch.from_numpy(np.ascontiguousarray(xu_rows[0, :14])).cuda()
    path = torch.from_numpy(np.ascontiguousarray(ee[: steps + N]).reshape(-1)).cuda()
    traj = torch.from_numpy(flatten_xu(xu[: steps + N])).cuda()
    return x0, path, traj, steps


def testfile_paths(root: str, start: int, goal: int):
    """testfiles/<start>_<goal>_traj.csv and _eepos.traj under root (runme.cu:145-150)."""
    return (os.path.join(root, "testfiles", f"{start}_{goal}_traj.csv"),
            os.path.join(root, "testfiles", f"{start}_{goal}_eepos.traj"))
