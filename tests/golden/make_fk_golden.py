"""Writes tests/golden/fk_0_0.npz from the reference's own fixture files
testfiles/0_0_traj.csv (21-column xu rows) and testfiles/0_0_eepos.traj
(FK xyz + rpy of each row, written by generateeepos.cu:74-118).

Only the data is kept: q (666 x 7, the first 7 columns of each xu row) and
ee (666 x 6).  Run once in the build container:
    python tests/golden/make_fk_golden.py /root/reference
"""
import os
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
xu = np.loadtxt(os.path.join(ref, "testfiles", "0_0_traj.csv"), delimiter=",")
ee = np.loadtxt(os.path.join(ref, "testfiles", "0_0_eepos.traj"), delimiter=",")
assert xu.shape[0] == ee.shape[0] and xu.shape[1] == 21 and ee.shape[1] == 6
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fk_0_0.npz")
np.savez_compressed(out, q=xu[:, :7], ee=ee, xu=xu)
print(out, xu.shape, ee.shape)
