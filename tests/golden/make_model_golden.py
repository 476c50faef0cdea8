"""Writes tests/golden/iiwa_spatial_inertia.npy: the 7 spatial inertia
matrices (6 x 6, link frame) of the model table that the reference's
generated code initialises (rbdfiles/iiwa_eepos_grid.cuh:909-1672, entries
252..503 of h_XImats, column-major per matrix).  Data only.
    python tests/golden/make_model_golden.py /root/reference
"""
import os
import re
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "rbdfiles", "iiwa_eepos_grid.cuh")).read()
start = src.index("T* init_XImats()")
body = src[start:src.index("return h_XImats", start)] if "return h_XImats" in src[start:] else src[start:start + 200000]
vals = {}
for m in re.finditer(r"h_XImats\[(\d+)\] = static_cast<T>\(([-0-9.e]+)\)", body):
    vals[int(m.group(1))] = float(m.group(2))
I = np.array([[vals.get(252 + 36 * i + c * 6 + r, 0.0) for c in range(6)] for i in range(7) for r in range(6)])
I = I.reshape(7, 6, 6)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "iiwa_spatial_inertia.npy")
np.save(out, I)
print(out, I.shape)
