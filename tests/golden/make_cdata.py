"""Writes tests/golden/cylinder_cdata.npy from the reference's input fixture
2d/cdata (5346 rows of x, y, g, e read by 2d/cylinder.cxx:82-90).  Run once
in a container that has the reference tree; the .npy is committed."""

import os
import sys

import numpy as np

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/2d/cdata"
d = np.loadtxt(src).reshape(-1, 4)
assert d.shape == (5346, 4), d.shape
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cylinder_cdata.npy")
np.save(out, d)
print(out, d.shape)
