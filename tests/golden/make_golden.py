"""Writes tests/golden/cfg1_ewald.npz: FP64 Ewald potentials (oracle cell-list Ewald,
s = 6, pinned by the Madelung constant and by agreement with the O(N^2) ewald_direct)
for the N = 1e4 uniform cubic workload (positions seed 1, charges seed 2).

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from periodicdmk_amd import workloads as W  # noqa: E402

cell = W.cubic_cell()
pos = W.uniform_positions(10_000, 1, cell)
q = W.neutral_uniform_charges(10_000, 2)
pot = O.ewald_cells(cell, pos, q, 6.0)
np.savez_compressed(os.path.join(HERE, "cfg1_ewald.npz"), ewald=pot,
                    input_check=np.array([pos.sum(), pos[0, 0], pos[-1, 2], q.sum(), q[0], q[-1]]))
print("wrote cfg1_ewald.npz", pot[:3])
