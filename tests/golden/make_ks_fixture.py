"""Writes tests/golden/ks_L22_M64_seed0_first2000_stats.npy: mean, std, max
|u|, first and last value of the first 2000 snapshots of the build's KS
generator (L = 22, M = 64, dt = 0.25, 10,000 steps discarded, stride 4,
seed 0).  Regenerate only when the generator's definition changes."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from kernel_evd_amd import ks_generate  # noqa: E402

X = ks_generate(2000, L_dom=22.0, M=64, n_discard=10000, stride=4, seed=0)
np.save(os.path.join(HERE, "ks_L22_M64_seed0_first2000_stats.npy"),
        np.array([X.mean(), X.std(), np.abs(X).max(), X[0, 0], X[-1, -1]]))
