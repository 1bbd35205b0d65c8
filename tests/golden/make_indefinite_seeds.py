"""Writes indefinite_seeds_4m.npy: the 1000 seed points (graph positions 0..999) of the
graph that build_graph makes for 4,194,304 points of the clustered 3-D mixture of the
headline workload (64 clusters, widths 1e-4..1e-1, seed 2024), n0 = 1000, k = 30.  With
the headline's Matern-3/2 table (cutoff 0.05, 2001 bins) and jitter 1e-5 their
interpolated covariance is indefinite.  Needs the GPU build; run from the repo root:

    python3 tests/golden/make_indefinite_seeds.py
"""
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import graphgp_amd as gp  # noqa: E402
from bench import clustered  # noqa: E402

n = 1 << 22
pts = torch.as_tensor(clustered(n, 3, 64, 1e-4, 1e-1, seed=2024), device="cuda")
g = gp.build_graph(pts, n0=1000, k=30)
np.save(os.path.join(os.path.dirname(os.path.abspath(__file__)), "indefinite_seeds_4m.npy"),
        g.points[:1000].cpu().numpy())
