"""Shared inputs for the tests: the reference's seeded data (via the JAX RNG restatement)
and the synthetic workloads used by the benchmark."""
import numpy as np

import jax_rng


def ref_points(shape, seed=137):
    """jax.random.normal(jax.random.key(seed), shape) under x64 (reference tests)."""
    return jax_rng.normal(jax_rng.key(seed), shape)


def clustered_points(n, d=3, n_clusters=64, wmin=1e-4, wmax=1e-1, seed=0, dtype=np.float32):
    """Mixture of isotropic Gaussians: centres U[0,1]^d, widths log-uniform in [wmin, wmax]."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (n_clusters, d))
    widths = np.exp(rng.uniform(np.log(wmin), np.log(wmax), n_clusters))
    lab = rng.integers(0, n_clusters, n)
    pts = centres[lab] + widths[lab, None] * rng.standard_normal((n, d))
    return pts.astype(dtype)
