"""Seeded ensembles of independent grids for batched stepping.

The ensemble workload: Cartesian periodic grids whose sizes are drawn by seed
from ENSEMBLE_SIZES, each with the sinusoid generator and its own seed, half
of them (by seed) with the over-pressure sphere.  The benchmark and the
ensemble tests draw their grids with this one rule.
"""
import numpy as np

import gamera_op as go

ENSEMBLE_SIZES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 64, 32)]


def ensemble_specs(n, seed):
    """n (kind, problem seed, shape) triples drawn from the seeded rule."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        shape = ENSEMBLE_SIZES[rng.integers(len(ENSEMBLE_SIZES))]
        kind = go.SINUSOID_SPHERE if rng.random() < 0.5 else go.SINUSOID
        specs.append((kind, int(rng.integers(1, 2**31)), shape))
    return specs
