"""Synthetic workloads of the benchmark configurations.

The batched-ensemble workload: Orszag-Tang instances ("otv") of sizes drawn
from ENSEMBLE_SIZES, B amplitude uniform in [0.8, 1.2] and the phases added
to the four sine terms uniform in [0, 2 pi), from a seeded generator, each
advanced to t = 0.5 with its own dt.
"""
import numpy as np

ENSEMBLE_SIZES = (128, 192, 256, 384, 512, 768, 1024)
ENSEMBLE_SEED = 2026


def ensemble_workload(n, seed=ENSEMBLE_SEED):
    """(sizes, params) of n instances: sizes[i] is the side of instance i's
    square grid, params[i] its five "otv" parameters."""
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.choice(ENSEMBLE_SIZES, size=n)]
    params = [(float(rng.uniform(0.8, 1.2)),) + tuple(float(x) for x in rng.uniform(0, 2 * np.pi, 4))
              for _ in range(n)]
    return sizes, params
