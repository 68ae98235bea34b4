"""Ensemble statistics of long particle runs, whose trajectories are chaotic
so that runs are compared by distributions rather than states: the kinetic
energy decay curve (FP64 reduction on the GPU), the contact-count histogram
(exact integer counts on the GPU) and the packing fraction of the settled
bed.  Also the log-normal radius distribution of polydisperse beds."""
from __future__ import annotations

import numpy as np


def packing_fraction(pos4: np.ndarray, radius=None, floor: float = -1.0,
                     half_width: float = 1.0, quantile: float = 0.95) -> float:
    """Sphere volume over bed volume below the `quantile` height of the
    sphere centres: the bed is the box cross-section (2*half_width)^2 from
    `floor` up to that height plus one radius.  Radii are pos.w when
    `radius` is None, else the given scalar or array."""
    pos4 = np.asarray(pos4, np.float64).reshape(-1, 4)
    r = pos4[:, 3] if radius is None else np.broadcast_to(np.asarray(radius, np.float64),
                                                          (len(pos4),))
    z = pos4[:, 2]
    h = float(np.quantile(z, quantile, method="lower"))
    below = z <= h
    top = h + float(np.max(r[below]))
    vol = float(np.sum(4.0 / 3.0 * np.pi * r[below] ** 3))
    return vol / ((2.0 * half_width) ** 2 * (top - floor))


def polydisperse_radii(n: int, median: float, sigma: float = 0.25, ratio: float = 3.0,
                       seed: int = 1973) -> np.ndarray:
    """Log-normal radii (median, sigma of log r) clipped to
    [median/sqrt(ratio), median*sqrt(ratio)], as float32."""
    rng = np.random.default_rng(seed)
    r = median * np.exp(sigma * rng.standard_normal(n))
    lo, hi = median / np.sqrt(ratio), median * np.sqrt(ratio)
    return np.clip(r, lo, hi).astype(np.float32)


def energy_curve(ps, dt: float, steps: int, every: int) -> np.ndarray:
    """Kinetic energy (FP64, GPU) after every `every` steps of ps.run."""
    out = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        ps.run(dt, k)
        done += k
        out.append(ps.stats_fp64()[0])
    return np.array(out)
