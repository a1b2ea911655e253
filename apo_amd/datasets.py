"""Seeded synthetic site sets (FP32 x, y, z, r).

The uniform stream is the reference's sample generator
(app/common/samples.cpp:64-85): a PCG hash of the draw index scaled to
[0, 1].  Draw index v = (seed << 28) + 4*i + c, so a seed selects a disjoint
block of the stream for any n < 2**26.  Values are formed in FP64 and rounded
to FP32 once.
"""
from __future__ import annotations

import numpy as np


def pcg_word(v: np.ndarray) -> np.ndarray:
    """The hash of app/common/samples.cpp:64-69 on uint32 draw indices."""
    v = np.asarray(v, dtype=np.uint32)
    with np.errstate(over="ignore"):
        state = v * np.uint32(747796405) + np.uint32(2891336453)
        shift = (state >> np.uint32(28)) + np.uint32(4)
        word = ((state >> shift) ^ state) * np.uint32(277803737)
    return (word >> np.uint32(22)) ^ word


def pcg_unit(v: np.ndarray) -> np.ndarray:
    return pcg_word(v).astype(np.float64) / 4294967295.0


def _draws(n: int, seed: int) -> np.ndarray:
    if n >= (1 << 26):
        raise ValueError("n must be below 2**26 for one seed block")
    base = np.uint32((seed << 28) & 0xFFFFFFFF)
    idx = (np.arange(n, dtype=np.uint32)[:, None] * np.uint32(4) + np.arange(4, dtype=np.uint32)[None, :])
    with np.errstate(over="ignore"):
        return pcg_unit(idx + base)


def uniform_sites(n: int, seed: int = 1, r_min: float = 0.002, r_max: float = 0.008) -> np.ndarray:
    """Centres uniform in [0,1]^3, radii uniform in [r_min, r_max]."""
    u = _draws(n, seed)
    out = np.empty((n, 4), np.float64)
    out[:, :3] = u[:, :3]
    out[:, 3] = r_min + (r_max - r_min) * u[:, 3]
    return out.astype(np.float32)


def loguniform_sites(n: int, seed: int = 3, r_lo: float = 1e-4, r_hi: float = 1e-2) -> np.ndarray:
    """Centres uniform in [0,1]^3, radii log-uniform in [r_lo, r_hi]."""
    u = _draws(n, seed)
    out = np.empty((n, 4), np.float64)
    out[:, :3] = u[:, :3]
    out[:, 3] = r_lo * np.power(r_hi / r_lo, u[:, 3])
    return out.astype(np.float32)
