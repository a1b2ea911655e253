"""Seeded synthetic site sets (FP32 x, y, z, r).

The uniform stream is the reference's sample generator
(app/common/samples.cpp:64-85): a PCG hash of the draw index scaled to
[0, 1].  Draw index v = (seed << 28) + 4*i + c, so a seed selects a disjoint
block of the stream for any n < 2**26.  Values are formed in FP64 and rounded
to FP32 once.
"""
from __future__ import annotations

import ctypes
import os

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


# Element radii (Å) and frequencies of the protein-like sets: C, N, O, S.
PROTEIN_RADII = np.array([1.70, 1.55, 1.52, 1.80])
PROTEIN_FREQ = np.array([0.62, 0.17, 0.19, 0.02])


def _samples_lib():
    global _SAMPLES
    if _SAMPLES is None:
        from . import build as _build
        if not os.path.exists(_build.SAMPLES):
            _build.build_samples()
        L = ctypes.CDLL(_build.SAMPLES)
        L.apo_samples_protein.restype = ctypes.c_int
        L.apo_samples_protein.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_float)]
        _SAMPLES = L
    return _SAMPLES


_SAMPLES = None


def protein_like_sites(n: int, seed: int = 2, density: float = 0.05, seed_tag: int = 0) -> np.ndarray:
    """Protein-like atoms in Å, packed at `density` atoms/Å^3 in a cube of
    side (n / density)^(1/3) (generator in apo_amd/csrc/samples.cpp).

    Chains are random walks of 3.8 Å steps with bond angle uniform in
    [80°, 150°] and torsion uniform in [0°, 360°); each residue carries 1-4
    side atoms at 1.5 Å in random directions.  An atom closer than 1.2 Å to
    an existing one is redrawn (up to 24 times, then the chain ends); a step
    leaving the cube also ends the chain.  Elements C/N/O/S are drawn with
    frequencies .62/.17/.19/.02 and get radii 1.70/1.55/1.52/1.80 Å.
    Randomness: PCG32 with state seed `seed` and stream `seed_tag`.
    """
    out = np.zeros((n, 4), np.float32)
    st = _samples_lib().apo_samples_protein(n, seed, seed_tag, density,
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if st != 0:
        raise ValueError("invalid protein-like sample arguments")
    return out


def molecule_sizes(count: int = 512, seed: int = 4, lo: int = 500, hi: int = 50000) -> np.ndarray:
    """Atom counts log-uniform in [lo, hi]: round(lo * (hi/lo)^u) with u the
    reference stream's draw (seed << 28) + 4m."""
    base = np.uint32((seed << 28) & 0xFFFFFFFF)
    with np.errstate(over="ignore"):
        u = pcg_unit(np.arange(count, dtype=np.uint32) * np.uint32(4) + base)
    return np.rint(lo * np.power(hi / lo, u)).astype(np.int64)


def batched_molecules(count: int = 512, seed: int = 4, lo: int = 500, hi: int = 50000,
                      density: float = 0.05):
    """A batch of independent protein-like molecules: molecule m has
    molecule_sizes(...)[m] atoms from protein_like_sites(seed=seed, seed_tag=m),
    each in its own frame.  Returns (sites (N, 4) float32, offsets (count+1,) uint32)."""
    sizes = molecule_sizes(count, seed, lo, hi)
    offsets = np.zeros(count + 1, np.uint32)
    offsets[1:] = np.cumsum(sizes)
    sites = np.empty((int(offsets[-1]), 4), np.float32)
    for m in range(count):
        sites[offsets[m]:offsets[m + 1]] = protein_like_sites(int(sizes[m]), seed=seed, density=density, seed_tag=m)
    return sites, offsets


def crowded_shell_sites(shell: int, seed: int = 7) -> np.ndarray:
    """One large site (r = 0.05 at the cube's centre) ringed by `shell` small
    sites (r in [0.001, 0.003]) on a thin shell at distance 0.062-0.07, plus
    200 scattered sites.  The large site's cell has one edge per pair of
    adjacent shell sites: about 200 edges at shell = 60, and more than the
    per-cell edge capacity at shell = 150 (the cell is flagged and
    truncated).  A stress case for the per-cell record capacities and for
    scans that take a cell's edges 64 at a time."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(shell, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ring = np.zeros((shell, 4), np.float32)
    ring[:, :3] = 0.5 + d * rng.uniform(0.062, 0.07, (shell, 1))
    ring[:, 3] = rng.uniform(0.001, 0.003, shell)
    bg = np.zeros((200, 4), np.float32)
    bg[:, :3] = rng.random((200, 3))
    bg[:, 3] = rng.uniform(0.01, 0.03, 200)
    return np.concatenate([[[0.5, 0.5, 0.5, 0.05]], ring, bg]).astype(np.float32)
