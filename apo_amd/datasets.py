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


# Element radii (Å) and frequencies of the protein-like sets: C, N, O, S.
PROTEIN_RADII = np.array([1.70, 1.55, 1.52, 1.80])
PROTEIN_FREQ = np.array([0.62, 0.17, 0.19, 0.02])


def protein_like_sites(n: int, seed: int = 2, density: float = 0.05, seed_tag: int = 0) -> np.ndarray:
    """Protein-like atoms in Å, packed at `density` atoms/Å^3 in a cube of
    side (n / density)^(1/3).

    Chains are random walks of 3.8 Å steps with bond angle uniform in
    [80°, 150°] and torsion uniform in [0°, 360°); each residue carries 1-4
    side atoms at 1.5 Å in random directions.  An atom closer than 1.2 Å to
    an existing one is redrawn (up to 24 times, then the chain ends); a step
    leaving the cube also ends the chain.  Elements C/N/O/S are drawn with
    frequencies .62/.17/.19/.02 and get radii 1.70/1.55/1.52/1.80 Å.
    Randomness: numpy's PCG64 seeded with (seed, seed_tag).
    """
    rng = np.random.Generator(np.random.PCG64([seed, seed_tag]))
    side = (n / density) ** (1.0 / 3.0)
    cell = 1.2
    dims = max(1, int(side / cell))
    grid: dict = {}
    pos = np.empty((n, 3), np.float64)
    count = 0

    def key(p):
        return (min(int(p[0] / cell), dims - 1), min(int(p[1] / cell), dims - 1), min(int(p[2] / cell), dims - 1))

    def free(p):
        if not (0.0 <= p[0] < side and 0.0 <= p[1] < side and 0.0 <= p[2] < side):
            return False
        kx, ky, kz = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((kx + dx, ky + dy, kz + dz), ()):
                        d = p - pos[q]
                        if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < cell * cell:
                            return False
        return True

    def add(p):
        nonlocal count
        pos[count] = p
        grid.setdefault(key(p), []).append(count)
        count += 1

    def unit():
        v = rng.normal(size=3)
        return v / np.linalg.norm(v)

    while count < n:
        p0 = rng.uniform(0.0, side, 3)
        if not free(p0):
            continue
        add(p0)
        prev, prev2 = p0, None
        residues = int(rng.integers(50, 400))
        for _ in range(residues):
            if count >= n:
                break
            placed = False
            for _try in range(24):
                if prev2 is None:
                    q = prev + 3.8 * unit()
                else:
                    # next backbone atom from bond angle and torsion (NeRF)
                    theta = np.radians(rng.uniform(80.0, 150.0))
                    phi = rng.uniform(0.0, 2.0 * np.pi)
                    b = prev - prev2
                    b /= np.linalg.norm(b)
                    a = unit()
                    nrm = np.cross(b, a)
                    nn = np.linalg.norm(nrm)
                    if nn < 1e-6:
                        continue
                    nrm /= nn
                    m = np.cross(nrm, b)
                    d = np.array([-np.cos(theta), np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi)])
                    q = prev + 3.8 * (d[0] * b + d[1] * m + d[2] * nrm)
                if free(q):
                    placed = True
                    break
            if not placed:
                break
            add(q)
            prev2, prev = prev, q
            for _s in range(int(rng.integers(1, 5))):
                if count >= n:
                    break
                for _try in range(24):
                    s = q + 1.5 * unit()
                    if free(s):
                        add(s)
                        break
    elem = rng.choice(4, size=n, p=PROTEIN_FREQ)
    out = np.empty((n, 4), np.float64)
    out[:, :3] = pos
    out[:, 3] = PROTEIN_RADII[elem]
    return out.astype(np.float32)
