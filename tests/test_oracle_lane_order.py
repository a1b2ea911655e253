"""The wave64 order against the reference's 32-lane order, on the CPU.

The HIP kernels (and the oracle by default) run the validation loops in
steps of 64 records and apply the contributors found in a step lowest lane
first; the reference runs 32-lane warps and visits ballot masks highest
lane first (forEach, apo/include/apo/gpu/utils.cuh:17-32).  The oracle runs
either order (lanes = 64 or 32); the emitted vertices (ids and FP64 bits)
must be the same on a 10k uniform set (C1), a 20k protein-like set (C2's
generator) and a 100k log-uniform set (C3's distribution), with and without
the completion mode.
"""
import numpy as np
import pytest

import oracle
from apo_amd import datasets

SETS = {
    "uniform_10k": lambda: datasets.uniform_sites(10_000, seed=1),
    "protein_20k": lambda: datasets.protein_like_sites(20_000, seed=2),
    "loguniform_100k": lambda: datasets.loguniform_sites(100_000, seed=3),
}


def _rows(d):
    x, ids = d.vertices()
    a = np.concatenate([ids.astype(np.uint64), x.view(np.uint64)], axis=1)
    return a[np.lexsort(a.T[::-1])]


@pytest.mark.parametrize("complete", [0, 4])
@pytest.mark.parametrize("name", sorted(SETS))
def test_vertex_set_does_not_depend_on_the_lane_order(name, complete):
    sites = SETS[name]()
    bound = 1 if len(sites) > 30_000 else 0
    wave64 = _rows(oracle.Diagram(sites, lanes=64, complete=complete, edge_bound=bound))
    warp32 = _rows(oracle.Diagram(sites, lanes=32, complete=complete, edge_bound=bound))
    assert len(wave64) > 0
    assert np.array_equal(wave64, warp32)
