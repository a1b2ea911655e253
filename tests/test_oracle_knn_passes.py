"""kNN continuation passes (knnValidationStep) in the oracle, pinned on the CPU.

The reference's AlgorithmGPU<K>(sites, knnValidationStep) runs kNN + cells
that many times (algorithm.inl:1535-1549); pass s > 0 restarts each list
after the previous pass's K-th gap, strictly, and a list the previous pass
did not fill becomes 0xffffffff x K (algorithm.inl:43-60).  The lists are
checked against a brute-force continuation over all sites, and the diagrams
against the brute-force Apollonius diagram on tiny sets.
"""
import numpy as np
import pytest

import oracle
from apo_amd import datasets

NONE = 0xFFFFFFFF


def _gaps_f32(s):
    """sphereDistance in FP32 (math.cuh:368-372): |ci - cj| - (ri + rj), the
    squared length summed x, y, z in that order, every operation rounded."""
    c = s[:, :3]
    d = c[:, None, :] - c[None, :, :]
    sq = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return np.sqrt(sq) - (s[:, 3][:, None] + s[:, 3][None, :])


def _brute_passes(s, k, passes):
    """The lists after `passes` passes by brute force over all sites."""
    n = len(s)
    g = _gaps_f32(s)
    np.fill_diagonal(g, np.inf)
    order = np.argsort(g, axis=1, kind="stable")
    lists = np.full((n, k), NONE, np.uint32)
    for p in range(passes):
        new = np.full((n, k), NONE, np.uint32)
        for i in range(n):
            if p == 0:
                start = -np.inf
            else:
                last = lists[i, k - 1]
                if last == NONE:
                    continue
                start = g[i, last]
            cand = order[i][g[i, order[i]] > start]
            cand = cand[np.isfinite(g[i, cand])][:k]
            new[i, :len(cand)] = cand
        lists = new
    return lists


@pytest.mark.parametrize("k,passes", [(16, 2), (16, 3), (1, 4), (4, 3), (8, 5)])
def test_continuation_lists_equal_brute_force(k, passes):
    sites = datasets.uniform_sites(1500, seed=31, r_min=0.002, r_max=0.02)
    d = oracle.Diagram(sites, k=k, knn_passes=passes, stage_mask=1)
    s, _ = d.sorted_sites()
    want = _brute_passes(s, k, passes)
    assert np.array_equal(d.knns(), want)


def test_exhausted_lists_end():
    # 40 sites, 16 per pass: the third pass finds 7, the fourth none (the
    # previous list was not full), the fifth repeats the empty list
    sites = datasets.uniform_sites(40, seed=7, r_min=0.01, r_max=0.03)
    for passes, filled in ((3, 7), (4, 0), (5, 0)):
        d = oracle.Diagram(sites, k=16, knn_passes=passes, stage_mask=1)
        got = d.knns()
        assert np.all((got != NONE).sum(axis=1) == filled), passes


def _rows(ids):
    a = np.sort(np.asarray(ids, np.uint32).reshape(-1, 4), axis=1)
    return a[np.lexsort(a.T[::-1])] if len(a) else a


@pytest.mark.parametrize("k,passes", [(16, 2), (1, 4), (4, 3), (16, 0), (2, 1)])
@pytest.mark.parametrize("n,seed", [(16, 2), (32, 4), (40, 5)])
def test_diagram_with_passes_equals_brute_force(k, passes, n, seed):
    sites = datasets.uniform_sites(n, seed=100 + seed, r_min=0.01, r_max=0.06)
    _, ids = oracle.Diagram(sites, k=k, knn_passes=passes).vertices()
    brute = oracle.bruteforce_vertices(sites)
    assert np.array_equal(_rows(ids), _rows(brute)), (k, passes, len(ids), len(brute))


def test_more_passes_certify_more_cells():
    sites = datasets.uniform_sites(5000, seed=9)
    uncertified = []
    for passes in (1, 2, 3, 4):
        d = oracle.Diagram(sites, k=16, knn_passes=passes, stage_mask=1)
        uncertified.append(int((d.cell_counts()["status"] == 0).sum()))
    assert uncertified == sorted(uncertified, reverse=True)
    assert uncertified[-1] < uncertified[0]


def test_zero_passes_start_every_cell_from_the_octahedron():
    sites = datasets.uniform_sites(300, seed=5)
    d = oracle.Diagram(sites, knn_passes=0, stage_mask=0)
    c = d.cell_counts()
    assert np.all(c["status"] == 0)
    assert np.all(c["vertex_nb"] == 8) and np.all(c["edge_nb"] == 12)
    assert np.all(c["known_radius"] == -np.finfo(np.float32).max)
    assert np.all(d.knns() == NONE)
