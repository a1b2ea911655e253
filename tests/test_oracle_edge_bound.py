"""The edge radius bound prunes the edge contributor search without changing
any result: the oracle built with the bound (on large search spheres, or on
every edge search) must equal the reference search record for record.

The bound is this build's own (the reference searches the whole sphere of a
closed trisector's extreme points, apo/include/apo/gpu/algorithm.inl:572-585),
so equality with the unpruned restatement is what pins it.
"""
import numpy as np
import pytest

import oracle
from apo_amd import datasets


def _same(a, b):
    ca, Va, Ea = a.all_cells()
    cb, Vb, Eb = b.all_cells()
    for key in ca:
        assert np.array_equal(ca[key], cb[key]), key
    assert np.array_equal(Va, Vb)
    assert np.array_equal(Ea, Eb)
    xa, ia = a.vertices()
    xb, ib = b.vertices()
    assert np.array_equal(ia, ib)
    assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))


CASES = {
    "uniform_3k": lambda: datasets.uniform_sites(3000, seed=1),
    "loguniform_6k": lambda: datasets.loguniform_sites(6000, seed=3),
    "loguniform_6k_seed9": lambda: datasets.loguniform_sites(6000, seed=9),
    "wide_radii_2k": lambda: datasets.uniform_sites(2000, seed=5, r_min=0.0, r_max=0.05),
    "protein_4k": lambda: datasets.protein_like_sites(4000, seed=2),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("mode", [1, 2])
def test_bounded_search_matches_reference_search(name, mode):
    sites = CASES[name]()
    ref = oracle.Diagram(sites, threads=4)
    got = oracle.Diagram(sites, threads=4, edge_bound=mode)
    _same(ref, got)



@pytest.mark.parametrize("name", ["loguniform", "protein"])
def test_bisector_bound_keeps_results(name):
    """The bisector searches' subtree test (a box that cannot hold a site
    within the cylinder is skipped) changes no result: every cell record,
    count, flag and vertex equals the reference's unpruned search."""
    sites = (datasets.loguniform_sites(20_000, seed=21) if name == "loguniform"
             else datasets.protein_like_sites(20_000, seed=22))
    a = oracle.Diagram(sites, bisector_bound=1)
    b = oracle.Diagram(sites, bisector_bound=0)
    ca, Va, Ea = a.all_cells()
    cb, Vb, Eb = b.all_cells()
    for key in ("status", "vertex_nb", "edge_nb", "flags"):
        assert np.array_equal(ca[key], cb[key]), key
    assert np.array_equal(Va, Vb) and np.array_equal(Ea, Eb)
    xa, ia = a.vertices()
    xb, ib = b.vertices()
    assert np.array_equal(ia, ib) and np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
