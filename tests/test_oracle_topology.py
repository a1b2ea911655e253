"""toTopology restatement (oracle.Diagram.topology, algorithm.inl:1732-1821)
checked against invariants of the diagram it describes: the quadrisector
set is the emitted vertex set, every face of a quadrisector is a
trisector, every pair of a trisector is a bisector, no artificial site, and
each cell's entries are strictly ascending."""
import numpy as np
import pytest

import oracle
from apo_amd import datasets


@pytest.fixture(scope="module", params=["uniform", "loguniform"])
def diagram(request):
    if request.param == "uniform":
        sites = datasets.uniform_sites(2500, seed=2)
    else:
        sites = datasets.loguniform_sites(4000, seed=5)
    return sites, oracle.Diagram(sites, threads=4)


def test_quadrisectors_are_the_vertex_set(diagram):
    sites, d = diagram
    t = d.topology()
    _, ids = d.vertices()
    assert {tuple(r) for r in np.sort(ids, axis=1)} == {tuple(sorted(r)) for r in t["quadrisectors"]}


def test_faces_and_edges_are_present(diagram):
    sites, d = diagram
    t = d.topology()
    tri = {tuple(sorted(r)) for r in t["trisectors"]}
    bis = {tuple(sorted(r)) for r in t["bisectors"]}
    for q in t["quadrisectors"]:
        i, j, k, l = q
        for f in ((i, j, k), (i, j, l), (i, k, l)):
            assert tuple(sorted(f)) in tri
    for r in t["trisectors"]:
        i, j, k = r
        assert tuple(sorted((i, j))) in bis and tuple(sorted((i, k))) in bis


def test_ids_and_order(diagram):
    sites, d = diagram
    n = len(sites)
    t = d.topology()
    _, perm = d.sorted_sites()
    rank = np.empty(n, np.int64)
    rank[perm] = np.arange(n)
    for key in ("bisectors", "trisectors", "quadrisectors"):
        a = t[key]
        assert a.max() < n
        s = rank[a]  # back to sorted space: entries are (i, ...) with i smallest
        assert np.all(s[:, 1:] > s[:, :1])
        # cells in ascending order, entries ascending within a cell
        keys = [tuple(r) for r in s]
        assert keys == sorted(keys)
        assert len(set(keys)) == len(keys)
