"""Pins the oracle's closed trisector edges (toHostFull, algorithm.inl:1309-1383)
with a brute force that does not reuse its cell records.

A closed edge of the diagram is a trisector of three sites that is an ellipse
(two extreme points of the empty-sphere radius, both found here by the exact
solve of the tangency equations in the plane of the three centres), carries
no Apollonius vertex (no empty sphere tangent to the three and a fourth site,
artificial sites included), whose sites are not buried, and lies in
the diagram, which for an ellipse without vertices holds exactly when its
extreme point of smaller radius is an empty sphere.  The oracle must emit that
set, each edge once, at that extreme point.

The JSON topology document (app/topology/topology.cpp:11-57) is checked to
parse back into the sets it was written from.
"""
import itertools
import json

import numpy as np
import pytest

import apo_amd
import oracle
from apo_amd import datasets

from test_oracle_geometry import _tangent_points


def _empty(point, spheres, skip, tol=1e-9):
    d = np.linalg.norm(spheres[:, :3] - point[:3], axis=1) - spheres[:, 3]
    d[list(skip)] = np.inf
    return bool(np.all(d >= point[3] - tol))


def _mixed_sites(n, seed):
    """Centres U[0,1]^3, radii log-uniform in [0.005, 0.3]: small spheres in
    the gaps of large ones, where closed edges occur."""
    rng = np.random.default_rng(1000 * n + seed)
    s = np.zeros((n, 4), np.float32)
    s[:, :3] = rng.random((n, 3))
    s[:, 3] = np.exp(rng.uniform(np.log(0.005), np.log(0.3), n))
    return s


# (n, seed, closed edges): sets with closed edges, and ones without
@pytest.mark.parametrize("n,seed,count", [(24, 21, 3), (32, 15, 3), (32, 1, 2), (16, 0, None), (24, 0, None)])
def test_closed_edges_equal_brute_force(n, seed, count):
    sites = _mixed_sites(n, seed)
    d = oracle.Diagram(sites)
    ex, eids = d.closed_edges()
    art = oracle.artificial_sites(sites)
    everything = np.concatenate([sites.astype(np.float64), art])

    def has_vertex(t):
        # an empty tangent sphere of the three sites and a fourth one, real
        # or artificial (the quadrisector solve is pinned against the exact
        # one in test_oracle_geometry.py)
        for m in range(len(everything)):
            if m in t:
                continue
            mask, x1, x2 = oracle.quadrisector(*(everything[q] for q in t), everything[m])
            for r, x in enumerate((x1, x2)):
                if (mask >> r) & 1 and _empty(x, everything, t + (m,)):
                    return True
        return False
    # a site inside another sphere is buried and has no cell
    sd = sites.astype(np.float64)
    gap = np.linalg.norm(sd[:, None, :3] - sd[None, :, :3], axis=2) + sd[:, None, 3] - sd[None, :, 3]
    np.fill_diagonal(gap, np.inf)
    buried = set(np.nonzero(np.any(gap <= 0.0, axis=1))[0].tolist())
    want = {}
    for t in itertools.combinations(range(n), 3):
        if buried.intersection(t):
            continue
        pts, _ = _tangent_points([everything[m] for m in t], plane=True)
        if len(pts) != 2:
            continue
        low = min(pts, key=lambda p: p[3])
        if _empty(low, everything, t) and not has_vertex(t):
            want[t] = low
    # ids are original indices (the emission rule orders them in Morton space)
    got = {tuple(sorted(int(x) for x in row[:3])): x for row, x in zip(eids, ex)}
    assert len(got) == len(eids), "an edge was emitted twice"
    assert np.all(eids[:, 3] == 0)
    assert set(got) == set(want), (sorted(set(got) ^ set(want)), len(got), len(want))
    if count is not None:
        assert len(want) == count
    for t, x in got.items():
        assert np.allclose(x, want[t], rtol=1e-9, atol=1e-11), (t, x, want[t])


def test_closed_edges_lie_on_topology_trisectors():
    sites = datasets.loguniform_sites(20_000, seed=3)
    d = oracle.Diagram(sites)
    ex, eids = d.closed_edges()
    tri = {tuple(sorted(int(v) for v in r)) for r in d.topology()["trisectors"]}
    assert len(eids) > 0
    for r in eids:
        assert tuple(sorted(int(v) for v in r[:3])) in tri
    # each extreme point is tangent to its three spheres
    s = sites.astype(np.float64)
    for m in range(3):
        dist = np.linalg.norm(ex[:, :3] - s[eids[:, m], :3], axis=1) - s[eids[:, m], 3]
        assert np.allclose(dist, ex[:, 3], rtol=1e-7, atol=1e-9)


def test_topology_json_round_trip():
    sites = datasets.uniform_sites(800, seed=32)
    topo = oracle.Diagram(sites).topology()
    text = apo_amd.topology_json(topo)
    assert text.startswith('{\n\t"bisectors": [\n\t\t[ ')
    doc = json.loads(text)
    assert list(doc) == ["bisectors", "trisectors", "quadrisectors"]
    for key in doc:
        assert np.array_equal(np.asarray(doc[key], np.uint32).reshape(topo[key].shape), topo[key])
    empty = apo_amd.topology_json({k: np.zeros((0, w), np.uint32) for k, w in
                                   (("bisectors", 2), ("trisectors", 3), ("quadrisectors", 4))})
    assert json.loads(empty) == {"bisectors": [], "trisectors": [], "quadrisectors": []}
