"""GPU parity against the CPU oracle, through the C ABI.

Integer and index results (Morton codes, permutation, LBVH topology, kNN
lists, cell records) must be bit-exact; FP32 boxes and FP64 vertices too,
because both sides evaluate the same IEEE expressions in the same order.
"""
import numpy as np
import pytest

import oracle
from apo_amd import AlgorithmGPU, datasets

pytestmark = pytest.mark.gpu

CASES = {
    "uniform_small_257": lambda: datasets.uniform_sites(257, seed=11, r_min=0.01, r_max=0.05),
    "uniform_10k": lambda: datasets.uniform_sites(10_000, seed=1),
    "loguniform_20k": lambda: datasets.loguniform_sites(20_000, seed=3),
    "loguniform_100k": lambda: datasets.loguniform_sites(100_000, seed=3),
}

# Above this size the oracle runs its edge searches with the edge radius
# bound, which tests/test_oracle_edge_bound.py pins to the reference search.
BOUND_ABOVE = 30_000


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", params=sorted(CASES))
def built(request):
    name = request.param
    sites = CASES[name]()
    _say(f"[case] {name} n={len(sites)} full pipeline")
    gpu = AlgorithmGPU(sites).build()
    orc = oracle.Diagram(sites, edge_bound=1 if len(sites) > BOUND_ABOVE else 0)
    yield name, sites, gpu, orc
    gpu.close()


def test_lbvh_bit_exact(built):
    name, sites, gpu, orc = built
    ref = oracle.lbvh(sites)
    got = gpu.lbvh()
    s, idx = gpu.sorted_sites()
    assert np.array_equal(got["codes"], ref["codes"]), name
    assert np.array_equal(idx, ref["indices"]), name
    assert np.array_equal(s, sites[ref["indices"]]), name
    for key in ("parent", "left", "right", "sizes"):
        assert np.array_equal(got[key], ref[key]), (name, key)
    assert np.array_equal(got["aabb"].view(np.uint32), ref["aabb"].view(np.uint32)), name
    assert np.array_equal(got["max_radius"], ref["max_radius"]), name
    assert np.array_equal(gpu.artificial_sites().view(np.uint64), oracle.artificial_sites(sites).view(np.uint64))


def test_knn_exact(built):
    name, sites, gpu, orc = built
    assert np.array_equal(gpu.knns(), orc.knns()), name


def test_cells_exact(built):
    name, sites, gpu, orc = built
    gc, gV, gE = gpu.all_cells()
    oc, oV, oE = orc.all_cells()
    assert np.array_equal(gc["status"], oc["status"]), name
    assert np.array_equal(gc["known_radius"].view(np.uint32), oc["known_radius"].view(np.uint32)), name
    assert np.array_equal(gc["vertex_nb"], oc["vertex_nb"]), name
    assert np.array_equal(gc["edge_nb"], oc["edge_nb"]), name
    assert np.array_equal(gV, oV), name
    assert np.array_equal(gE, oE), name
    assert np.array_equal(gc["flags"], oc["flags"]), name


def test_vertices_exact(built):
    name, sites, gpu, orc = built
    g = gpu.to_host()
    ox, oi = orc.vertices()
    assert len(g.vertices) > 0
    assert np.array_equal(g.vertices_id, oi), name
    assert np.array_equal(g.vertices.view(np.uint64), ox.view(np.uint64)), name


def test_empty_sphere(built):
    name, sites, gpu, orc = built
    assert gpu.empty_sphere_violations() == 0, name
    if len(sites) <= 10_000:
        g = gpu.to_host()
        assert oracle.empty_sphere_violations(sites, g.vertices, g.vertices_id) == 0, name


def test_checker_counts_planted_violations(built):
    # move a third of the vertex centres: the tangent-site radius is
    # recomputed from the moved centre, so other sites fall inside; the BVH
    # checker must report exactly the brute-force count of the oracle
    name, sites, gpu, orc = built
    if len(sites) > 10_000:
        pytest.skip("brute-force reference limited to small sets")
    g = gpu.to_host()
    x = g.vertices.copy()
    rng = np.random.default_rng(7)
    pick = rng.random(len(x)) < 0.33
    x[pick, :3] += rng.uniform(-0.03, 0.03, (pick.sum(), 3))
    want = oracle.empty_sphere_violations(sites, x, g.vertices_id)
    got = gpu.empty_sphere_violations(x, g.vertices_id)
    assert want > 0
    assert got == want, (name, got, want)
