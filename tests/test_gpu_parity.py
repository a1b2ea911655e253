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
    "protein_100k": lambda: datasets.protein_like_sites(100_000, seed=2),
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


def _violations_kdtree(sites, xyzw, ids, tol=1e-9):
    """Independent FP64 count of the criterion of app/precision/precision.cpp:
    26-100 through a scipy KD-tree: a vertex is violated when a site other
    than its 4 tangent sites has weighted distance below the largest of
    theirs minus tol."""
    from scipy.spatial import cKDTree
    xyz = sites[:, :3].astype(np.float64)
    r = sites[:, 3].astype(np.float64)
    tree = cKDTree(xyz)
    c = xyzw[:, :3]
    own = np.linalg.norm(xyz[ids] - c[:, None, :], axis=2) - r[ids]
    wmax = own.max(axis=1)
    bad = 0
    reach = wmax + r.max() + 1e-6
    for t, cand in enumerate(tree.query_ball_point(c, reach)):
        cand = np.asarray(cand, np.int64)
        cand = cand[~np.isin(cand, ids[t])]
        if len(cand) and np.any(np.linalg.norm(xyz[cand] - c[t], axis=1) - r[cand] < wmax[t] - tol):
            bad += 1
    return bad


def test_empty_sphere(built):
    # protein-like sets overlap heavily; the reference's method then emits a
    # few vertices (negative radius) that another site intrudes on, which its
    # own precision check counts rather than excludes: the count must match
    # an independent one, and be 0 elsewhere
    name, sites, gpu, orc = built
    g = gpu.to_host()
    got = gpu.empty_sphere_violations()
    want = _violations_kdtree(sites, g.vertices, g.vertices_id)
    print(f"[empty-sphere] {name}: {got} of {len(g.vertices)} vertices", flush=True)
    assert got == want, (name, got, want)
    if not name.startswith("protein"):
        assert got == 0, name
    if len(sites) <= 10_000:
        assert oracle.empty_sphere_violations(sites, g.vertices, g.vertices_id) == want, name


def test_checker_counts_planted_violations(built):
    # move vertex centres: the tangent-site radius is recomputed from the
    # moved centre, so other sites fall inside; the BVH checker must report
    # exactly the brute-force count of the oracle (up to 10k sites) and the
    # independent KD-tree count (every size, on at most 20k moved vertices)
    name, sites, gpu, orc = built
    g = gpu.to_host()
    rng = np.random.default_rng(7)
    pick = rng.random(len(g.vertices)) < min(0.33, 20_000 / len(g.vertices))
    x = g.vertices[pick].copy()
    ids = np.ascontiguousarray(g.vertices_id[pick])
    scale = float(np.ptp(sites[:, :3], axis=0).max())
    x[:, :3] += rng.uniform(-0.03, 0.03, (len(x), 3)) * scale
    got = gpu.empty_sphere_violations(x, ids)
    want = _violations_kdtree(sites, x, ids)
    print(f"[planted] {name}: {got} of {len(x)} moved vertices", flush=True)
    assert want > 0
    assert got == want, (name, got, want)
    if len(sites) <= 10_000:
        assert oracle.empty_sphere_violations(sites, x, ids) == want, name


def test_topology_sets_exact(built):
    # toTopology (algorithm.inl:1732-1821): same entries, same order
    name, sites, gpu, orc = built
    got, want = gpu.topology(), orc.topology()
    for key in ("bisectors", "trisectors", "quadrisectors"):
        assert np.array_equal(got[key], want[key]), (name, key)


def test_cell_topology_csr(built):
    # per-site records in original order and ids: the oracle's cells mapped
    name, sites, gpu, orc = built
    n = len(sites)
    _, perm = orc.sorted_sites()
    oc, oV, oE = orc.all_cells()
    vo, v, eo, e = gpu.cell_topology()
    mapped = np.concatenate([perm, np.arange(n, n + 6, dtype=np.uint32)])
    vstart = np.concatenate([[0], np.cumsum(oc["vertex_nb"].astype(np.int64))])
    estart = np.concatenate([[0], np.cumsum(oc["edge_nb"].astype(np.int64))])
    for i in np.random.default_rng(3).choice(n, size=min(n, 300), replace=False):
        o = perm[i]
        if oc["status"][i] == 0xFF:
            assert vo[o + 1] == vo[o] and eo[o + 1] == eo[o]
            continue
        rv = oV[vstart[i]:vstart[i + 1]].copy()
        re = oE[estart[i]:estart[i + 1]].copy()
        want_v = np.stack([mapped[rv[:, 0]], mapped[rv[:, 1]], mapped[rv[:, 2]], rv[:, 3] & 0xFF], axis=1)
        want_e = np.stack([mapped[re[:, 0]], mapped[re[:, 1]], re[:, 2], np.zeros(len(re), np.uint32)], axis=1)
        assert np.array_equal(v[vo[o]:vo[o + 1]], want_v), (name, i)
        assert np.array_equal(e[eo[o]:eo[o + 1]], want_e), (name, i)


def test_closed_edges_exact(built):
    # toHostFull's closed trisector edges (algorithm.inl:1309-1383): same
    # edges, same order, same FP64 extreme points
    name, sites, gpu, orc = built
    g = gpu.to_host_full()
    ox, oi = orc.closed_edges()
    print(f"[closed-edges] {name}: {len(oi)}", flush=True)
    assert np.array_equal(g.vertices_id, orc.vertices()[1]), name
    assert np.array_equal(g.closed_edges_id, oi), name
    assert np.array_equal(g.closed_edges_min.view(np.uint64), ox.view(np.uint64)), name
