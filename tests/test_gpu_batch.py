"""Batched builds of independent molecules against the oracle run on each
molecule alone, through the C ABI.

A batch is one launch per stage over all molecules (offset-indexed LBVH,
persistent work-queue kernels); molecule m's results must be exactly the
oracle's diagram of its sites, with site ids shifted by offsets[m], LBVH
nodes renumbered into the batch layout (internal nodes of molecule m from
offsets[m] - m, leaves after the n - m internal nodes) and artificial sites
renumbered to n + 6m + a.  The full-size batch (512 molecules, 5.3 M atoms)
is checked in the benchmark with the same properties.
"""
import numpy as np
import pytest

import oracle
from apo_amd import AlgorithmGPU, ApoError, datasets

pytestmark = pytest.mark.gpu

NONE = np.uint32(0xFFFFFFFF)


def _tiny_mix():
    # the smallest molecules (2 and 3 sites: no vertex at all), then a few more
    sizes = [2, 3, 5, 17, 300, 64]
    parts = [datasets.protein_like_sites(n, seed=9, seed_tag=t) for t, n in enumerate(sizes)]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    return np.concatenate(parts), offsets


CASES = {
    # 12 molecules of 200..4000 atoms: one molecule digit in the key (8 sort passes)
    "proteins_12": lambda: datasets.batched_molecules(count=12, seed=4, lo=200, hi=4000),
    # 40 molecules: a ninth sort pass over the molecule index
    "proteins_40": lambda: datasets.batched_molecules(count=40, seed=5, lo=60, hi=900),
    "tiny_mix": _tiny_mix,
}


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", params=sorted(CASES))
def batch(request):
    name = request.param
    sites, offsets = CASES[name]()
    _say(f"[case] {name} molecules={len(offsets) - 1} n={len(sites)} batched build")
    gpu = AlgorithmGPU(sites, offsets=offsets).build()
    mols = [oracle.Diagram(sites[offsets[m]:offsets[m + 1]]) for m in range(len(offsets) - 1)]
    yield name, sites, offsets, gpu, mols
    gpu.close()


def _map_ids(a, base, n_local, n_total, m):
    """Oracle site ids of molecule m (local, artificial n_local + a) to batch ids."""
    a = a.astype(np.int64)
    out = np.where(a < n_local, a + base, n_total + 6 * m + (a - n_local))
    return out.astype(np.uint32)


def test_vertices_exact(batch):
    name, sites, offsets, gpu, mols = batch
    g = gpu.to_host()
    xs, ids, counts = [], [], [0]
    for m, d in enumerate(mols):
        x, i = d.vertices()
        xs.append(x)
        ids.append(i + offsets[m])
        counts.append(counts[-1] + len(x))
    assert np.array_equal(g.vertices_id, np.concatenate(ids)), name
    assert np.array_equal(g.vertices.view(np.uint64), np.concatenate(xs).view(np.uint64)), name
    assert np.array_equal(gpu.batch_vertex_offsets(), np.asarray(counts, np.uint64)), name


def test_lbvh_bit_exact(batch):
    name, sites, offsets, gpu, mols = batch
    n, M = len(sites), len(offsets) - 1
    internal = n - M
    got = gpu.lbvh()
    srt, idx = gpu.sorted_sites()
    for m in range(M):
        a, b = int(offsets[m]), int(offsets[m + 1])
        nm = b - a
        ref = oracle.lbvh(sites[a:b])
        assert np.array_equal(got["codes"][a:b], ref["codes"]), (name, m)
        assert np.array_equal(idx[a:b], ref["indices"] + a), (name, m)
        # node renumbering: local internal k -> a - m + k, local leaf nm-1+j -> internal + a + j
        loc = np.arange(2 * nm - 1)
        glob = np.where(loc < nm - 1, a - m + loc, internal + a + (loc - (nm - 1)))

        def remap(v):
            v = np.asarray(v, np.int64)
            return np.where(v < 0, -1, np.where(v < nm - 1, a - m + v, internal + a + (v - (nm - 1))))

        assert np.array_equal(got["parent"][glob], remap(ref["parent"])), (name, m)
        inner = glob[: nm - 1]
        assert np.array_equal(got["left"][inner], remap(ref["left"][: nm - 1])), (name, m)
        assert np.array_equal(got["right"][inner], remap(ref["right"][: nm - 1])), (name, m)
        assert np.array_equal(got["sizes"][glob], ref["sizes"]), (name, m)
        assert np.array_equal(got["aabb"][glob].view(np.uint32), ref["aabb"].view(np.uint32)), (name, m)
        assert np.array_equal(got["max_radius"][glob], ref["max_radius"]), (name, m)


def test_knn_and_cells_exact(batch):
    name, sites, offsets, gpu, mols = batch
    n, M = len(sites), len(offsets) - 1
    knn = gpu.knns()
    gc, gV, gE = gpu.all_cells()
    vstart = np.concatenate([[0], np.cumsum(gc["vertex_nb"].astype(np.int64))])
    estart = np.concatenate([[0], np.cumsum(gc["edge_nb"].astype(np.int64))])
    for m, d in enumerate(mols):
        a, b = int(offsets[m]), int(offsets[m + 1])
        nm = b - a
        ok = d.knns()
        want = np.where(ok == NONE, NONE, ok + np.uint32(a))
        assert np.array_equal(knn[a:b], want), (name, m)
        oc, oV, oE = d.all_cells()
        for key in ("status", "vertex_nb", "edge_nb", "flags"):
            assert np.array_equal(gc[key][a:b], oc[key]), (name, m, key)
        assert np.array_equal(gc["known_radius"][a:b].view(np.uint32), oc["known_radius"].view(np.uint32)), (name, m)
        V = gV[vstart[a]:vstart[b]]
        E = gE[estart[a]:estart[b]]
        wantV = oV.copy()
        wantE = oE.copy()
        for c in range(3):
            wantV[:, c] = _map_ids(oV[:, c], a, nm, n, m)
        for c in range(2):
            wantE[:, c] = _map_ids(oE[:, c], a, nm, n, m)
        assert np.array_equal(V, wantV), (name, m)
        assert np.array_equal(E, wantE), (name, m)


def test_empty_sphere_per_molecule(batch):
    # the LBVH check walks each vertex's own molecule tree; its count is the
    # sum of the oracle's brute-force counts over the molecules
    name, sites, offsets, gpu, mols = batch
    g = gpu.to_host()
    got = gpu.empty_sphere_violations()
    vo = gpu.batch_vertex_offsets()
    want = 0
    for m in range(len(offsets) - 1):
        a, b = int(offsets[m]), int(offsets[m + 1])
        x = g.vertices[vo[m]:vo[m + 1]]
        i = g.vertices_id[vo[m]:vo[m + 1]] - np.uint32(a)
        want += oracle.empty_sphere_violations(sites[a:b], x, i) if len(x) else 0
    print(f"[empty-sphere] {name}: {got} of {len(g.vertices)} vertices", flush=True)
    assert got == want, (name, got, want)


def test_graph_replay_matches_eager(batch):
    # the build captured in a HIP graph: first call captures, the next two replay
    name, sites, offsets, gpu, mols = batch
    ref = gpu.to_host()
    alg = AlgorithmGPU(sites, offsets=offsets, graph=True)
    try:
        for rep in range(3):
            _say(f"[graph] {name} replay {rep}")
            got = alg.build().to_host()
            assert np.array_equal(got.vertices_id, ref.vertices_id), (name, rep)
            assert np.array_equal(got.vertices.view(np.uint64), ref.vertices.view(np.uint64)), (name, rep)
            assert np.array_equal(alg.batch_vertex_offsets(), gpu.batch_vertex_offsets()), (name, rep)
            assert alg.stats()["ms_total"] > 0.0
    finally:
        alg.close()


def test_single_molecule_batch_is_a_plain_build():
    sites = datasets.uniform_sites(3000, seed=21)
    a = AlgorithmGPU(sites).build().to_host()
    b = AlgorithmGPU(sites, offsets=np.array([0, len(sites)], np.uint32)).build().to_host()
    assert np.array_equal(a.vertices_id, b.vertices_id)
    assert np.array_equal(a.vertices.view(np.uint64), b.vertices.view(np.uint64))


def test_graph_single_build_and_reshape():
    # plain build in graph mode: replays of one shape, then a new size recaptures
    s1 = datasets.loguniform_sites(20_000, seed=31)
    s2 = datasets.loguniform_sites(12_000, seed=32)
    ref1 = AlgorithmGPU(s1).build().to_host()
    ref2 = AlgorithmGPU(s2).build().to_host()
    alg = AlgorithmGPU(s1, graph=True)
    try:
        for rep in range(2):
            got = alg.build().to_host()
            assert np.array_equal(got.vertices_id, ref1.vertices_id), rep
        alg.sites = np.ascontiguousarray(s2)
        got = alg.build().to_host()
        assert np.array_equal(got.vertices_id, ref2.vertices_id)
        assert np.array_equal(got.vertices.view(np.uint64), ref2.vertices.view(np.uint64))
    finally:
        alg.close()


@pytest.mark.parametrize("offsets", [[0, 1, 10], [1, 10], [0, 5, 9], [0, 4, 3, 10]])
def test_invalid_offsets_rejected(offsets):
    sites = datasets.uniform_sites(10, seed=2)
    alg = AlgorithmGPU(sites, offsets=np.asarray(offsets, np.uint32))
    try:
        with pytest.raises((ApoError, ValueError)):
            alg.build()
    finally:
        alg.close()
