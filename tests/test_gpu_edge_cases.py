"""Edge cases of the input on the GPU, each against the oracle.

Tiny site counts (2, 3, 5, 17: at and around the 16-neighbour list),
exact duplicates, concentric spheres, one sphere holding nearly all others,
coplanar centres, zero radii, large and negative coordinates.  Every case
must build without a device guard and give the oracle's vertices and
closed edges bit for bit.  Invalid arguments must fail with a status.
"""
import numpy as np
import pytest

import oracle
from apo_amd import AlgorithmGPU, ApoError, datasets

pytestmark = pytest.mark.gpu


def _rand(n, seed, rmin=0.01, rmax=0.05):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 4), np.float32)
    s[:, :3] = rng.random((n, 3))
    s[:, 3] = rng.uniform(rmin, rmax, n)
    return s


def _duplicates():
    s = _rand(200, 1)
    return np.concatenate([s, s[::10]]).astype(np.float32)


def _concentric():
    s = _rand(150, 2)
    inner = s[:10].copy()
    inner[:, 3] *= 0.5
    return np.concatenate([s, inner]).astype(np.float32)


def _giant():
    s = _rand(300, 3, 0.005, 0.01)
    return np.concatenate([s, [[0.5, 0.5, 0.5, 0.45]]]).astype(np.float32)


def _coplanar():
    s = _rand(300, 4)
    s[:, 2] = 0.5
    return s


def _zero_radii():
    s = _rand(300, 5)
    s[:, 3] = 0.0
    return s


def _far_negative():
    s = _rand(400, 6)
    s[:, :3] = s[:, :3] * 1000.0 - 5.0e4
    s[:, 3] *= 1000.0
    return s


CASES = {
    "n2": lambda: _rand(2, 10),
    "n3": lambda: _rand(3, 11),
    "n5": lambda: _rand(5, 12),
    "n17": lambda: _rand(17, 13),
    "duplicates": _duplicates,
    "concentric": _concentric,
    "giant": _giant,
    "coplanar": _coplanar,
    "zero_radii": _zero_radii,
    "far_negative": _far_negative,
    "protein_like_small": lambda: datasets.protein_like_sites(600, seed=21),
    "crowded_shell_60": lambda: datasets.crowded_shell_sites(60),
    "crowded_shell_150": lambda: datasets.crowded_shell_sites(150),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_edge_case_matches_oracle(name):
    sites = CASES[name]()
    print(f"[case] edge case {name} n={len(sites)}", flush=True)
    gpu = AlgorithmGPU(sites).build()
    try:
        g = gpu.to_host_full()
    finally:
        gpu.close()
    orc = oracle.Diagram(sites)
    ox, oi = orc.vertices()
    ex, ei = orc.closed_edges()
    assert np.array_equal(g.vertices_id, oi), name
    assert np.array_equal(g.vertices.view(np.uint64), ox.view(np.uint64)), name
    assert np.array_equal(g.closed_edges_id, ei), name
    assert np.array_equal(g.closed_edges_min.view(np.uint64), ex.view(np.uint64)), name


@pytest.mark.parametrize("shell", [60, 150])
def test_crowded_cell_records_match_oracle(shell):
    """A cell of about 200 edges (four 64-edge steps of every per-edge scan)
    and one past the edge capacity (flagged and truncated): every cell
    record, count and flag equals the oracle's."""
    sites = datasets.crowded_shell_sites(shell)
    print(f"[case] crowded shell {shell} n={len(sites)} cell records", flush=True)
    gpu = AlgorithmGPU(sites).build()
    try:
        gc, gV, gE = gpu.all_cells()
    finally:
        gpu.close()
    oc, oV, oE = oracle.Diagram(sites).all_cells()
    for key in ("status", "vertex_nb", "edge_nb", "flags"):
        assert np.array_equal(gc[key], oc[key]), key
    assert np.array_equal(gc["known_radius"].view(np.uint32), oc["known_radius"].view(np.uint32))
    assert np.array_equal(gV, oV)
    assert np.array_equal(gE, oE)
    assert int(gc["edge_nb"].max()) > 192


def test_invalid_arguments_fail_with_a_status():
    with pytest.raises(ApoError):
        AlgorithmGPU(_rand(1, 1)).build()
    gpu = AlgorithmGPU(_rand(50, 2), classes=np.zeros(49, np.uint8))
    try:
        with pytest.raises(ApoError):
            gpu.build()
    finally:
        gpu.close()
    with pytest.raises(ApoError):
        AlgorithmGPU(_rand(50, 2), classes=np.full(50, 9, np.uint8))
    for k in (0, 65):
        with pytest.raises(ApoError):
            AlgorithmGPU(_rand(50, 2), k=k)
    # the learned proposal runs one kNN pass
    gpu = AlgorithmGPU(_rand(50, 2), classes=np.zeros(50, np.uint8), knn_validation_step=2)
    try:
        with pytest.raises(ApoError):
            gpu.build()
    finally:
        gpu.close()
