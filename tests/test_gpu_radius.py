"""Range path (per-site search radii, apo_proposal_set_radii) on the GPU.

* Parity: the range-query lists, every cell record, status, known radius and
  output vertex equal the oracle's range path bit for bit, with radii drawn
  at random (many lists cut short, many cells left to validation) and with
  radii from the cells' final radii.
* apo_cell_radii: twice a cell's radius, given back as its search radius,
  certifies the cell; with completion the range path then emits the plain
  completed build's vertex rows.
"""
import numpy as np
import pytest

import oracle
from apo_amd import AlgorithmGPU, datasets

pytestmark = pytest.mark.gpu


def _rows(x, ids):
    a = np.concatenate([ids.astype(np.uint64), x.view(np.uint64)], axis=1)
    return a[np.lexsort(a.T[::-1])]


def _compare(gpu, orc):
    gc, gV, gE = gpu.all_cells()
    oc, oV, oE = orc.all_cells()
    for key in ("status", "vertex_nb", "edge_nb", "flags"):
        assert np.array_equal(gc[key], oc[key]), key
    assert np.array_equal(gc["known_radius"].view(np.uint32), oc["known_radius"].view(np.uint32))
    assert np.array_equal(gV, oV) and np.array_equal(gE, oE)
    g = gpu.to_host()
    ox, oi = orc.vertices()
    assert np.array_equal(g.vertices_id, oi)
    assert np.array_equal(g.vertices.view(np.uint64), ox.view(np.uint64))


CASES = {
    "uniform_20k": lambda: datasets.uniform_sites(20_000, seed=11),
    "loguniform_50k": lambda: datasets.loguniform_sites(50_000, seed=12),
    "protein_50k": lambda: datasets.protein_like_sites(50_000, seed=13),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_random_radii_match_oracle(name):
    sites = CASES[name]()
    n = len(sites)
    rng = np.random.default_rng(5)
    # scale of a typical neighbour gap: cube side over the cube root of n
    span = float(np.ptp(sites[:, :3], axis=0).max()) / n ** (1 / 3)
    radii = (rng.uniform(-0.2, 3.0, n) * span).astype(np.float32)
    print(f"[case] radius {name}", flush=True)
    gpu = AlgorithmGPU(sites, radii=radii).build()
    try:
        orc = oracle.Diagram(sites, radii=radii, edge_bound=1 if n > 30_000 else 0)
        _compare(gpu, orc)
        st = gpu.cell_counts()["status"]
        print(f"[radius] {name}: certified {int((st == 1).sum())} of {n}", flush=True)
    finally:
        gpu.close()


def test_cell_radii_certify_and_complete():
    sites = datasets.loguniform_sites(50_000, seed=14)
    print("[case] cell radii 50k", flush=True)
    plain = AlgorithmGPU(sites, complete=4).build()
    try:
        d = plain.to_host()
        ref = _rows(d.vertices, d.vertices_id)
        cr = plain.cell_radii()
        status = plain.cell_counts()["status"]
    finally:
        plain.close()
    buried = cr == np.float32(-np.finfo(np.float32).max)  # input order; status is in sorted order
    assert int(buried.sum()) == int((status == 0xFF).sum())
    radii = (2 * np.where(buried, np.float32(0), cr) * np.float32(1 + 1e-3)).astype(np.float32)
    g = AlgorithmGPU(sites, radii=radii, complete=4).build()
    try:
        dd = g.to_host()
        assert np.array_equal(_rows(dd.vertices, dd.vertices_id), ref)
        cert = int((g.cell_counts()["status"] == 1).sum())
        print(f"[radius] exact radii: certified {cert} of {len(sites)}", flush=True)
        orc = oracle.Diagram(sites, radii=radii, complete=4, edge_bound=1)
        _compare(g, orc)
    finally:
        g.close()
