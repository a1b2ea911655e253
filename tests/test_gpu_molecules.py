"""Diagrams of the reference's sample proteins on the GPU.

The reference's topology and precision apps load these MMTF files and
joggle them by 1e-3 before a build (app/topology/topology_apo.cu,
app/precision/precision_apo.cpp).  The GPU build of each must equal the
oracle's, bit for bit: vertices, closed edges and topology sets; and the
interval validator must give the oracle's counts.
"""
import os

import numpy as np
import pytest

import oracle
from apo_amd import AlgorithmGPU, molecules

pytestmark = pytest.mark.gpu

MMTF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mmtf")


@pytest.mark.parametrize("name", ["5ZCK", "1AGA", "3DIK", "101M", "1A3F"])
def test_protein_matches_oracle(name):
    sites = molecules.joggle(molecules.load_mmtf(os.path.join(MMTF, f"{name}.mmtf")), 1e-3)
    print(f"[case] protein {name} n={len(sites)}", flush=True)
    gpu = AlgorithmGPU(sites).build()
    try:
        g = gpu.to_host_full()
        topo = gpu.topology()
        prec = gpu.precision()
    finally:
        gpu.close()
    orc = oracle.Diagram(sites)
    ox, oi = orc.vertices()
    ex, ei = orc.closed_edges()
    assert len(oi) > 0
    assert np.array_equal(g.vertices_id, oi)
    assert np.array_equal(g.vertices.view(np.uint64), ox.view(np.uint64))
    assert np.array_equal(g.closed_edges_id, ei)
    assert np.array_equal(g.closed_edges_min.view(np.uint64), ex.view(np.uint64))
    want = orc.topology()
    for key in ("bisectors", "trisectors", "quadrisectors"):
        assert np.array_equal(topo[key], want[key]), key
    assert (prec["bad_quadrisectors"], prec["uncertain_quadrisectors"]) == \
        oracle.precision_counts(sites, ox, oi, arity=4)
    assert (prec["bad_trisectors"], prec["uncertain_trisectors"]) == \
        oracle.precision_counts(sites, ex, ei, arity=3)
