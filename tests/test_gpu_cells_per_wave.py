"""The cell pass with two cells per wave (config.cells_per_wave = 2, half a
wave per cell) stores the records of the one-cell-per-wave pass, which the
parity tests pin to the oracle: every mode of the pass (first pass, resumed
kNN passes, learned classes up to 64 neighbours, ranged lists, batches,
capacity overflow, completion) compared cell for cell and vertex for
vertex.  Reference: apo/include/apo/gpu/algorithm.inl:101-206 (getCells)."""
import numpy as np
import pytest

from apo_amd import AlgorithmGPU, ApoError, datasets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[cells_per_wave] {msg}", flush=True)


def _same(a, b, what):
    ha, hb = a.to_host(), b.to_host()
    assert np.array_equal(ha.vertices_id, hb.vertices_id), what
    assert np.array_equal(ha.vertices.view(np.uint64), hb.vertices.view(np.uint64)), what
    (ca, va, ea), (cb, vb, eb) = a.all_cells(), b.all_cells()
    for key in ca:
        assert np.array_equal(ca[key].view(np.uint8), cb[key].view(np.uint8)), (what, key)
    assert np.array_equal(va, vb) and np.array_equal(ea, eb), what


CASES = {
    "uniform_20k": lambda: (datasets.uniform_sites(20_000, seed=81), {}),
    "loguniform_30k": lambda: (datasets.loguniform_sites(30_000, seed=82), {}),
    "protein_20k": lambda: (datasets.protein_like_sites(20_000, seed=83), {}),
    "k4_passes3": lambda: (datasets.uniform_sites(10_000, seed=84), {"k": 4, "knn_validation_step": 3}),
    "k64": lambda: (datasets.loguniform_sites(10_000, seed=85), {"k": 64}),
    "complete": lambda: (datasets.loguniform_sites(20_000, seed=86), {"complete": 2}),
    "crowded": lambda: (datasets.crowded_shell_sites(210, seed=87), {}),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_pair_pass_equals_single_pass(name):
    sites, kw = CASES[name]()
    _say(f"{name}: {len(sites)} sites {kw}")
    one = AlgorithmGPU(sites, cells_per_wave=1, **kw).build()
    two = AlgorithmGPU(sites, cells_per_wave=2, **kw).build()
    try:
        _same(one, two, name)
        s1, s2 = one.stats(), two.stats()
        assert s1["uncertified"] == s2["uncertified"] and s1["flagged_cells"] == s2["flagged_cells"]
    finally:
        one.close()
        two.close()


def test_pair_pass_learned_ranged_and_batch():
    sites = datasets.loguniform_sites(20_000, seed=88)
    rng = np.random.default_rng(88)
    classes = rng.integers(0, 8, len(sites)).astype(np.uint8)
    _say("learned classes, 20k sites")
    kw = dict(proposal_max_k=64, proposal_fallback_k=48)
    one = AlgorithmGPU(sites, classes=classes, cells_per_wave=1, **kw).build()
    two = AlgorithmGPU(sites, classes=classes, cells_per_wave=2, **kw).build()
    try:
        _same(one, two, "classes")
    finally:
        one.close()
        two.close()
    sites = datasets.uniform_sites(20_000, seed=90)
    n = len(sites)
    span = float(np.ptp(sites[:, :3], axis=0).max()) / n ** (1 / 3)
    radii = (np.random.default_rng(90).uniform(-0.2, 3.0, n) * span).astype(np.float32)
    _say("ranged lists, 20k sites")
    one = AlgorithmGPU(sites, radii=radii, cells_per_wave=1).build()
    two = AlgorithmGPU(sites, radii=radii, cells_per_wave=2).build()
    try:
        _same(one, two, "ranged")
    finally:
        one.close()
        two.close()
    mols, offsets = datasets.batched_molecules(count=12, seed=89, lo=300, hi=3000)
    _say(f"batch of 12 molecules, {len(mols)} atoms")
    one = AlgorithmGPU(mols, offsets=offsets, cells_per_wave=1).build()
    two = AlgorithmGPU(mols, offsets=offsets, cells_per_wave=2).build()
    try:
        _same(one, two, "batch")
    finally:
        one.close()
        two.close()


def test_pair_pass_rejects_other_values():
    with pytest.raises(ApoError):
        AlgorithmGPU(datasets.uniform_sites(100, seed=1), cells_per_wave=3)
