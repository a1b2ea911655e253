"""Autotuner: the search on a synthetic score (CPU) and, on the GPU, that
every tuned knob leaves the output unchanged."""
import numpy as np
import pytest

from apo_amd import AlgorithmGPU, autotune, datasets


def test_coordinate_descent_finds_separable_minimum():
    target = {"a": 3, "b": 0.5, "c": 8}
    cands = {"a": [1, 2, 3, 4], "b": [0.0, 0.5, 1.0], "c": [2, 4, 8, 16]}
    calls = []

    def score(cfg, knob):
        calls.append(knob)
        return sum((cfg[k] - target[k]) ** 2 for k in target)

    best, log = autotune.coordinate_descent(["a", "b", "c"], cands, {"a": 1, "b": 1.0, "c": 2}, score)
    assert best == target
    assert len(log) == sum(len(v) for v in cands.values())
    assert calls == ["a"] * 4 + ["b"] * 3 + ["c"] * 4


def test_ties_break_deterministically():
    best, _ = autotune.coordinate_descent(["x"], {"x": [4, 2, 8]}, {"x": 4}, lambda cfg, k: 1.0)
    assert best["x"] == 2


def test_wave_candidates_cover_full_occupancy():
    assert autotune._wave_candidates(8) == [1, 2, 3, 4, 6, 8]
    assert autotune._wave_candidates(5) == [1, 2, 3, 4, 5]


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "TUNED_DIR", str(tmp_path))
    rec = {"plain": {"waves_edge": 4, "cover_ratio": 1.0}}
    autotune.save("unit", rec, device="dev x")
    assert autotune.load("unit", device="dev x") == rec
    assert autotune.load("other", device="dev x") is None


@pytest.mark.gpu
def test_knobs_do_not_change_the_output():
    sites = datasets.loguniform_sites(30_000, seed=71)
    ref = AlgorithmGPU(sites).build().to_host()
    variants = [dict(cover_ratio=r) for r in (-1.0,) + autotune.COVER_RATIOS]
    variants += [dict(waves_cells=1, waves_vertex=1, waves_edge=2, waves_bisector=3),
                 dict(waves_per_cu=2), dict(waves_edge=16, waves_vertex=16)]
    for v in variants:
        print(f"[autotune] knobs {v}", flush=True)
        got = AlgorithmGPU(sites, **v).build().to_host()
        assert np.array_equal(got.vertices_id, ref.vertices_id), v
        assert np.array_equal(got.vertices.view(np.uint64), ref.vertices.view(np.uint64)), v


@pytest.mark.gpu
def test_tuner_on_a_small_workload():
    sites = datasets.loguniform_sites(20_000, seed=72)
    res = autotune.tune(sites, "unit_20k", repeats=1, cache=False)
    assert res["same_vertices_as_default"]
    occ = res["occupancy"]
    assert all(o >= 1 for o in occ)
    for knob, o in zip(("waves_cells", "waves_vertex", "waves_edge", "waves_bisector"), occ):
        assert 1 <= res["plain"][knob] <= o
    assert res["plain"]["cover_ratio"] in autotune.COVER_RATIOS
