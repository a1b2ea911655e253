"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces every fixture bit for bit, so a change to
the restatement cannot pass silently.  GPU: the HIP library reproduces the
fixtures through the C ABI without running the oracle on the GPU box.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from apo_amd import AlgorithmGPU

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SINGLE = sorted(p for p in glob.glob(os.path.join(HERE, "*.npz")) if not os.path.basename(p).startswith("batch"))
BATCH = sorted(glob.glob(os.path.join(HERE, "batch_*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("path", SINGLE, ids=os.path.basename)
def test_oracle_reproduces_fixture(path):
    g = _load(path)
    b = oracle.lbvh(g["sites"])
    for key in ("codes", "indices", "parent", "left", "right", "sizes"):
        assert np.array_equal(b[key], g[key]), key
    assert np.array_equal(b["aabb"].view(np.uint32), g["aabb"].view(np.uint32))
    assert np.array_equal(b["max_radius"].view(np.uint32), g["max_radius"].view(np.uint32))
    d = oracle.Diagram(g["sites"])
    x, i = d.vertices()
    assert np.array_equal(i, g["ids"])
    assert np.array_equal(x.view(np.uint64), g["xyzw"].view(np.uint64))
    assert np.array_equal(d.knns(), g["knn"])
    t = d.topology()
    for key in ("bisectors", "trisectors", "quadrisectors"):
        assert np.array_equal(t[key], g[key]), key


@pytest.mark.parametrize("path", BATCH, ids=os.path.basename)
def test_oracle_reproduces_batch_fixture(path):
    g = _load(path)
    off = g["offsets"]
    ids = []
    for m in range(len(off) - 1):
        _, i = oracle.Diagram(g["sites"][off[m]:off[m + 1]]).vertices()
        ids.append(i + off[m])
    assert np.array_equal(np.concatenate(ids), g["ids"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", SINGLE, ids=os.path.basename)
def test_gpu_reproduces_fixture(path):
    g = _load(path)
    print(f"[golden] {os.path.basename(path)} n={len(g['sites'])}", flush=True)
    alg = AlgorithmGPU(g["sites"]).build()
    try:
        got = alg.lbvh()
        _, idx = alg.sorted_sites()
        assert np.array_equal(got["codes"], g["codes"])
        assert np.array_equal(idx, g["indices"])
        for key in ("parent", "left", "right", "sizes"):
            assert np.array_equal(got[key], g[key]), key
        assert np.array_equal(got["aabb"].view(np.uint32), g["aabb"].view(np.uint32))
        assert np.array_equal(alg.knns(), g["knn"])
        d = alg.to_host()
        assert np.array_equal(d.vertices_id, g["ids"])
        assert np.array_equal(d.vertices.view(np.uint64), g["xyzw"].view(np.uint64))
        t = alg.topology()
        for key in ("bisectors", "trisectors", "quadrisectors"):
            assert np.array_equal(t[key], g[key]), key
    finally:
        alg.close()


@pytest.mark.gpu
@pytest.mark.parametrize("path", BATCH, ids=os.path.basename)
def test_gpu_reproduces_batch_fixture(path):
    g = _load(path)
    print(f"[golden] {os.path.basename(path)} molecules={len(g['offsets']) - 1}", flush=True)
    alg = AlgorithmGPU(g["sites"], offsets=g["offsets"]).build()
    try:
        d = alg.to_host()
        assert np.array_equal(d.vertices_id, g["ids"])
        assert np.array_equal(d.vertices.view(np.uint64), g["xyzw"].view(np.uint64))
    finally:
        alg.close()
