"""Pins the oracle's Morton / sort / Karras / refit / kNN against closed forms,
independent constructions and brute force (CPU only).

The reference (PlathC/apo) ships no tests or golden outputs for this path;
these checks are chosen so that a dropped term, a swapped axis or a wrong
tie-break fails one of them.
"""
import numpy as np
import pytest

import oracle
from apo_amd import datasets


def test_morton_hand_vectors():
    # one bit per axis lands at positions 0 (x), 1 (y), 2 (z) -- lbvh.cu:11-26
    assert oracle.morton64(1, 0, 0) == 1
    assert oracle.morton64(0, 1, 0) == 2
    assert oracle.morton64(0, 0, 1) == 4
    assert oracle.morton64(1023, 1023, 1023) == 0x3FFFFFFF
    # the high 10 bits of each axis form bits 30..59 -- lbvh.cu:28-39
    assert oracle.morton64(1 << 10, 0, 0) == 1 << 30
    assert oracle.morton64(0, 0, 1 << 19) == 1 << (30 + 27 + 2)
    assert oracle.morton64((1 << 20) - 1, (1 << 20) - 1, (1 << 20) - 1) == (1 << 60) - 1


def test_morton_matches_bitwise_interleave():
    rng = np.random.default_rng(5)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 20, 3))
        expect = 0
        for b in range(20):
            expect |= ((x >> b) & 1) << (3 * b)
            expect |= ((y >> b) & 1) << (3 * b + 1)
            expect |= ((z >> b) & 1) << (3 * b + 2)
        assert oracle.morton64(x, y, z) == expect


def test_morton_quantisation_rule():
    # centres at the box corners map to 0 and 2^20-1 on every axis
    s = np.array([[0.1, 0.2, 0.3, 0.05], [0.9, 0.8, 0.7, 0.05], [0.5, 0.5, 0.5, 0.01]], np.float32)
    box, mm = oracle.scene_box(s)
    assert np.allclose(box, [0.05, 0.15, 0.25, 0.95, 0.85, 0.75], atol=1e-7)
    assert mm[0] == np.float64(np.float32(0.01)) and mm[1] == np.float64(np.float32(0.05))
    codes = oracle.morton_codes(s)
    # recompute in FP32 exactly as lbvh.cu:41-48, 287-288 (IEEE division, no FMA)
    scale = np.float32(1) / (box[3:] - box[:3])
    q = ((s[:, :3] - box[:3]) * scale).astype(np.float32)
    u = (q * np.float32(1048575.0)).astype(np.float32).astype(np.uint32)
    for i in range(3):
        assert codes[i] == oracle.morton64(*(int(v) for v in u[i]))


def _karras_reference(codes):
    """Independent top-down build: split each range at the highest differing
    bit of its first and last code, equal codes split by index (the
    reference's 64 + clz(i ^ j) tie-break)."""
    n = len(codes)
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]  # 96-bit key, unique
    left = {}
    right = {}
    counter = [0]

    def build(lo, hi):
        if lo == hi:
            return n - 1 + lo
        a, b = keys[lo], keys[hi]
        bit = (a ^ b).bit_length() - 1
        # first index in (lo, hi] whose key has `bit` set
        split = lo
        while split < hi and not ((keys[split + 1] >> bit) & 1):
            split += 1
        node = None
        return (lo, hi, split, build(lo, split), build(split + 1, hi))

    tree = build(0, n - 1)
    return tree


def _collect(tree, n, out):
    if isinstance(tree, int):
        return (tree - (n - 1), tree - (n - 1))
    lo, hi, split, l, r = tree
    out.append((lo, hi, split))
    _collect(l, n, out)
    _collect(r, n, out)
    return (lo, hi)


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (17, 3), (64, 4), (257, 5), (1000, 6)])
def test_karras_topology_matches_topdown_split(n, seed):
    s = datasets.uniform_sites(n, seed=seed)
    b = oracle.lbvh(s)
    codes = b["codes"]
    assert np.all(codes[1:] >= codes[:-1])
    ranges = []
    _collect(_karras_reference(codes), n, ranges)
    # every internal node of the oracle covers one top-down range and splits it identically
    def leaf_range(node):
        if node >= n - 1:
            return (node - (n - 1), node - (n - 1))
        lo = leaf_range(b["left"][node])[0]
        hi = leaf_range(b["right"][node])[1]
        return (lo, hi)
    got = []
    for i in range(n - 1):
        lo, _ = leaf_range(b["left"][i])
        _, hi = leaf_range(b["right"][i])
        split = leaf_range(b["left"][i])[1]
        got.append((lo, hi, split))
    assert sorted(got) == sorted(ranges)
    assert b["parent"][0] == -1
    for node in range(1, 2 * n - 1):
        p = b["parent"][node]
        assert b["left"][p] == node or b["right"][p] == node


def test_sort_is_stable_with_duplicate_codes():
    # duplicated centres give equal codes; ties must keep input order
    base = datasets.uniform_sites(300, seed=9)
    s = np.concatenate([base, base[::-1], base[:50]])
    b = oracle.lbvh(s)
    codes = b["codes"]
    idx = b["indices"]
    assert np.all(codes[1:] >= codes[:-1])
    expect = np.argsort(oracle.morton_codes(s), kind="stable")
    assert np.array_equal(idx, expect)


def test_refit_matches_brute_force():
    n = 500
    s = datasets.loguniform_sites(n, seed=4)
    b = oracle.lbvh(s)
    idx = b["indices"]
    sd = s.astype(np.float64)[idx]
    leaf_min = (sd[:, :3] - sd[:, 3:4]).astype(np.float32)
    leaf_max = (sd[:, :3] + sd[:, 3:4]).astype(np.float32)

    def leaves(node):
        if node >= n - 1:
            return [node - (n - 1)]
        return leaves(b["left"][node]) + leaves(b["right"][node])

    for node in range(2 * n - 1):
        L = leaves(node)
        assert b["sizes"][node] == len(L)
        assert np.array_equal(b["aabb"][node, :3], leaf_min[L].min(0))
        assert np.array_equal(b["aabb"][node, 3:], leaf_max[L].max(0))
        assert b["max_radius"][node] == s[idx[L], 3].max()


@pytest.mark.parametrize("gen,n", [(datasets.uniform_sites, 2000), (datasets.loguniform_sites, 2000)])
def test_knn_matches_brute_force(gen, n):
    s = gen(n)
    b = oracle.lbvh(s)
    sorted_sites = s[b["indices"]]
    ids, dist, depth = oracle.knn(sorted_sites, b, 16)
    bids, bdist = oracle.knn_brute(sorted_sites, 16)
    assert depth <= 32
    # identical gaps; ids identical wherever the gap is not tied
    assert np.array_equal(dist, bdist)
    untied = np.ones_like(dist, bool)
    untied[:, 1:] &= dist[:, 1:] != dist[:, :-1]
    untied[:, :-1] &= dist[:, :-1] != dist[:, 1:]
    assert np.array_equal(ids[untied], bids[untied])
    assert np.all(ids != np.arange(n)[:, None])
