"""The range path of the oracle (per-site search radii), pinned on the CPU.

* The lists equal a brute-force range query: the sites whose FP32 gap is
  below max(radius, the float above -r_i), nearest first, at most 64.
* The known radius a cell takes from a list that is not full is sound: no
  site whose exact (float64) gap lies below it is missing from the list.
* With radii twice the cells' final radius (plus a margin) the diagram
  equals the plain build's, and on small sets the brute-force diagram.
"""
import numpy as np
import pytest

import oracle
from apo_amd import datasets

NONE = 0xFFFFFFFF


def _gaps32(s):
    c = s[:, :3]
    d = c[:, None, :] - c[None, :, :]
    sq = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return np.sqrt(sq) - (s[:, 3][:, None] + s[:, 3][None, :])


def _top(b, r):
    neg = np.float32(-r)
    above = np.nextafter(neg, np.float32(np.inf))
    return max(np.float32(b), above)


def _radius_lists(seed, n=1500):
    sites = datasets.uniform_sites(n, seed=seed, r_min=0.002, r_max=0.03)
    rng = np.random.default_rng(seed)
    radii = rng.uniform(-0.01, 0.12, n).astype(np.float32)
    d = oracle.Diagram(sites, radii=radii, stage_mask=1)
    s, perm = d.sorted_sites()
    out = np.zeros((n, 64), np.uint32)
    oracle.lib().orc_get_knns(d._h, out.ctypes.data_as(oracle.u32p))
    return s, radii[perm], out


@pytest.mark.parametrize("seed", [1, 2])
def test_range_lists_equal_brute_force(seed):
    s, r, lists = _radius_lists(seed)
    g = _gaps32(s)
    np.fill_diagonal(g, np.inf)
    for i in range(len(s)):
        top = _top(r[i], s[i, 3])
        cand = np.nonzero(g[i] < top)[0]
        cand = cand[np.argsort(g[i, cand], kind="stable")][:64]
        want = np.full(64, NONE, np.uint32)
        want[:len(cand)] = cand
        assert np.array_equal(lists[i], want), i


def test_range_known_radius_is_sound():
    s, r, lists = _radius_lists(3)
    S = s.astype(np.float64)
    g64 = np.linalg.norm(S[:, None, :3] - S[None, :, :3], axis=2) - S[:, 3][:, None] - S[:, 3][None, :]
    np.fill_diagonal(g64, np.inf)
    checked = 0
    for i in range(len(s)):
        if lists[i, -1] != NONE:
            continue  # a full list takes no known radius from its range
        top = float(_top(r[i], s[i, 3]))
        scale = abs(top) + abs(S[i, 0]) + abs(S[i, 1]) + abs(S[i, 2]) + abs(S[i, 3])
        known = top - scale * 2.0 ** -18
        inside = set(np.nonzero(g64[i] < known)[0].tolist())
        assert inside <= set(lists[i][lists[i] != NONE].tolist()), i
        checked += 1
    assert checked > 500


def _rows(d):
    x, ids = d.vertices()
    a = np.concatenate([ids.astype(np.uint64), x.view(np.uint64)], axis=1)
    return a[np.lexsort(a.T[::-1])]


def _exact_radii(sites):
    """Twice the final cell radius of every site, plus a margin: the
    oracle's cells of a completed build, radius from its vertex records is
    not exported, so the GPU path's apo_cell_radii is mirrored by the plain
    diagram's largest vertex radius per site (a lower bound: cells then need
    validation where trisector extremes are larger)."""
    d = oracle.Diagram(sites, complete=4)
    x, ids = d.vertices()
    w = np.full(len(sites), -np.inf)
    for col in range(4):
        np.maximum.at(w, ids[:, col], x[:, 3])
    return np.where(np.isfinite(w), 2.0 * w * (1 + 1e-3) + 1e-6, 3e38).astype(np.float32), d


def test_range_path_gives_the_completed_diagram():
    sites = datasets.loguniform_sites(20_000, seed=7)
    radii, plain = _exact_radii(sites)
    ranged = oracle.Diagram(sites, radii=radii, complete=4)
    assert np.array_equal(_rows(ranged), _rows(plain))
    c = ranged.cell_counts()
    assert int((c["status"] == 1).sum()) > 0


@pytest.mark.parametrize("n,seed", [(16, 2), (32, 4), (40, 5)])
def test_range_path_equals_brute_force(n, seed):
    sites = datasets.uniform_sites(n, seed=100 + seed, r_min=0.01, r_max=0.06)
    rng = np.random.default_rng(seed)
    radii = rng.uniform(-0.05, 0.5, n).astype(np.float32)
    _, ids = oracle.Diagram(sites, radii=radii).vertices()
    a = np.sort(ids, axis=1)
    a = a[np.lexsort(a.T[::-1])]
    b = oracle.bruteforce_vertices(sites)
    b = np.sort(b, axis=1)
    b = b[np.lexsort(b.T[::-1])]
    assert np.array_equal(a, b)
