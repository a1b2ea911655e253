"""Pins the oracle's interval validator (app/precision/precision.cpp:14-163)
without reusing its arithmetic.

* Each directed-rounding operation against exact rational arithmetic: the
  lower and upper results enclose the exact value and are the two adjacent
  doubles around it (or both equal to it when it is representable).
* The interval weighted distance encloses the exact |p - c| - r.
* Counts bracketed by plain FP64 counts: a point bad under intervals is
  violated in FP64 with any tolerance above the rounding error, and a point
  violated by more than that margin is bad.
* Closed forms: sites exactly as far as the point's own sites make it
  uncertain, never bad; a site closer by far more than the rounding error
  makes it bad.
* On valid diagrams (log-uniform sets) no vertex and no closed edge is bad.
"""
from decimal import Decimal, getcontext
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from apo_amd import datasets

getcontext().prec = 80


def _adjacent(lo, hi, exact):
    if Fr(lo) == exact:
        return lo == hi
    return Fr(lo) < exact < Fr(hi) and hi == np.nextafter(lo, np.inf)


@pytest.mark.parametrize("seed", range(4))
def test_directed_rounding_against_exact_arithmetic(seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(400) * 10.0 ** rng.integers(-8, 8, 400)
    b = rng.standard_normal(400) * 10.0 ** rng.integers(-8, 8, 400)
    # exactly representable results must come back unchanged
    a[:20] = np.round(a[:20] * 64) / 64
    b[:20] = np.round(b[:20] * 64) / 64
    for x, y in zip(a, b):
        lo, hi = oracle.round_dir("add", x, y)
        assert _adjacent(lo, hi, Fr(x) + Fr(y)), (x, y, lo, hi)
        lo, hi = oracle.round_dir("mul", x, y)
        assert _adjacent(lo, hi, Fr(x) * Fr(y)), (x, y, lo, hi)
        v = abs(x)
        lo, hi = oracle.round_dir("sqrt", v)
        assert Fr(lo) ** 2 <= Fr(v) <= Fr(hi) ** 2, (v, lo, hi)
        assert lo == hi if Fr(lo) ** 2 == Fr(v) else hi == np.nextafter(lo, np.inf)
    for v in (4.0, 2.25, 2.0 ** -40, 0.0):
        assert oracle.round_dir("sqrt", v)[0] == oracle.round_dir("sqrt", v)[1]


def test_interval_distance_encloses_exact_value():
    rng = np.random.default_rng(5)
    for _ in range(300):
        s = rng.random(4) * np.array([1, 1, 1, 0.01])
        p = rng.random(3)
        lo, hi = oracle.interval_distance(s, p)
        S = sum((Fr(s[i]) - Fr(p[i])) ** 2 for i in range(3))
        exact = Decimal(S.numerator) / Decimal(S.denominator)
        d = exact.sqrt() - Decimal(Fr(s[3]).numerator) / Decimal(Fr(s[3]).denominator)
        assert Decimal(lo) <= d <= Decimal(hi), (lo, d, hi)
        assert hi - lo <= 8 * np.spacing(abs(hi) + 1.0)


def _fp64_count(sites, xyzw, ids, tol):
    return oracle.empty_sphere_violations(sites, xyzw, ids, tol=tol)


@pytest.mark.parametrize("gen", ["uniform", "loguniform"])
def test_counts_bracketed_by_fp64(gen):
    sites = (datasets.uniform_sites(2500, seed=61) if gen == "uniform"
             else datasets.loguniform_sites(2500, seed=62))
    x, ids = oracle.Diagram(sites).vertices()
    bad, unc = oracle.precision_counts(sites, x, ids)
    assert bad == 0
    assert unc <= len(x) // 1000
    moved = x.copy()
    rng = np.random.default_rng(9)
    pick = rng.random(len(x)) < 0.3
    moved[pick, :3] += rng.uniform(-0.02, 0.02, (pick.sum(), 3))
    bad, unc = oracle.precision_counts(sites, moved, ids)
    lo_count = _fp64_count(sites, moved, ids, 1e-12)
    hi_count = _fp64_count(sites, moved, ids, -1e-12)
    assert 0 < lo_count <= bad <= hi_count, (lo_count, bad, hi_count)
    assert bad + unc >= hi_count


def _cube_case():
    """8 zero-radius sites on the corners of a cube around (5, 5, 5), three
    more probes, and 200 filler sites far away; points at the cube centre."""
    rng = np.random.default_rng(4)
    corners = np.array([[5 + dx, 5 + dy, 5 + dz, 0.0] for dx in (-1, 1) for dy in (-1, 1) for dz in (-1, 1)])
    h = np.sqrt(3.0)
    probes = np.array([[5 + h, 5, 5, 1e-6],       # closer than the corners by 1e-6: bad
                       [5, 5 + h, 5, 0.0],        # FP32 rounds it 9e-8 farther than the corners: no effect
                       [5, 5, 5 + 3.0, 0.0]])      # farther: no effect
    filler = np.concatenate([rng.random((200, 3)), rng.uniform(0.001, 0.002, (200, 1))], axis=1)
    sites = np.concatenate([corners, probes, filler]).astype(np.float32)
    p = np.array([5.0, 5.0, 5.0, 0.0])
    return sites, p


def test_cospherical_sites_are_uncertain_not_bad():
    sites, p = _cube_case()
    ids_only_corners = np.array([[0, 3, 5, 6]], np.uint32)
    # the other four corners tie with the own sites
    probe_bad = sites.copy()
    probe_bad[8, 3] = 0.0  # drop the bad probe
    assert oracle.precision_counts(probe_bad, p[None], ids_only_corners) == (0, 1)
    # with the 1e-6 probe in place the point is bad
    assert oracle.precision_counts(sites, p[None], ids_only_corners) == (1, 0)
    # moved clearly inside the own sites' sphere: neither
    far = sites.copy()
    far[[1, 2, 4, 7, 8, 9], :3] += 0.5 * np.sign(far[[1, 2, 4, 7, 8, 9], :3] - 5.0)
    assert oracle.precision_counts(far, p[None], ids_only_corners) == (0, 0)


def test_closed_edges_of_valid_diagrams_are_not_bad():
    sites = datasets.loguniform_sites(5000, seed=63)
    d = oracle.Diagram(sites)
    ex, eids = d.closed_edges()
    assert len(ex) > 0
    assert oracle.precision_counts(sites, ex, eids, arity=3)[0] == 0
    # the same points checked as if a fourth site were part of them change nothing
    ids4 = eids.copy()
    ids4[:, 3] = eids[:, 2]
    assert oracle.precision_counts(sites, ex, ids4, arity=4) == oracle.precision_counts(sites, ex, eids, arity=3)
