"""Pins the oracle's FP64 geometry and cell construction with checks that do
not reuse its formulas.

* Apollonius vertices (quadrisector, math.cuh:703-780) and trisector extreme
  points (math.cuh:636-701) against an exact solve: the tangency equations
  |x - p_m| = r_m + w are reduced to a linear system in x with w as a
  parameter, solved in exact rationals (fractions.Fraction), and the
  remaining quadratic in w is solved with 60-digit decimals.
* Special cases that reduce to textbook routines: equal radii give the
  circumsphere of the 4 centres (minus the common radius), zero radii the
  Delaunay circumsphere.
* Common tangent planes of 3 spheres (math.cuh:577-634): each returned unit
  normal n makes n.p_m + r_m equal for the three spheres.
* sphereFromPoints (math.cuh:533-575): the 4 points lie on the sphere.
* The cell method (topology_update.inl, algorithm.inl) against brute force
  on small general-position sets: every quadruple of sites is solved and
  kept when no other site (artificial ones included) intrudes; the
  diagram's vertex set must equal it.
* Certification soundness (algorithm.inl:197-201): a cell certified after
  k neighbours keeps its records with more neighbours and through the
  validation stages.
* Labels of the learned proposal (benchmark_contributing_nearest.cu:75-113):
  label <= K exactly when the K-neighbour cell pass certifies the cell.
"""
from decimal import Decimal, getcontext
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from apo_amd import datasets

getcontext().prec = 60


def _solve3(A, b):
    """Exact 3x3 solve by Cramer's rule over Fractions."""
    def det(M):
        return (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
                - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
                + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]))
    D = det(A)
    out = []
    for k in range(3):
        M = [row[:] for row in A]
        for r in range(3):
            M[r][k] = b[r]
        out.append(det(M) / D)
    return out


def _dec(f: Fr) -> Decimal:
    return Decimal(f.numerator) / Decimal(f.denominator)


def _tangent_points(spheres, plane=False):
    """Points x with |x - p_m| = r_m + w for the given spheres (4, or 3 with x
    in the plane of their centres), w >= -r_1; exact up to the final sqrt."""
    P = [[Fr(float(v)) for v in s[:3]] for s in spheres]
    R = [Fr(float(s[3])) for s in spheres]
    rows, b, c = [], [], []
    for m in range(1, len(spheres)):
        rows.append([2 * (P[m][k] - P[0][k]) for k in range(3)])
        b.append(sum(P[m][k] ** 2 for k in range(3)) - sum(P[0][k] ** 2 for k in range(3)) - R[m] ** 2 + R[0] ** 2)
        c.append(-2 * (R[m] - R[0]))
    if plane:
        e1 = [P[1][k] - P[0][k] for k in range(3)]
        e2 = [P[2][k] - P[0][k] for k in range(3)]
        nrm = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
        rows.append(nrm)
        b.append(sum(nrm[k] * P[0][k] for k in range(3)))
        c.append(Fr(0))
    u = _solve3(rows, b)
    v = _solve3(rows, c)
    d = [u[k] - P[0][k] for k in range(3)]
    qa = sum(x * x for x in v) - 1
    qb = 2 * (sum(v[k] * d[k] for k in range(3)) - R[0])
    qc = sum(x * x for x in d) - R[0] ** 2
    disc = qb * qb - 4 * qa * qc
    if disc < 0:
        return [], disc
    s = _dec(disc).sqrt()
    out = []
    for sgn in (1, -1):
        w = (-_dec(qb) + sgn * s) / (2 * _dec(qa))
        if w + _dec(R[0]) < 0:
            continue
        x = [_dec(u[k]) + w * _dec(v[k]) for k in range(3)]
        out.append(np.array([float(x[0]), float(x[1]), float(x[2]), float(w)]))
    return out, disc


def _random_spheres(rng, count, rmax=0.2):
    s = np.zeros((count, 4))
    s[:, :3] = rng.random((count, 3))
    s[:, 3] = rng.random(count) * rmax
    return s.astype(np.float32).astype(np.float64)


def _match(got, want, tol):
    assert len(got) == len(want), (got, want)
    used = set()
    for g in got:
        hit = [q for q, w in enumerate(want) if q not in used and np.all(np.abs(g - w) <= tol * (1 + np.abs(w)))]
        assert hit, (g, want)
        used.add(hit[0])


def test_quadrisector_matches_exact_solve():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(300):
        s = _random_spheres(rng, 4)
        want, disc = _tangent_points(s)
        if abs(float(disc)) < 1e-6:  # skip near-double roots: the reference keeps one of them
            continue
        mask, x1, x2 = oracle.quadrisector(*s)
        got = [x for bit, x in ((1, x1), (2, x2)) if mask & bit]
        _match(got, want, 1e-7)
        # the returned radius is the tangent distance to each of the 4 spheres
        for x in got:
            d = np.linalg.norm(s[:, :3] - x[:3], axis=1) - s[:, 3]
            assert np.allclose(d, x[3], rtol=1e-7, atol=1e-9)
        checked += 1
    assert checked > 200


def test_trisector_matches_exact_solve():
    rng = np.random.default_rng(12)
    checked = 0
    for _ in range(300):
        s = _random_spheres(rng, 3)
        want, disc = _tangent_points(s, plane=True)
        if abs(float(disc)) < 1e-6:
            continue
        count, x1, x2 = oracle.trisector(*s)
        got = [x1, x2][:count]
        _match(got, want, 1e-7)
        checked += 1
    assert checked > 200


def _circumsphere(p):
    A = 2.0 * (p[1:] - p[0])
    b = (p[1:] ** 2).sum(axis=1) - (p[0] ** 2).sum()
    c = np.linalg.solve(A, b)
    return c, np.linalg.norm(p[0] - c)


@pytest.mark.parametrize("radius", [0.0, 0.05])
def test_equal_radii_give_the_circumsphere(radius):
    # all radii equal: the vertex is the circumcentre of the 4 centres and its
    # radius the circumradius minus the common radius (Delaunay for r = 0)
    rng = np.random.default_rng(13)
    for _ in range(100):
        p = rng.random((4, 3))
        s = np.concatenate([p, np.full((4, 1), radius)], axis=1)
        c, R = _circumsphere(p)
        mask, x1, x2 = oracle.quadrisector(*s)
        got = [x for bit, x in ((1, x1), (2, x2)) if mask & bit]
        assert len(got) == 1
        assert np.allclose(got[0][:3], c, rtol=1e-8, atol=1e-10)
        assert np.isclose(got[0][3], R - radius, rtol=1e-8, atol=1e-10)


def test_sphere_from_points_is_the_circumsphere():
    rng = np.random.default_rng(14)
    for _ in range(100):
        p = rng.random((4, 3))
        out = oracle.sphere_from_points(*p)
        c, R = _circumsphere(p)
        assert np.allclose(out[:3], c, rtol=1e-8, atol=1e-10)
        assert np.isclose(out[3], R, rtol=1e-8)


def test_tangent_planes_touch_all_three_spheres():
    rng = np.random.default_rng(15)
    checked = 0
    for _ in range(200):
        s = _random_spheres(rng, 3, rmax=0.05)
        n1, n2 = oracle.tangent_planes(*s)
        if not (np.all(np.isfinite(n1)) and np.all(np.isfinite(n2))):
            continue
        for n in (n1, n2):
            assert np.isclose(np.linalg.norm(n), 1.0, atol=1e-9)
            off = s[:, :3] @ n + s[:, 3]
            assert np.ptp(off) < 1e-8 * (1 + np.abs(off).max()), off
        assert not np.allclose(n1, n2)
        checked += 1
    assert checked > 150


def _rows(ids):
    a = np.sort(np.asarray(ids, np.uint32).reshape(-1, 4), axis=1)
    return a[np.lexsort(a.T[::-1])] if len(a) else a


@pytest.mark.parametrize("n,seed", [(10, 1), (16, 2), (24, 3), (32, 4), (40, 5)])
def test_diagram_equals_brute_force(n, seed):
    sites = datasets.uniform_sites(n, seed=100 + seed, r_min=0.01, r_max=0.06)
    _, ids = oracle.Diagram(sites).vertices()
    brute = oracle.bruteforce_vertices(sites)
    assert len(brute) > 0
    assert np.array_equal(_rows(ids), _rows(brute)), (n, len(ids), len(brute))


@pytest.mark.parametrize("gen,k", [("uniform", 16), ("uniform", 32), ("loguniform", 32)])
def test_certified_cells_do_not_change(gen, k):
    # a cell certified after k neighbours is final: more neighbours (64) and
    # the full pipeline (validation included) leave its vertex records as
    # they are
    sites = (datasets.uniform_sites(20_000, seed=201) if gen == "uniform"
             else datasets.loguniform_sites(20_000, seed=202))
    cells = []
    for d in (oracle.Diagram(sites, k=k, stage_mask=1), oracle.Diagram(sites, k=64, stage_mask=1),
              oracle.Diagram(sites)):
        c, V, _ = d.all_cells()
        start = np.concatenate([[0], np.cumsum(c["vertex_nb"].astype(np.int64))])
        cells.append((c, V, start))
    c16 = cells[0][0]
    certified = np.nonzero(c16["status"] == 1)[0]
    assert len(certified) > 100

    def recs(k, i):
        c, V, start = cells[k]
        r = V[start[i]:start[i + 1]].copy()
        r[:, 3] &= 0xFF
        return r[np.lexsort(r.T[::-1])]

    for i in certified:
        a = recs(0, i)
        assert np.array_equal(a, recs(1, i)), i
        assert np.array_equal(a, recs(2, i)), i


def test_labels_agree_with_k_neighbour_certification():
    sites = datasets.loguniform_sites(3000, seed=41)
    labels = oracle.certified_k(sites, 64)
    for K in (16, 24, 32, 48, 64):
        d = oracle.Diagram(sites, k=K, stage_mask=1)
        st = d.cell_counts()["status"]
        _, perm = d.sorted_sites()
        lab = labels[perm]
        live = st != 0xFF
        want = (lab >= 1) & (lab <= K)
        assert np.array_equal(st[live] == 1, want[live]), K


def test_pcg_stream_two_implementations():
    # the sample stream (samples.cpp:64-85) and the joggle (utils.cpp:6-25)
    # use the same hash: the C oracle and numpy must agree bit for bit
    v = np.arange(0, 1 << 20, 997, dtype=np.uint32)
    lib = oracle.lib()
    got = np.array([lib.orc_pcg_unit(int(x)) for x in v])
    assert np.array_equal(got, datasets.pcg_unit(v))
