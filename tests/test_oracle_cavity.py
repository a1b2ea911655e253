"""Pins the oracle's cavity hit ratios (app/cavity/utils.cu:17-65, 201-242)
without reusing its arithmetic.

* Directions: the Halton construction recomputed in float64 from the radical
  inverses in bases 2 and 3 (van der Corput), unit length.
* One sphere seen from the vertex: a ray hits exactly when its angle to the
  sphere's centre is below asin(r / D); counted in float64 from the
  directions, away from grazing rays.
* A vertex enclosed by a shell of spheres is hit by every ray; a vertex with
  every sphere behind tmin = 0.01 or off every ray is hit by none.
"""
import numpy as np

import oracle


def _radical_inverse(k, base):
    f, r = 1.0 / base, 0.0
    while k:
        r += f * (k % base)
        k //= base
        f /= base
    return r


def test_directions_are_the_halton_points():
    d = oracle.cavity_directions(256)
    for k in range(256):
        t = 2.0 * _radical_inverse(k, 2) - 1.0
        phi = 4.0 * np.pi * _radical_inverse(k, 3)
        st = np.sqrt(max(0.0, 1.0 - t * t))
        want = np.array([st * np.cos(phi), st * np.sin(phi), t])
        assert np.allclose(d[k], want / np.linalg.norm(want), atol=2e-6), (k, d[k], want)
    assert np.allclose(np.linalg.norm(d.astype(np.float64), axis=1), 1.0, atol=1e-6)


def test_single_sphere_solid_angle():
    dirs = oracle.cavity_directions(64).astype(np.float64)
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(40):
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        D, r = rng.uniform(2.0, 5.0), rng.uniform(0.3, 1.5)
        site = np.array([[*(u * D), r]], np.float32)
        cosang = dirs @ (site[0, :3].astype(np.float64) / np.linalg.norm(site[0, :3]))
        limit = np.sqrt(1.0 - (r / np.linalg.norm(site[0, :3])) ** 2)
        if np.min(np.abs(cosang - limit)) < 1e-5:
            continue  # a grazing ray: FP32 decides it
        want = np.sum(cosang > limit) / 64.0
        got = oracle.cavity_hit_ratios(site, np.zeros((1, 4)))
        assert got[0] == np.float32(want), (got, want)
        checked += 1
    assert checked > 30


def test_enclosed_and_open_vertices():
    dirs = oracle.cavity_directions(64)
    # a dense shell of radius 3 around the origin: every ray is stopped
    g = np.random.default_rng(1).standard_normal((4000, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    shell = np.concatenate([3.0 * g, np.full((4000, 1), 0.3)], axis=1).astype(np.float32)
    assert oracle.cavity_hit_ratios(shell, np.zeros((1, 4)))[0] == 1.0
    # spheres whose every intersection lies before tmin, or off every ray
    tiny = np.array([[0.0, 0.0, 0.0, 0.005], [50.0, 50.0, 50.0, 0.5]], np.float32)
    assert oracle.cavity_hit_ratios(tiny, np.zeros((1, 4)))[0] == 0.0
    assert len(dirs) == 64
