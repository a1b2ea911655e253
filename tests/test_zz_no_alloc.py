"""apo_build_device_static: the build that allocates nothing, through the C ABI.

Its output must equal apo_build_device's (vertices and ids bit for bit) on
two calls in a row, the second with fewer sites, also when the calls are
interleaved with ordinary builds on the same context (whose own buffers it
must leave intact).  Sizes above APO_STATIC_MAX_SITES and too-small output
buffers fail with a status.
"""
import ctypes

import numpy as np
import pytest

from apo_amd import _lib, datasets

pytestmark = pytest.mark.gpu

MAX_SITES = 4096
ERR_INVALID, ERR_CAPACITY = 3, 5


class _Hip:
    """Device memory through the HIP runtime libapo_hip.so is linked to (by
    its soname, so the loaded copy), independent of any other binding."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.live = []

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(nbytes, 1)) == 0
        self.live.append(p)
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


@pytest.fixture
def hip():
    _lib.lib()  # the library (and its HIP runtime) first
    h = _Hip()
    yield h
    h.free()


def _ctx(**fields):
    cfg = _lib.Config()
    _lib.lib().apo_default_config(ctypes.byref(cfg))
    for k, v in fields.items():
        setattr(cfg, k, v)
    return _lib.create(cfg)


def _reference(ctx, hip, sites):
    """The existing entry point: apo_build_device + apo_get_vertices."""
    dev = hip.upload(sites)
    L = _lib.lib()
    _lib.check(ctx, L.apo_build_device(ctx, dev, len(sites)))
    cnt = ctypes.c_uint64()
    _lib.check(ctx, L.apo_vertex_count(ctx, ctypes.byref(cnt)))
    x = np.zeros((cnt.value, 4), np.float64)
    i = np.zeros((cnt.value, 4), np.uint32)
    _lib.check(ctx, L.apo_get_vertices(ctx, _lib.dptr(x), _lib.uptr(i), cnt.value))
    return x, i


def _static(ctx, hip, sites, capacity):
    dev = hip.upload(sites)
    x = hip.alloc(capacity * 32)
    i = hip.alloc(capacity * 16)
    cnt = ctypes.c_uint64()
    st = _lib.lib().apo_build_device_static(ctx, dev, len(sites), x, i, capacity, ctypes.byref(cnt))
    return st, int(cnt.value), hip.download(x, (capacity, 4), np.float64), hip.download(i, (capacity, 4), np.uint32)


@pytest.mark.parametrize("cfg", [{}, {"k": 4, "knn_passes": 3}])
def test_static_build_equals_build_device(hip, cfg):
    ctx = _ctx(**cfg)
    try:
        for n, seed in ((MAX_SITES, 1), (1000, 2)):   # two calls in a row, the second smaller
            sites = datasets.uniform_sites(n, seed=seed)
            print(f"[case] static build n={n} cfg={cfg}", flush=True)
            st, cnt, x, i = _static(ctx, hip, sites, 12 * n)
            _lib.check(ctx, st)
            want_x, want_i = _reference(ctx, hip, sites)
            assert cnt == len(want_i) > 0
            assert np.array_equal(i[:cnt], want_i)
            assert np.array_equal(x[:cnt].view(np.uint64), want_x.view(np.uint64))
    finally:
        _lib.lib().apo_destroy(ctx)


def test_static_build_leaves_the_context_build_intact(hip):
    # an ordinary build, a static build of other sites, then the ordinary
    # build's buffers again: a new ordinary build gives its result unchanged
    ctx = _ctx()
    try:
        a = datasets.loguniform_sites(3000, seed=5)
        b = datasets.protein_like_sites(2500, seed=6)
        ax, ai = _reference(ctx, hip, a)
        st, cnt, x, i = _static(ctx, hip, b, 20 * len(b))
        _lib.check(ctx, st)
        # the context holds no build after a static one
        c = ctypes.c_uint64()
        assert _lib.lib().apo_vertex_count(ctx, ctypes.byref(c)) == 4
        ax2, ai2 = _reference(ctx, hip, a)
        assert np.array_equal(ai2, ai) and np.array_equal(ax2.view(np.uint64), ax.view(np.uint64))
        bx, bi = _reference(ctx, hip, b)
        assert cnt == len(bi)
        assert np.array_equal(i[:cnt], bi) and np.array_equal(x[:cnt].view(np.uint64), bx.view(np.uint64))
    finally:
        _lib.lib().apo_destroy(ctx)


def test_static_build_limits(hip):
    ctx = _ctx()
    try:
        st, _, _, _ = _static(ctx, hip, datasets.uniform_sites(MAX_SITES + 1, seed=3), 16)
        assert st == ERR_INVALID
        sites = datasets.uniform_sites(500, seed=4)
        want_x, want_i = _reference(ctx, hip, sites)
        st, cnt, x, i = _static(ctx, hip, sites, 100)
        assert st == ERR_CAPACITY and cnt == len(want_i)
        assert np.array_equal(i[:100], want_i[:100])
    finally:
        _lib.lib().apo_destroy(ctx)
