"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle restates PlathC/apo's construction on the host (see
oracle/apo_oracle.cpp).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
i32p = ctypes.POINTER(ctypes.c_int32)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)
u8p = ctypes.POINTER(ctypes.c_uint8)


class OrcConfig(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint32) for name in (
        "k", "vertex_capacity", "edge_capacity", "new_edge_capacity", "bisector_capacity",
        "max_iterations", "max_restarts", "threads", "stage_mask", "edge_bound", "fallback_k", "knn_passes",
        "complete", "lanes", "bisector_bound")]


def build() -> str:
    """Compile oracle/liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_pcg_unit.restype = ctypes.c_double
        L.orc_pcg_unit.argtypes = [ctypes.c_uint32]
        L.orc_pcg_ldexp.restype = ctypes.c_double
        L.orc_pcg_ldexp.argtypes = [ctypes.c_uint32]
        L.orc_quadrisector.restype = ctypes.c_uint32
        L.orc_quadrisector.argtypes = [f64p] * 6
        L.orc_trisector.restype = ctypes.c_uint32
        L.orc_trisector.argtypes = [f64p] * 5
        L.orc_tangent_planes.restype = None
        L.orc_tangent_planes.argtypes = [f64p] * 5
        L.orc_sphere_from_points.restype = None
        L.orc_sphere_from_points.argtypes = [f64p] * 5
        L.orc_scene_box.restype = None
        L.orc_scene_box.argtypes = [f32p, ctypes.c_uint32, f32p, f64p]
        L.orc_morton64.restype = ctypes.c_uint64
        L.orc_morton64.argtypes = [ctypes.c_uint32] * 3
        L.orc_artificial_sites.restype = None
        L.orc_artificial_sites.argtypes = [f32p, ctypes.c_uint32, f64p]
        L.orc_morton_codes.restype = None
        L.orc_morton_codes.argtypes = [f32p, ctypes.c_uint32, f32p, u64p]
        L.orc_lbvh_build.restype = ctypes.c_int
        L.orc_lbvh_build.argtypes = [f32p, ctypes.c_uint32, u64p, u32p, i32p, i32p, i32p, f32p, u32p, f32p]
        L.orc_knn.restype = ctypes.c_int
        L.orc_knn.argtypes = [f32p, ctypes.c_uint32, i32p, i32p, f32p, ctypes.c_uint32, u32p, f32p]
        L.orc_knn_brute.restype = None
        L.orc_knn_brute.argtypes = [f32p, ctypes.c_uint32, ctypes.c_uint32, u32p, f32p]
        L.orc_default_config.restype = None
        L.orc_default_config.argtypes = [ctypes.POINTER(OrcConfig)]
        L.orc_build.restype = ctypes.c_void_p
        L.orc_build.argtypes = [f32p, ctypes.c_uint32, ctypes.POINTER(OrcConfig)]
        L.orc_build_proposal.restype = ctypes.c_void_p
        L.orc_build_proposal.argtypes = [f32p, ctypes.c_uint32, ctypes.POINTER(OrcConfig), u8p]
        L.orc_build_radii.restype = ctypes.c_void_p
        L.orc_build_radii.argtypes = [f32p, ctypes.c_uint32, ctypes.POINTER(OrcConfig), f32p]
        L.orc_free.restype = None
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_vertex_count.restype = ctypes.c_uint64
        L.orc_vertex_count.argtypes = [ctypes.c_void_p]
        L.orc_get_vertices.restype = None
        L.orc_get_vertices.argtypes = [ctypes.c_void_p, f64p, u32p]
        L.orc_closed_edge_count.restype = ctypes.c_uint64
        L.orc_closed_edge_count.argtypes = [ctypes.c_void_p]
        L.orc_get_closed_edges.restype = None
        L.orc_get_closed_edges.argtypes = [ctypes.c_void_p, f64p, u32p]
        L.orc_get_sorted_sites.restype = None
        L.orc_get_sorted_sites.argtypes = [ctypes.c_void_p, f32p, u32p]
        L.orc_get_cell_counts.restype = None
        L.orc_get_cell_counts.argtypes = [ctypes.c_void_p, u32p, u32p, u8p, f32p, u32p]
        L.orc_get_cell.restype = None
        L.orc_get_cell.argtypes = [ctypes.c_void_p, ctypes.c_uint32, u32p, u32p]
        L.orc_get_all_cells.restype = None
        L.orc_get_all_cells.argtypes = [ctypes.c_void_p, u32p, u32p]
        L.orc_topology.restype = None
        L.orc_topology.argtypes = [ctypes.c_void_p, u64p, u32p, u32p, u32p]
        L.orc_get_knns.restype = None
        L.orc_get_knns.argtypes = [ctypes.c_void_p, u32p]
        L.orc_box_reaches_half_space.restype = ctypes.c_int
        L.orc_box_reaches_half_space.argtypes = [f32p, f32p, f32p]
        L.orc_get_stats.restype = None
        L.orc_get_stats.argtypes = [ctypes.c_void_p, u64p]
        L.orc_certified_k.restype = None
        L.orc_certified_k.argtypes = [f32p, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32]
        L.orc_empty_sphere_violations.restype = ctypes.c_uint64
        L.orc_empty_sphere_violations.argtypes = [f32p, ctypes.c_uint32, f64p, u32p, ctypes.c_uint64,
                                                  ctypes.c_double, ctypes.c_uint32]
        L.orc_bruteforce_vertices.restype = ctypes.c_uint64
        L.orc_bruteforce_vertices.argtypes = [f32p, ctypes.c_uint32, u32p, ctypes.c_uint64, ctypes.c_double]
        L.orc_cavity_directions.restype = None
        L.orc_cavity_directions.argtypes = [ctypes.c_uint32, f32p]
        L.orc_cavity_hit_ratios.restype = None
        L.orc_cavity_hit_ratios.argtypes = [f32p, ctypes.c_uint32, f64p, ctypes.c_uint64, ctypes.c_uint32,
                                            ctypes.c_uint32, f32p]
        L.orc_precision_counts.restype = None
        L.orc_precision_counts.argtypes = [f32p, ctypes.c_uint32, f64p, u32p, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_uint32, u64p]
        L.orc_round_dir.restype = None
        L.orc_round_dir.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, f64p]
        L.orc_interval_distance.restype = None
        L.orc_interval_distance.argtypes = [f64p, f64p, f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _sites(xyzr) -> np.ndarray:
    a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
    return a


def quadrisector(s1, s2, s3, s4):
    s = [np.ascontiguousarray(x, dtype=np.float64) for x in (s1, s2, s3, s4)]
    x1 = np.zeros(4)
    x2 = np.zeros(4)
    m = lib().orc_quadrisector(*[_p(a, f64p) for a in s], _p(x1, f64p), _p(x2, f64p))
    return int(m), x1, x2


def trisector(s1, s2, s3):
    s = [np.ascontiguousarray(x, dtype=np.float64) for x in (s1, s2, s3)]
    x1 = np.zeros(4)
    x2 = np.zeros(4)
    m = lib().orc_trisector(*[_p(a, f64p) for a in s], _p(x1, f64p), _p(x2, f64p))
    return int(m), x1, x2


def tangent_planes(s1, s2, s3):
    s = [np.ascontiguousarray(x, dtype=np.float64) for x in (s1, s2, s3)]
    n1 = np.zeros(3)
    n2 = np.zeros(3)
    lib().orc_tangent_planes(*[_p(a, f64p) for a in s], _p(n1, f64p), _p(n2, f64p))
    return n1, n2


def sphere_from_points(p1, p2, p3, p4):
    p = [np.ascontiguousarray(x, dtype=np.float64) for x in (p1, p2, p3, p4)]
    out = np.zeros(4)
    lib().orc_sphere_from_points(*[_p(a, f64p) for a in p], _p(out, f64p))
    return out


def scene_box(xyzr):
    s = _sites(xyzr)
    box = np.zeros(6, np.float32)
    mm = np.zeros(2)
    lib().orc_scene_box(_p(s, f32p), len(s), _p(box, f32p), _p(mm, f64p))
    return box, mm


def morton64(x, y, z) -> int:
    return int(lib().orc_morton64(x, y, z))


def artificial_sites(xyzr):
    s = _sites(xyzr)
    out = np.zeros((6, 4), np.float64)
    lib().orc_artificial_sites(_p(s, f32p), len(s), _p(out, f64p))
    return out


def morton_codes(xyzr):
    s = _sites(xyzr)
    box, _ = scene_box(s)
    codes = np.zeros(len(s), np.uint64)
    lib().orc_morton_codes(_p(s, f32p), len(s), _p(box, f32p), _p(codes, u64p))
    return codes


def lbvh(xyzr) -> dict:
    s = _sites(xyzr)
    n = len(s)
    nodes = 2 * n - 1
    out = dict(codes=np.zeros(n, np.uint64), indices=np.zeros(n, np.uint32),
               parent=np.zeros(nodes, np.int32), left=np.zeros(nodes, np.int32),
               right=np.zeros(nodes, np.int32), aabb=np.zeros((nodes, 6), np.float32),
               sizes=np.zeros(nodes, np.uint32), max_radius=np.zeros(nodes, np.float32))
    rc = lib().orc_lbvh_build(_p(s, f32p), n, _p(out["codes"], u64p), _p(out["indices"], u32p),
                              _p(out["parent"], i32p), _p(out["left"], i32p), _p(out["right"], i32p),
                              _p(out["aabb"], f32p), _p(out["sizes"], u32p), _p(out["max_radius"], f32p))
    if rc != 0:
        raise ValueError("oracle lbvh needs n >= 2")
    return out


def knn(sorted_xyzr, bvh: dict, k: int = 16):
    s = _sites(sorted_xyzr)
    n = len(s)
    ids = np.zeros((n, k), np.uint32)
    dist = np.zeros((n, k), np.float32)
    depth = lib().orc_knn(_p(s, f32p), n, _p(bvh["left"], i32p), _p(bvh["right"], i32p),
                          _p(np.ascontiguousarray(bvh["aabb"]), f32p), k, _p(ids, u32p), _p(dist, f32p))
    return ids, dist, int(depth)


def knn_brute(sorted_xyzr, k: int = 16):
    s = _sites(sorted_xyzr)
    n = len(s)
    ids = np.zeros((n, k), np.uint32)
    dist = np.zeros((n, k), np.float32)
    lib().orc_knn_brute(_p(s, f32p), n, k, _p(ids, u32p), _p(dist, f32p))
    return ids, dist


def default_config(**kw) -> OrcConfig:
    cfg = OrcConfig()
    lib().orc_default_config(ctypes.byref(cfg))
    for key, val in kw.items():
        setattr(cfg, key, val)
    return cfg


class Diagram:
    """Result of the oracle's full construction."""

    def __init__(self, xyzr, classes=None, radii=None, **cfg):
        """classes: per-site proposal classes (input order) for the learned
        path; radii: per-site search radii (input order) for the range path."""
        self.sites = _sites(xyzr)
        self.n = len(self.sites)
        self.cfg = default_config(**cfg)
        if radii is not None:
            self._radii = np.ascontiguousarray(radii, np.float32)
            assert len(self._radii) == self.n
            self._h = lib().orc_build_radii(_p(self.sites, f32p), self.n, ctypes.byref(self.cfg),
                                            _p(self._radii, f32p))
        elif classes is None:
            self._h = lib().orc_build(_p(self.sites, f32p), self.n, ctypes.byref(self.cfg))
        else:
            self._classes = np.ascontiguousarray(classes, np.uint8)
            assert len(self._classes) == self.n
            self._h = lib().orc_build_proposal(_p(self.sites, f32p), self.n, ctypes.byref(self.cfg),
                                               _p(self._classes, u8p))
        if not self._h:
            raise ValueError("oracle build needs n >= 2")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_free(h)
            self._h = None

    def vertices(self):
        v = int(lib().orc_vertex_count(self._h))
        xyzw = np.zeros((v, 4), np.float64)
        ids = np.zeros((v, 4), np.uint32)
        if v:
            lib().orc_get_vertices(self._h, _p(xyzw, f64p), _p(ids, u32p))
        return xyzw, ids

    def closed_edges(self):
        """toHostFull's closed trisector edges: (x, y, z, w) and (i, j, k, 0)."""
        e = int(lib().orc_closed_edge_count(self._h))
        xyzw = np.zeros((e, 4), np.float64)
        ids = np.zeros((e, 4), np.uint32)
        if e:
            lib().orc_get_closed_edges(self._h, _p(xyzw, f64p), _p(ids, u32p))
        return xyzw, ids

    def sorted_sites(self):
        s = np.zeros((self.n, 4), np.float32)
        idx = np.zeros(self.n, np.uint32)
        lib().orc_get_sorted_sites(self._h, _p(s, f32p), _p(idx, u32p))
        return s, idx

    def cell_counts(self):
        n = self.n
        vn = np.zeros(n, np.uint32)
        en = np.zeros(n, np.uint32)
        st = np.zeros(n, np.uint8)
        kr = np.zeros(n, np.float32)
        fl = np.zeros(n, np.uint32)
        lib().orc_get_cell_counts(self._h, _p(vn, u32p), _p(en, u32p), _p(st, u8p), _p(kr, f32p), _p(fl, u32p))
        return dict(vertex_nb=vn, edge_nb=en, status=st, known_radius=kr, flags=fl)

    def all_cells(self):
        c = self.cell_counts()
        V = np.zeros((int(c["vertex_nb"].sum()), 4), np.uint32)
        E = np.zeros((int(c["edge_nb"].sum()), 4), np.uint32)
        lib().orc_get_all_cells(self._h, _p(V, u32p), _p(E, u32p))
        return c, V, E

    def topology(self) -> dict:
        """toTopology sets (algorithm.inl:1732-1821), original ids."""
        c = np.zeros(3, np.uint64)
        lib().orc_topology(self._h, _p(c, u64p), None, None, None)
        b = np.zeros((int(c[0]), 2), np.uint32)
        t = np.zeros((int(c[1]), 3), np.uint32)
        q = np.zeros((int(c[2]), 4), np.uint32)
        lib().orc_topology(self._h, _p(c, u64p), _p(b, u32p), _p(t, u32p), _p(q, u32p))
        return {"bisectors": b, "trisectors": t, "quadrisectors": q}

    def knns(self):
        out = np.zeros((self.n, self.cfg.k), np.uint32)
        lib().orc_get_knns(self._h, _p(out, u32p))
        return out

    def stats(self):
        s = np.zeros(12, np.uint64)
        lib().orc_get_stats(self._h, _p(s, u64p))
        keys = ["uncertified_after_cells", "validation_tasks", "validation_iterations", "flagged_cells",
                "vertices", "buried", "bisector_changed", "completion_rounds", "completion_updated_cells",
                "completion_unrepairable", "completion_validation_tasks", "completion_rebuilt_cells"]
        return dict(zip(keys, (int(x) for x in s)))


def box_reaches_half_space(box6, n, p) -> bool:
    b = np.ascontiguousarray(box6, np.float32)
    nn = np.ascontiguousarray(n, np.float32)
    pp = np.ascontiguousarray(p, np.float32)
    return bool(lib().orc_box_reaches_half_space(_p(b, f32p), _p(nn, f32p), _p(pp, f32p)))


def certified_k(xyzr, kmax: int = 64, threads: int = 0):
    s = _sites(xyzr)
    out = np.zeros(len(s), np.uint32)
    lib().orc_certified_k(_p(s, f32p), len(s), kmax, _p(out, u32p), threads)
    return out


def empty_sphere_violations(xyzr, xyzw, ids, tol: float = 1e-9, threads: int = 0) -> int:
    s = _sites(xyzr)
    xyzw = np.ascontiguousarray(xyzw, np.float64)
    ids = np.ascontiguousarray(ids, np.uint32)
    return int(lib().orc_empty_sphere_violations(_p(s, f32p), len(s), _p(xyzw, f64p), _p(ids, u32p),
                                                 len(ids), tol, threads))


def bruteforce_vertices(xyzr, tol: float = 1e-9, cap: int = 1 << 16):
    s = _sites(xyzr)
    ids = np.zeros((cap, 4), np.uint32)
    cnt = int(lib().orc_bruteforce_vertices(_p(s, f32p), len(s), _p(ids, u32p), cap, tol))
    return ids[:min(cnt, cap)]


def precision_counts(xyzr, xyzw, ids, arity: int = 4, threads: int = 0):
    """Interval validator of app/precision/precision.cpp:14-163 by a scan of
    every site: (bad, uncertain) points.  arity 4: vertices, 3: closed edges."""
    s = _sites(xyzr)
    x = np.ascontiguousarray(xyzw, np.float64).reshape(-1, 4)
    i = np.ascontiguousarray(ids, np.uint32).reshape(-1, 4)
    out = np.zeros(2, np.uint64)
    lib().orc_precision_counts(_p(s, f32p), len(s), _p(x, f64p), _p(i, u32p), len(x), arity, threads,
                               _p(out, u64p))
    return int(out[0]), int(out[1])


def round_dir(op: str, a: float, b: float = 0.0):
    """(down, up) of a+b, a*b or sqrt(a) under the switched rounding mode."""
    out = np.zeros(2)
    lib().orc_round_dir({"add": 0, "mul": 1, "sqrt": 2}[op], a, b, _p(out, f64p))
    return float(out[0]), float(out[1])


def interval_distance(site, point):
    s = np.ascontiguousarray(site, np.float64)
    p = np.ascontiguousarray(point, np.float64)
    out = np.zeros(2)
    lib().orc_interval_distance(_p(s, f64p), _p(p, f64p), _p(out, f64p))
    return float(out[0]), float(out[1])


def cavity_directions(samples: int = 64):
    out = np.zeros((samples, 3), np.float32)
    lib().orc_cavity_directions(samples, _p(out, f32p))
    return out


def cavity_hit_ratios(xyzr, xyzw, samples: int = 64, threads: int = 0):
    """Hit ratio of every vertex (app/cavity/utils.cu:17-65), every ray
    against every site."""
    s = _sites(xyzr)
    x = np.ascontiguousarray(xyzw, np.float64).reshape(-1, 4)
    out = np.zeros(len(x), np.float32)
    lib().orc_cavity_hit_ratios(_p(s, f32p), len(s), _p(x, f64p), len(x), samples, threads, _p(out, f32p))
    return out
