"""MI355X (gfx950) Apollonius-diagram engine: Python host interface.

This mirrors the construction API of PlathC/apo
(apo/include/apo/gpu/algorithm.hpp:9-59, algorithm.cuh:12-46):

    AlgorithmGPU(sites, k=16, knn_validation_step=1).build()
                                                -> AlgorithmGPU<K>(sites, knnValidationStep).build
    AlgorithmGPU.to_host()                      -> AlgorithmGPU::toHost -> VertexDiagram
    AlgorithmGPU.to_host_full()                 -> AlgorithmGPU::toHostFull -> FullDiagram
    AlgorithmGPU.topology()                     -> AlgorithmGPU::toTopology -> Topology sets
    compute_apollonius_diagram(sites)           -> apo::gpu::computeApolloniusDiagram
    compute_full_apollonius_diagram(sites)      -> apo::gpu::computeFullApolloniusDiagram
    topology_json(topology)                     -> apo::toCsv (app/topology/topology.cpp:11-57)
    AlgorithmGPU(sites, offsets=o).build()      -> one build of many independent molecules

Everything runs in the HIP library apo_amd/lib/libapo_hip.so through its C
ABI (include/apo_c.h).  There is no CPU fallback: if the library is missing
or no GPU is present, construction raises.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ApoError, library_path  # noqa: F401

__all__ = ["AlgorithmGPU", "VertexDiagram", "FullDiagram", "compute_apollonius_diagram",
           "compute_full_apollonius_diagram", "topology_json", "ApoError", "library_path"]


@dataclass
class VertexDiagram:
    """Host result of apo's toHost (algorithm.hpp:10-16): vertices as (x, y, z, w)
    with w the empty-sphere radius, and the 4 tangent sites (original indices)."""
    vertices: np.ndarray      # (V, 4) float64
    vertices_id: np.ndarray   # (V, 4) uint32


@dataclass
class FullDiagram:
    """Host result of apo's toHostFull (algorithm.hpp:18-26): the vertices of
    VertexDiagram plus every closed trisector edge as its extreme point of
    smaller empty-sphere radius (x, y, z, w) and its 3 sites (i, j, k, 0)."""
    vertices: np.ndarray          # (V, 4) float64
    vertices_id: np.ndarray       # (V, 4) uint32
    closed_edges_min: np.ndarray  # (E, 4) float64
    closed_edges_id: np.ndarray   # (E, 4) uint32


class AlgorithmGPU:
    """One construction context on one GPU (AlgorithmGPU<K>, algorithm.cuh:12-46).

    k: neighbours per kNN pass (the template argument K, 1..64).
    knn_validation_step: kNN + cell passes (the constructor's second
    argument, algorithm.cuh:16); each later pass continues the neighbour
    lists after the previous pass's last neighbour.
    offsets: for a batch, the molecule boundaries (m + 1 values, molecule i is
    sites[offsets[i]:offsets[i+1]]), built in one pass of offset-indexed
    LBVHs and persistent work-queue kernels.  graph: capture the device build
    in a HIP graph and replay it on later builds of the same shape.  Other
    keywords set fields of apo_config (include/apo_c.h), e.g. cover_ratio,
    waves_edge or proposal_fallback_k."""

    def __init__(self, sites, k: int = 16, max_iterations: int = 10, max_restarts: int = 100,
                 stage_mask: int = 7, waves_per_cu: int = 0, device: int = 0, proposal=None,
                 proposal_cost: float = 0.0, offsets=None, graph: bool = False, classes=None,
                 knn_validation_step: int = 1, radii=None, **tuning):
        self.sites = np.ascontiguousarray(np.asarray(sites, dtype=np.float32).reshape(-1, 4))
        self.offsets = None if offsets is None else np.ascontiguousarray(offsets, np.uint32)
        self.k = int(k)
        cfg = _lib.Config()
        _lib.lib().apo_default_config(_lib.byref(cfg))
        cfg.k = self.k
        cfg.max_iterations = max_iterations
        cfg.max_restarts = max_restarts
        cfg.stage_mask = stage_mask
        cfg.waves_per_cu = waves_per_cu
        cfg.device = device
        cfg.proposal_cost = proposal_cost
        cfg.graph = 1 if graph else 0
        cfg.knn_passes = int(knn_validation_step)
        # further apo_config fields (cover_ratio, waves_edge, proposal_fallback_k ...)
        names = {f for f, _ in cfg._fields_}
        for key, val in tuning.items():
            if key not in names:
                raise TypeError(f"unknown configuration field {key}")
            setattr(cfg, key, val)
        self._ctx = _lib.create(cfg)
        self.built = False
        if proposal is not None:
            self._proposal = np.ascontiguousarray(proposal, np.float32)
            _lib.check(self._ctx, _lib.lib().apo_proposal_set(self._ctx, _lib.fptr(self._proposal)))
        if radii is not None:
            # per-site search radii (input order): range-query cell pass
            self._radii = np.ascontiguousarray(radii, np.float32)
            _lib.check(self._ctx, _lib.lib().apo_proposal_set_radii(self._ctx, _lib.fptr(self._radii),
                                                                    len(self._radii)))
        if classes is not None:
            # per-site proposal classes (input order) instead of the MLP's
            self._classes = np.ascontiguousarray(classes, np.uint8)
            _lib.check(self._ctx, _lib.lib().apo_proposal_set_classes(self._ctx, _lib.u8ptr(self._classes),
                                                                      len(self._classes)))

    def close(self):
        if getattr(self, "_ctx", None):
            _lib.lib().apo_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    # -- AlgorithmGPU::build ------------------------------------------------
    def build(self):
        if self.offsets is None:
            st = _lib.lib().apo_build(self._ctx, _lib.fptr(self.sites), len(self.sites))
        else:
            if len(self.offsets) < 2 or int(self.offsets[-1]) != len(self.sites):
                raise ValueError("offsets must end at the number of sites")
            st = _lib.lib().apo_build_batched(self._ctx, _lib.fptr(self.sites), _lib.uptr(self.offsets),
                                              len(self.offsets) - 1)
        _lib.check(self._ctx, st)
        self.built = True
        return self

    @property
    def molecules(self) -> int:
        return 1 if self.offsets is None else len(self.offsets) - 1

    def batch_vertex_offsets(self) -> np.ndarray:
        """First output vertex of each molecule (molecules + 1 values)."""
        out = np.zeros(self.molecules + 1, np.uint64)
        _lib.check(self._ctx, _lib.lib().apo_batch_vertex_offsets(self._ctx, _lib.u64ptr(out)))
        return out

    # -- AlgorithmGPU::toHost -----------------------------------------------
    def to_host(self) -> VertexDiagram:
        L = _lib.lib()
        cnt = _lib.c_uint64()
        _lib.check(self._ctx, L.apo_vertex_count(self._ctx, _lib.byref(cnt)))
        v = int(cnt.value)
        xyzw = np.zeros((v, 4), np.float64)
        ids = np.zeros((v, 4), np.uint32)
        _lib.check(self._ctx, L.apo_get_vertices(self._ctx, _lib.dptr(xyzw), _lib.uptr(ids), v))
        return VertexDiagram(xyzw, ids)

    # -- AlgorithmGPU::toHostFull -------------------------------------------
    def to_host_full(self) -> FullDiagram:
        """Vertices and closed trisector edges (algorithm.inl:1648-1730)."""
        L = _lib.lib()
        vd = self.to_host()
        cnt = _lib.c_uint64()
        _lib.check(self._ctx, L.apo_closed_edge_count(self._ctx, _lib.byref(cnt)))
        e = int(cnt.value)
        xyzw = np.zeros((e, 4), np.float64)
        ids = np.zeros((e, 4), np.uint32)
        if e:
            _lib.check(self._ctx, L.apo_get_closed_edges(self._ctx, _lib.dptr(xyzw), _lib.uptr(ids), e))
        return FullDiagram(vd.vertices, vd.vertices_id, xyzw, ids)

    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(self._ctx, _lib.lib().apo_get_stats(self._ctx, _lib.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_ if name != "reserved"}

    def empty_sphere_violations(self, xyzw=None, ids=None, tol: float = 1e-9) -> int:
        """FP64 empty-sphere check through the LBVH (the criterion of
        app/precision/precision.cpp:26-100).  Default: the build's own output."""
        cnt = _lib.c_uint64()
        if xyzw is None:
            st = _lib.lib().apo_check_empty_spheres(self._ctx, None, None, 0, tol, _lib.byref(cnt))
        else:
            x = np.ascontiguousarray(xyzw, np.float64)
            i = np.ascontiguousarray(ids, np.uint32)
            st = _lib.lib().apo_check_empty_spheres(self._ctx, _lib.dptr(x), _lib.uptr(i), len(x), tol,
                                                    _lib.byref(cnt))
        _lib.check(self._ctx, st)
        return int(cnt.value)

    def precision(self) -> dict:
        """Interval validator of app/precision/precision.cpp:14-163 on the
        build's vertices and closed edges: counts of points, of bad ones (some
        other site certainly closer) and of uncertain (near-degenerate) ones."""
        r = _lib.Precision()
        _lib.check(self._ctx, _lib.lib().apo_precision_check(self._ctx, _lib.byref(r)))
        return {name: int(getattr(r, name)) for name, _ in r._fields_}

    def precision_points(self, xyzw, ids, arity: int = 4):
        """The interval validator on given points: (bad, uncertain)."""
        x = np.ascontiguousarray(xyzw, np.float64).reshape(-1, 4)
        i = np.ascontiguousarray(ids, np.uint32).reshape(-1, 4)
        bad, unc = _lib.c_uint64(), _lib.c_uint64()
        _lib.check(self._ctx, _lib.lib().apo_precision_check_points(self._ctx, _lib.dptr(x), _lib.uptr(i), len(x),
                                                                    arity, _lib.byref(bad), _lib.byref(unc)))
        return int(bad.value), int(unc.value)

    # -- cavity detection (app/cavity/utils.cu) -------------------------------
    def hit_ratios(self, samples: int = 64) -> np.ndarray:
        """Fraction of `samples` rays from each output vertex that hit a site."""
        cnt = _lib.c_uint64()
        _lib.check(self._ctx, _lib.lib().apo_vertex_count(self._ctx, _lib.byref(cnt)))
        out = np.zeros(int(cnt.value), np.float32)
        _lib.check(self._ctx, _lib.lib().apo_cavity_hit_ratios(self._ctx, samples, _lib.fptr(out)))
        return out

    def cavity_vertices(self, hit_ratio_threshold: float = 0.5, radius_threshold: float = 5.0) -> np.ndarray:
        """Vertices (FP32 x, y, z, w) passing both thresholds, after hit_ratios()."""
        L = _lib.lib()
        cnt = _lib.c_uint64()
        _lib.check(self._ctx, L.apo_cavity_vertices(self._ctx, hit_ratio_threshold, radius_threshold, None, 0,
                                                    _lib.byref(cnt)))
        out = np.zeros((int(cnt.value), 4), np.float32)
        if len(out):
            _lib.check(self._ctx, L.apo_cavity_vertices(self._ctx, hit_ratio_threshold, radius_threshold,
                                                        _lib.fptr(out), len(out), _lib.byref(cnt)))
        return out

    # -- device state (public members of AlgorithmGPU) ------------------------
    def sorted_sites(self):
        n = len(self.sites)
        xyzr = np.zeros((n, 4), np.float32)
        idx = np.zeros(n, np.uint32)
        _lib.check(self._ctx, _lib.lib().apo_get_sorted_sites(self._ctx, _lib.fptr(xyzr), _lib.uptr(idx)))
        return xyzr, idx

    def lbvh(self) -> dict:
        n = len(self.sites)
        nodes = 2 * n - self.molecules
        out = dict(codes=np.zeros(n, np.uint64), parent=np.zeros(nodes, np.int32),
                   left=np.zeros(nodes, np.int32), right=np.zeros(nodes, np.int32),
                   aabb=np.zeros((nodes, 6), np.float32), sizes=np.zeros(nodes, np.uint32),
                   max_radius=np.zeros(nodes, np.float32))
        _lib.check(self._ctx, _lib.lib().apo_get_lbvh(
            self._ctx, _lib.u64ptr(out["codes"]), _lib.iptr(out["parent"]), _lib.iptr(out["left"]),
            _lib.iptr(out["right"]), _lib.fptr(out["aabb"]), _lib.uptr(out["sizes"]), _lib.fptr(out["max_radius"])))
        return out

    def artificial_sites(self) -> np.ndarray:
        out = np.zeros((6, 4), np.float64)
        _lib.check(self._ctx, _lib.lib().apo_get_artificial_sites(self._ctx, _lib.dptr(out)))
        return out

    def knns(self) -> np.ndarray:
        out = np.zeros((len(self.sites), self.k), np.uint32)
        _lib.check(self._ctx, _lib.lib().apo_get_knns(self._ctx, _lib.uptr(out)))
        return out

    def cell_counts(self) -> dict:
        n = len(self.sites)
        vn = np.zeros(n, np.uint32)
        en = np.zeros(n, np.uint32)
        st = np.zeros(n, np.uint8)
        kr = np.zeros(n, np.float32)
        fl = np.zeros(n, np.uint32)
        _lib.check(self._ctx, _lib.lib().apo_get_cell_counts(self._ctx, _lib.uptr(vn), _lib.uptr(en),
                                                             _lib.u8ptr(st), _lib.fptr(kr), _lib.uptr(fl)))
        return dict(vertex_nb=vn, edge_nb=en, status=st, known_radius=kr, flags=fl)

    def all_cells(self):
        c = self.cell_counts()
        V = np.zeros((int(c["vertex_nb"].sum()), 4), np.uint32)
        E = np.zeros((int(c["edge_nb"].sum()), 4), np.uint32)
        _lib.check(self._ctx, _lib.lib().apo_get_all_cells(self._ctx, _lib.uptr(V), _lib.uptr(E)))
        return c, V, E


    def cell_radii(self) -> np.ndarray:
        """Per-site cell radius of the last build (input order, FP32 rounded
        up): twice it is the smallest search radius that certifies the cell."""
        out = np.zeros(len(self.sites), np.float32)
        _lib.check(self._ctx, _lib.lib().apo_cell_radii(self._ctx, _lib.fptr(out)))
        return out

    def proposal_radii(self) -> np.ndarray:
        """Search radii of the last build's sites (input order) when it ran
        the range-query cell pass (proposal_radius, or radii given)."""
        out = np.zeros(len(self.sites), np.float32)
        _lib.check(self._ctx, _lib.lib().apo_proposal_radii(self._ctx, _lib.fptr(out)))
        return out

    def proposal_classes(self) -> np.ndarray:
        """Classes the learned proposal gave each site in the last build."""
        out = np.zeros(len(self.sites), np.uint8)
        _lib.check(self._ctx, _lib.lib().apo_proposal_classes(self._ctx, _lib.u8ptr(out)))
        return out

    # -- AlgorithmGPU::toTopology ---------------------------------------------
    def topology(self) -> dict:
        """Bisectors (m, 2), trisectors (m, 3), quadrisectors (m, 4) in original
        ids, in the order of AlgorithmGPU::toTopology (algorithm.inl:1732-1821)."""
        L = _lib.lib()
        counts = np.zeros(3, np.uint64)
        cp = counts.ctypes.data_as(_lib._u64p)
        _lib.check(self._ctx, L.apo_get_topology_sets(self._ctx, None, 0, None, 0, None, 0, cp))
        b = np.zeros((int(counts[0]), 2), np.uint32)
        t = np.zeros((int(counts[1]), 3), np.uint32)
        q = np.zeros((int(counts[2]), 4), np.uint32)
        _lib.check(self._ctx, L.apo_get_topology_sets(self._ctx, _lib.uptr(b), len(b), _lib.uptr(t), len(t),
                                                      _lib.uptr(q), len(q), cp))
        return {"bisectors": b, "trisectors": t, "quadrisectors": q}

    def cell_topology(self):
        """Per-site cell records in original order and ids: (vertex_offsets,
        vertices (m, 4), edge_offsets, edges (m, 4))."""
        L = _lib.lib()
        n = len(self.sites)
        vo = np.zeros(n + 1, np.uint64)
        eo = np.zeros(n + 1, np.uint64)
        _lib.check(self._ctx, L.apo_get_cell_topology(self._ctx, vo.ctypes.data_as(_lib._u64p), None, 0,
                                                      eo.ctypes.data_as(_lib._u64p), None, 0))
        v = np.zeros((int(vo[-1]), 4), np.uint32)
        e = np.zeros((int(eo[-1]), 4), np.uint32)
        _lib.check(self._ctx, L.apo_get_cell_topology(self._ctx, vo.ctypes.data_as(_lib._u64p), _lib.uptr(v), len(v),
                                                      eo.ctypes.data_as(_lib._u64p), _lib.uptr(e), len(e)))
        return vo, v, eo, e


def compute_apollonius_diagram(sites, k: int = 16) -> VertexDiagram:
    """apo::gpu::computeApolloniusDiagram (apo/src/gpu/algorithm.cu:6-15)."""
    alg = AlgorithmGPU(sites, k=k)
    try:
        return alg.build().to_host()
    finally:
        alg.close()


def compute_full_apollonius_diagram(sites, k: int = 16) -> FullDiagram:
    """apo::gpu::computeFullApolloniusDiagram (apo/src/gpu/algorithm.cu:17-26)."""
    alg = AlgorithmGPU(sites, k=k)
    try:
        return alg.build().to_host_full()
    finally:
        alg.close()


def topology_json(topology: dict) -> str:
    """The topology document of the reference's topology app (apo::toCsv,
    app/topology/topology.cpp:11-57): a JSON object with "bisectors",
    "trisectors" and "quadrisectors" arrays, one tab-indented entry per line,
    in the order AlgorithmGPU.topology() returns them."""
    out = ["{\n"]
    for key, width, last in (("bisectors", 2, False), ("trisectors", 3, False), ("quadrisectors", 4, True)):
        rows = np.asarray(topology[key], np.uint32).reshape(-1, width)
        out.append(f'\t"{key}": [\n')
        m = len(rows)
        for r, row in enumerate(rows):
            out.append("\t\t[ " + ", ".join(str(int(x)) for x in row) + " ]" + ("" if r == m - 1 else ",") + "\n")
        out.append("\t]\n" if last else "\t],\n")
    out.append("}\n")
    return "".join(out)
