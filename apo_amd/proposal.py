"""Learned proposal: labels, training and inference.

An MLP (52 -> 64 -> 64 -> 9) predicts for each site how many nearest
neighbours its cell needs before the plain security-radius test certifies
it (8 class logits) and the search radius it needs (output 8: log2 of the
radius over d16, the centre distance of its 16th nearest neighbour); builds
with the proposal run that many updates per cell (or, with
proposal_radius, every site below the radius through a range query) and
send the cells the test does not certify through the plain validation.
Class labels are the reference's certified k (app/benchmarks/
benchmark_contributing_nearest.cu:75-113), computed by the library's label
pass on the GPU; radius targets are twice each cell's final radius (the
radius that certifies it) from a completed plain build; training runs in
HIP kernels (apo_proposal_train).
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib

FEATURES = 52
CLASSES = 8
OUTPUTS = 9   # the class logits, then the radius head
PARAMS = 8137
# neighbours asked for by each class; the last (not certified within 64) runs 16
CLASS_K = (20, 24, 28, 32, 40, 48, 64, 16)
LABEL_BURIED = 255

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(_HERE, "models", "proposal_c5.f32")


def label_class(labels: np.ndarray) -> np.ndarray:
    """Class of each label (neighbours used at certification, 0 = never)."""
    labels = np.asarray(labels)
    out = np.full(labels.shape, CLASSES - 1, np.uint8)
    for c in range(CLASSES - 2, -1, -1):
        out[(labels >= 1) & (labels <= CLASS_K[c])] = c
    return out


def _ctx(k: int = 16):
    cfg = _lib.Config()
    _lib.lib().apo_default_config(_lib.byref(cfg))
    cfg.k = k
    return _lib.create(cfg)


def labels(sites: np.ndarray):
    """(features (n, 52), labels (n,)) of a site set, in input order."""
    s = np.ascontiguousarray(np.asarray(sites, np.float32).reshape(-1, 4))
    n = len(s)
    feat = np.zeros((n, FEATURES), np.float32)
    lab = np.zeros(n, np.uint8)
    ctx = _ctx()
    try:
        _lib.check(ctx, _lib.lib().apo_proposal_labels(ctx, _lib.fptr(s), n, _lib.fptr(feat), _lib.u8ptr(lab)))
    finally:
        _lib.lib().apo_destroy(ctx)
    return feat, lab


# Radius-head targets are clipped to [RADIUS_LO, RADIUS_HI] (log2 of the
# radius over d16): a range of 2 d16 already holds about 128 sites, more
# than any neighbour list, so larger radii all mean "not certified by a
# list" and need no finer prediction.
RADIUS_LO, RADIUS_HI = -3.0, 1.0


def radius_targets(sites: np.ndarray, margin: float = 1e-3) -> np.ndarray:
    """Radius-head targets of a site set (input order): log2 of twice each
    cell's final radius, (1 + margin) above, over d16, clipped to
    [RADIUS_LO, RADIUS_HI]; NaN for buried sites.  One plain build with
    completion (apo_cell_radii) and its 16 nearest."""
    from . import AlgorithmGPU
    s = np.ascontiguousarray(np.asarray(sites, np.float32).reshape(-1, 4))
    g = AlgorithmGPU(s, k=16, complete=4).build()
    try:
        cr = g.cell_radii()
        knn = g.knns()
        ss, idx = g.sorted_sites()
    finally:
        g.close()
    # d16: centre distance of each sorted site's last listed neighbour
    have = knn < len(s)
    last = np.where(have.any(axis=1), have.shape[1] - 1 - np.argmax(have[:, ::-1], axis=1), 0)
    j = knn[np.arange(len(s)), last]
    ok = have.any(axis=1)
    d = np.zeros(len(s), np.float32)
    d[ok] = np.linalg.norm(ss[j[ok], :3] - ss[ok, :3], axis=1)
    d16 = np.zeros(len(s), np.float32)
    d16[idx] = d
    t = np.full(len(s), np.nan, np.float32)
    live = (cr > -np.finfo(np.float32).max) & (d16 > 0)
    t[live] = np.clip(np.log2(2.0 * np.maximum(cr[live].astype(np.float64), 1e-30) * (1.0 + margin) / d16[live]),
                      RADIUS_LO, RADIUS_HI)
    return t


def train(features: np.ndarray, labels_: np.ndarray, radius: np.ndarray | None = None, epochs: int = 8,
          batch: int = 2048, lr: float = 3e-3, seed: int = 0, init: np.ndarray | None = None):
    """Train on the GPU; returns (params (8137,), mean loss per epoch).
    radius: radius-head targets (radius_targets), or None for none."""
    f = np.ascontiguousarray(features, np.float32)
    y = np.ascontiguousarray(labels_, np.uint8)
    r = None if radius is None else np.ascontiguousarray(radius, np.float32)
    params = np.zeros(PARAMS, np.float32)
    loss = np.zeros(epochs, np.float32)
    ini = None if init is None else np.ascontiguousarray(init, np.float32)
    ctx = _ctx()
    try:
        _lib.check(ctx, _lib.lib().apo_proposal_train(ctx, _lib.fptr(f), _lib.u8ptr(y),
                                                       None if r is None else _lib.fptr(r), len(y), epochs, batch,
                                                       lr, seed, None if ini is None else _lib.fptr(ini),
                                                       _lib.fptr(params), _lib.fptr(loss)))
    finally:
        _lib.lib().apo_destroy(ctx)
    return params, loss


def gradient(params: np.ndarray, features: np.ndarray, classes: np.ndarray, radius: np.ndarray | None = None):
    """(gradient (8137,), mean loss) of the training loss over the rows at
    `params` (cross-entropy, plus the radius head's weighted squared error
    where `radius` has a target), from the training kernels."""
    p = np.ascontiguousarray(params, np.float32)
    f = np.ascontiguousarray(features, np.float32)
    c = np.ascontiguousarray(classes, np.uint8)
    r = None if radius is None else np.ascontiguousarray(radius, np.float32)
    g = np.zeros(PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    ctx = _ctx()
    try:
        _lib.check(ctx, _lib.lib().apo_proposal_gradient(ctx, _lib.fptr(p), _lib.fptr(f), _lib.u8ptr(c),
                                                          None if r is None else _lib.fptr(r), len(c),
                                                          _lib.fptr(g), _lib.fptr(loss)))
    finally:
        _lib.lib().apo_destroy(ctx)
    return g, float(loss[0])


def infer(params: np.ndarray, features: np.ndarray):
    """Classes and the 9 outputs (8 logits, radius head) of feature rows
    through the matrix-core inference."""
    p = np.ascontiguousarray(params, np.float32)
    f = np.ascontiguousarray(features, np.float32)
    cls = np.zeros(len(f), np.uint8)
    logits = np.zeros((len(f), OUTPUTS), np.float32)
    ctx = _ctx()
    try:
        L = _lib.lib()
        _lib.check(ctx, L.apo_proposal_set(ctx, _lib.fptr(p)))
        _lib.check(ctx, L.apo_proposal_infer(ctx, _lib.fptr(f), len(f), _lib.u8ptr(cls), _lib.fptr(logits)))
    finally:
        _lib.lib().apo_destroy(ctx)
    return cls, logits


def load(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    """Parameters from a raw float32 file (no pickle).  A JSON header beside
    it (same name, .json) records the training recipe and a SHA-256 of the
    parameters, which is checked."""
    p = np.fromfile(path, dtype="<f4")
    if p.size != PARAMS:
        raise ValueError(f"{path}: expected {PARAMS} float32 parameters, found {p.size}")
    header = os.path.splitext(path)[0] + ".json"
    if os.path.exists(header):
        import hashlib
        import json
        with open(header) as f:
            want = json.load(f).get("sha256")
        if want and hashlib.sha256(p.tobytes()).hexdigest() != want:
            raise ValueError(f"{path}: parameters do not match the checksum in {header}")
    return p


def save(params: np.ndarray, path: str = DEFAULT_WEIGHTS) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.ascontiguousarray(params, "<f4").tofile(path)
