"""The learned neighbour-count proposal on the GPU.

* The label pass reproduces the oracle's certified k (the restatement of
  app/benchmarks/benchmark_contributing_nearest.cu:75-113).
* The matrix-core inference (BF16 weights and activations, FP32
  accumulation) agrees with a float64 numpy evaluation of the same
  BF16-rounded network: logits within 2e-2 and the same class wherever the
  top two logits differ by more than that.
* Training lowers the loss and beats the majority class.
* A build with the proposal equals, record for record, the oracle run with
  the same per-site classes, and has no empty-sphere violation; so does a
  build with classes given by the caller (the exact labels' classes) and a
  capped neighbour list (proposal_max_k).
"""
import numpy as np
import pytest

import oracle
from apo_amd import AlgorithmGPU, datasets, proposal

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def training():
    sites = datasets.loguniform_sites(20_000, seed=51)
    _say("[case] label pass, 20k log-uniform sites")
    feat, lab = proposal.labels(sites)
    return sites, feat, lab


def test_label_pass_matches_oracle(training):
    sites, feat, lab = training
    want = oracle.certified_k(sites, kmax=64).astype(np.int64)
    # oracle: 0 buried, 0xffffffff never; GPU: 255 buried, 0 never
    got = lab.astype(np.int64)
    got_o = np.where(got == 255, 0, np.where(got == 0, 0xFFFFFFFF, got))
    assert np.array_equal(got_o, want)
    assert np.all(np.isfinite(feat))


def _bf16(x):
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _numpy_logits(p, x):
    f, h = proposal.FEATURES, 64
    o = 0
    W1 = p[o:o + f * h].reshape(f, h); o += f * h
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    W3 = p[o:o + h * 8].reshape(h, 8); o += h * 8
    b3 = p[o:o + 8]
    a = _bf16(x)
    a = _bf16(np.maximum(a @ _bf16(W1) + b1, 0.0))
    a = _bf16(np.maximum(a @ _bf16(W2) + b2, 0.0))
    return a @ _bf16(W3) + b3


def test_mfma_inference_matches_numpy(training):
    sites, feat, lab = training
    rng = np.random.default_rng(0)
    params = rng.normal(0.0, 0.3, proposal.PARAMS).astype(np.float32)
    x = feat[:4000]
    cls, logits = proposal.infer(params, x)
    ref = _numpy_logits(params.astype(np.float64), x)
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1.0
    assert np.all(np.abs(logits - ref) <= 2e-2 * scale)
    srt = np.sort(ref, axis=1)
    clear = (srt[:, -1] - srt[:, -2]) > 4e-2 * scale[:, 0]
    assert np.array_equal(cls[clear], np.argmax(ref, axis=1)[clear])


def test_training_lowers_loss(training):
    sites, feat, lab = training
    params, loss = proposal.train(feat, lab, epochs=6, batch=1024, lr=3e-3, seed=1)
    _say(f"[train] loss per epoch {np.round(loss, 4).tolist()}")
    assert np.all(np.isfinite(loss)) and loss[-1] < loss[0]
    keep = lab != proposal.LABEL_BURIED
    cls, _ = proposal.infer(params, feat[keep])
    acc = float(np.mean(cls == proposal.label_class(lab[keep])))
    base = float(np.max(np.bincount(proposal.label_class(lab[keep]), minlength=8))) / keep.sum()
    _say(f"[train] accuracy {acc:.3f} (majority class {base:.3f})")
    assert acc > base


def test_learned_build_matches_oracle(training):
    sites, feat, lab = training
    params, _ = proposal.train(feat, lab, epochs=4, batch=1024, lr=3e-3, seed=2)
    test_sites = datasets.loguniform_sites(20_000, seed=52)
    _say("[case] learned build, 20k log-uniform sites")
    g = AlgorithmGPU(test_sites, proposal=params).build()
    cls = g.proposal_classes()
    st = g.stats()
    _say(f"[learned] classes {np.bincount(cls, minlength=8).tolist()} certified by the cell pass "
         f"{st['cells_certified']}")
    orc = oracle.Diagram(test_sites, classes=cls, edge_bound=0)
    gc, gV, gE = g.all_cells()
    oc, oV, oE = orc.all_cells()
    for key in ("status", "vertex_nb", "edge_nb", "flags"):
        assert np.array_equal(gc[key], oc[key]), key
    assert np.array_equal(gV, oV) and np.array_equal(gE, oE)
    gx = g.to_host()
    ox, oi = orc.vertices()
    assert np.array_equal(gx.vertices_id, oi)
    assert np.array_equal(gx.vertices.view(np.uint64), ox.view(np.uint64))
    assert g.empty_sphere_violations() == 0
    g.close()


def test_given_classes_and_capped_lists_match_oracle(training):
    sites, feat, lab = training
    _say("[case] learned build with the labels' classes, lists capped at 32")
    cls = proposal.label_class(lab)
    cls[cls > 3] = 7  # classes of more than 32 neighbours become the last class
    g = AlgorithmGPU(sites, classes=cls, proposal_max_k=32, proposal_fallback_k=32).build()
    try:
        assert np.array_equal(g.proposal_classes(), cls)
        st = g.stats()
        # every cell the labels certify within 32 neighbours is certified by the cell pass
        assert st["cells_certified"] == int(np.sum((lab >= 1) & (lab <= 32)))
        orc = oracle.Diagram(sites, classes=cls, edge_bound=0, fallback_k=32)
        gc, gV, gE = g.all_cells()
        oc, oV, oE = orc.all_cells()
        for key in ("status", "vertex_nb", "edge_nb", "flags"):
            assert np.array_equal(gc[key], oc[key]), key
        assert np.array_equal(gV, oV) and np.array_equal(gE, oE)
        gx = g.to_host()
        ox, oi = orc.vertices()
        assert np.array_equal(gx.vertices_id, oi)
        assert np.array_equal(gx.vertices.view(np.uint64), ox.view(np.uint64))
    finally:
        g.close()
