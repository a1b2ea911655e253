"""The learned proposal (neighbour-count classes and the radius head) on the GPU.

* The label pass reproduces the oracle's certified k (the restatement of
  app/benchmarks/benchmark_contributing_nearest.cu:75-113).
* The matrix-core inference (BF16 weights and activations, FP32
  accumulation) agrees with a float64 numpy evaluation of the same
  BF16-rounded network: logits within 2e-2 and the same class wherever the
  top two logits differ by more than that.
* One batch's gradient from the training kernels (BF16 forward pass as in
  inference, FP32 backward pass on the matrix cores) agrees with a float64
  numpy backward pass of the same BF16-rounded network to 1e-4 relative
  (Frobenius norm per tensor); the loss to 1e-4; with and without radius
  targets (some rows without one).
* The 52 features of every site equal a numpy computation from the sites
  and their 16 nearest neighbours (the kNN lists of a build).
* Training lowers the loss and beats the majority class.
* A build with the proposal equals, record for record, the oracle run with
  the same per-site classes, and has no empty-sphere violation; so does a
  build with classes given by the caller (the exact labels' classes) and a
  capped neighbour list (proposal_max_k), and a build with the radius head
  (proposal_radius) against the oracle's range path with the radii the
  head gave (d16 x 2^output, checked against the inference outputs).
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
    W3 = p[o:o + h * 9].reshape(h, 9); o += h * 9
    b3 = p[o:o + 9]
    a = _bf16(x)
    a = _bf16(np.maximum(a @ _bf16(W1) + b1, 0.0))
    a = _bf16(np.maximum(a @ _bf16(W2) + b2, 0.0))
    return a @ _bf16(W3) + b3


def _split(p):
    f, h = proposal.FEATURES, 64
    o, out = 0, []
    for shape in ((f, h), (h,), (h, h), (h,), (h, 9), (9,)):
        size = int(np.prod(shape))
        out.append(p[o:o + size].reshape(shape))
        o += size
    return out


def _numpy_gradient(p, x, y, t=None):
    """Float64 backward pass of the network the training kernels run:
    BF16-rounded weights, features and hidden activations; cross-entropy
    over the 8 logits plus 0.25 x the radius head's squared error against t
    (NaN: no target)."""
    W1, b1, W2, b2, W3, b3 = _split(p)
    W1, W2, W3 = _bf16(W1), _bf16(W2), _bf16(W3)
    a0 = _bf16(x)
    h1 = _bf16(np.maximum(a0 @ W1 + b1, 0.0))
    h2 = _bf16(np.maximum(h1 @ W2 + b2, 0.0))
    out = h2 @ W3 + b3
    z = out[:, :8] - out[:, :8].max(axis=1, keepdims=True)
    e = np.exp(z)
    prob = e / e.sum(axis=1, keepdims=True)
    n = len(y)
    loss_rows = -np.log(prob[np.arange(n), y])
    dz = np.zeros_like(out)
    dz[:, :8] = prob
    dz[np.arange(n), y] -= 1.0
    if t is not None:
        has = ~np.isnan(t)
        err = np.where(has, out[:, 8] - np.where(has, t, 0.0), 0.0)
        loss_rows = loss_rows + 0.25 * err * err
        dz[:, 8] = 2.0 * 0.25 * err
    loss = float(np.mean(loss_rows))
    dz /= n
    d2 = (dz @ W3.T) * (h2 > 0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    grads = [a0.T @ d1, d1.sum(0), h1.T @ d2, d2.sum(0), h2.T @ dz, dz.sum(0)]
    return grads, loss


@pytest.mark.parametrize("with_radius", [False, True])
def test_training_gradient_matches_float64(training, with_radius):
    sites, feat, lab = training
    rng = np.random.default_rng(3)
    params = rng.normal(0.0, 0.2, proposal.PARAMS).astype(np.float32)
    keep = np.nonzero(lab != proposal.LABEL_BURIED)[0][:4000]  # 250 tiles of 16 rows, the last one partial
    x = feat[keep][:3990]
    y = proposal.label_class(lab[keep][:3990])
    t = None
    if with_radius:
        t = rng.normal(0.5, 1.0, len(y)).astype(np.float32)
        t[rng.random(len(y)) < 0.2] = np.nan  # rows without a target
    g, loss = proposal.gradient(params, x, y, t)
    ref, ref_loss = _numpy_gradient(params.astype(np.float64), x.astype(np.float64), y.astype(np.int64),
                                    None if t is None else t.astype(np.float64))
    _say(f"[grad] loss {loss:.6f} (float64 {ref_loss:.6f})")
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss)
    for name, got, want in zip(("W1", "b1", "W2", "b2", "W3", "b3"), _split(g.astype(np.float64)), ref):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        _say(f"[grad] {name}: relative error {err:.2e}")
        assert err < 1e-4, name


def test_features_match_numpy(training):
    sites, feat, lab = training
    g = AlgorithmGPU(sites, k=16, stage_mask=8).build()  # the LBVH and the 16-NN lists only
    try:
        s, perm = g.sorted_sites()
        knn = g.knns()
    finally:
        g.close()
    n = len(s)
    want = np.zeros((n, proposal.FEATURES), np.float32)
    f32 = np.float32
    for i in range(n):
        si = s[i]
        nb = knn[i]
        d16 = f32(0)
        for t in range(15, -1, -1):
            if nb[t] < n:
                dv = s[nb[t], :3] - si[:3]
                d16 = np.sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
                break
        inv = f32(1) / d16 if d16 > 0 else f32(0)
        row = want[i]
        row[0] = si[3] * inv
        larger, rsum, have = f32(0), f32(0), 0
        gmin, gmax = f32(np.finfo(np.float32).max), f32(-np.finfo(np.float32).max)
        for t in range(16):
            j = nb[t]
            if j >= n:
                continue
            o = s[j]
            dv = o[:3] - si[:3]
            d = np.sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
            gg = ((d - si[3]) - o[3]) * inv
            row[1 + 3 * t], row[2 + 3 * t], row[3 + 3 * t] = d * inv, o[3] * inv, gg
            larger += f32(1) if o[3] > si[3] else f32(0)
            rsum += o[3]
            gmin, gmax = min(gmin, gg), max(gmax, gg)
            have += 1
        row[49] = larger / f32(have) if have else 0
        lr = f32(0)
        if have and rsum > 0 and si[3] > 0:
            lr = f32(np.log2(si[3] / (rsum / f32(have)))) * f32(0.125)
        row[50] = min(f32(1), max(f32(-1), lr))
        row[51] = gmax - gmin if have else 0
    got = feat[perm]  # input order -> sorted order
    assert np.allclose(got, want, rtol=1e-6, atol=1e-6)


def test_mfma_inference_matches_numpy(training):
    sites, feat, lab = training
    rng = np.random.default_rng(0)
    params = rng.normal(0.0, 0.3, proposal.PARAMS).astype(np.float32)
    x = feat[:4000]
    cls, logits = proposal.infer(params, x)
    ref = _numpy_logits(params.astype(np.float64), x)
    assert logits.shape == (len(x), proposal.OUTPUTS)
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1.0
    assert np.all(np.abs(logits - ref) <= 2e-2 * scale)
    srt = np.sort(ref[:, :8], axis=1)
    clear = (srt[:, -1] - srt[:, -2]) > 4e-2 * scale[:, 0]
    assert np.array_equal(cls[clear], np.argmax(ref[:, :8], axis=1)[clear])


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


def test_radius_head_build_matches_oracle(training):
    """Radius head: targets from a completed plain build, a short training,
    then a build with proposal_radius: its radii are d16 x 2^output of the
    inference pass, and every record equals the oracle's range path run with
    those radii."""
    sites, feat, lab = training
    _say("[case] radius targets, 20k log-uniform sites")
    t = proposal.radius_targets(sites)
    assert np.isnan(t).sum() == (lab == proposal.LABEL_BURIED).sum()
    params, loss = proposal.train(feat, lab, t, epochs=6, batch=1024, lr=3e-3, seed=3)
    _say(f"[radius] loss per epoch {np.round(loss, 4).tolist()}")
    assert loss[-1] < loss[0]
    test_sites = datasets.loguniform_sites(20_000, seed=53)
    g = AlgorithmGPU(test_sites, proposal=params, proposal_radius=1, proposal_max_k=64,
                     proposal_fallback_k=64).build()
    try:
        radii = g.proposal_radii()
        # the same radii from the inference outputs and the sites' d16
        tf, _ = proposal.labels(test_sites)
        _, out = proposal.infer(params, tf)
        tt = proposal.radius_targets(test_sites)
        live = ~np.isnan(tt)
        err = np.abs(out[live, 8] - tt[live])
        _say(f"[radius] head error (log2): median {np.median(err):.3f}, 90% {np.quantile(err, 0.9):.3f}")
        st = g.stats()
        _say(f"[radius] certified by the cell pass {st['cells_certified']} of {len(test_sites)}")
        orc = oracle.Diagram(test_sites, radii=radii)
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
    finally:
        g.close()
