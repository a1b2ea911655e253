#!/usr/bin/env python3
"""Throughput of the gfx950 Apollonius-diagram construction on one MI355X.

Workload: 1,000,000 sites, centres uniform in [0,1]^3, radii log-uniform in
[1e-4, 1e-2] (FP32, seed 3).  One step = one full construction (scene box,
LBVH, kNN, cell update, vertex/edge/bisector validation, vertex output
compaction) from sites resident in HBM to the compacted vertex buffer in
HBM.  value = cells per second over the timed steps, each one replay of the
build captured in a HIP graph; the eager steps before them give the stage
split and the per-kernel times of the roofline.

Prints one JSON line.  --gpus above 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=60_000, help="sites in the CPU-baseline sub-cube")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal comparison")
    ap.add_argument("--proposal-cost", type=float, default=None, help="override the tuned proposal cost")
    ap.add_argument("--no-c5", action="store_true", help="skip the 5M-site learned evaluation")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-molecules workload")
    ap.add_argument("--lbvh-only", action="store_true",
                    help="only the LBVH-alone timings at 1M, 4M and 16M sites (for kernel traces)")
    ap.add_argument("--batch-molecules", type=int, default=512)
    return ap.parse_args()


def _vertices(ctx, L, _lib, v):
    xyzw = np.zeros((v, 4), np.float64)
    ids = np.zeros((v, 4), np.uint32)
    _lib.check(ctx, L.apo_get_vertices(ctx, _lib.dptr(xyzw), _lib.uptr(ids), v))
    return xyzw, ids


def graph_replay(dev, n, tuning, steps, warmup, say):
    """The same build captured once in a HIP graph and replayed: one graph
    launch per step, bracketed by device synchronisations."""
    import torch
    from apo_amd import _lib
    ctx = _ctx(tuning, graph=True)
    L = _lib.lib()
    ptr = _lib.ctypes.c_void_p(dev.data_ptr())
    for _ in range(max(warmup, 1)):  # the first build captures the graph
        _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
    wall, dev_ms = 0.0, 0.0
    s = _lib.Stats()
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(s)))
        dev_ms += s.ms_total
    ms = 1000.0 * wall / steps
    say(f"graph replay: {ms:.1f} ms per step")
    L.apo_destroy(ctx)
    return {"cells_per_s": n / (ms / 1000.0), "ms_per_step": ms, "device_ms_per_replay": dev_ms / steps,
            "vertices": int(s.vertices)}


def batch_workload(count, steps, warmup, say):
    """512 independent protein-like molecules, atom counts log-uniform in
    [500, 50000] (seed 4), in one batched build: offset-indexed LBVH and
    persistent work-queue kernels, captured in a HIP graph and replayed.
    Checks: the graph replay equals the eager build, and a spread of
    molecules equals their own single-molecule builds, vertex for vertex."""
    import torch
    from apo_amd import _lib, datasets
    t0 = time.perf_counter()
    sites, offsets = datasets.batched_molecules(count=count, seed=4, lo=500, hi=50000)
    n = len(sites)
    say(f"batch: {count} molecules, {n} atoms generated in {time.perf_counter() - t0:.1f} s")
    dev = torch.from_numpy(sites).to("cuda")
    torch.cuda.synchronize()
    L = _lib.lib()
    ptr = _lib.ctypes.c_void_p(dev.data_ptr())
    off = np.ascontiguousarray(offsets, np.uint32)

    from apo_amd import autotune
    rec = autotune.load("batch_128")
    tuning = dict(rec["plain"]) if rec else {}

    def make(graph):
        return _ctx(tuning, graph=graph)

    # eager build: stage split
    ctx = make(False)
    _lib.check(ctx, L.apo_build_batched_device(ctx, ptr, _lib.uptr(off), count))
    s = _lib.Stats()
    _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(s)))
    stage = {k: float(getattr(s, k)) for k in ("ms_total", "ms_lbvh", "ms_knn", "ms_cells", "ms_validation",
                                                 "ms_bisector", "ms_output")}
    v_eager = int(s.vertices)
    x_e, i_e = _vertices(ctx, L, _lib, v_eager)
    vo = np.zeros(count + 1, np.uint64)
    _lib.check(ctx, L.apo_batch_vertex_offsets(ctx, _lib.u64ptr(vo)))
    cnt = _lib.c_uint64()
    _lib.check(ctx, L.apo_check_empty_spheres(ctx, None, None, 0, 1e-9, _lib.byref(cnt)))
    L.apo_destroy(ctx)
    say(f"batch eager build: {stage['ms_total']:.1f} ms, {v_eager} vertices")

    # graph replays: the timed steps
    gctx = make(True)
    for _ in range(max(warmup, 1)):
        _lib.check(gctx, L.apo_build_batched_device(gctx, ptr, _lib.uptr(off), count))
    wall, dev_ms = 0.0, 0.0
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(gctx, L.apo_build_batched_device(gctx, ptr, _lib.uptr(off), count))
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        _lib.check(gctx, L.apo_get_stats(gctx, _lib.byref(s)))
        dev_ms += s.ms_total
    ms = 1000.0 * wall / steps
    x_g, i_g = _vertices(gctx, L, _lib, int(s.vertices))
    L.apo_destroy(gctx)
    same_graph = bool(np.array_equal(i_g, i_e) and np.array_equal(x_g.view(np.uint64), x_e.view(np.uint64)))
    say(f"batch graph replay: {ms:.1f} ms per step, identical to eager: {same_graph}")

    # spot check: molecules across the size range against single builds
    sizes = np.diff(offsets.astype(np.int64))
    order = np.argsort(sizes)
    picks = sorted({int(order[q]) for q in np.linspace(0, count - 1, 6).astype(int)})
    single = make(False)
    matched = 0
    for m in picks:
        a, b = int(offsets[m]), int(offsets[m + 1])
        sub = np.ascontiguousarray(sites[a:b])
        _lib.check(single, L.apo_build(single, _lib.fptr(sub), b - a))
        sv = _lib.c_uint64()
        _lib.check(single, L.apo_vertex_count(single, _lib.byref(sv)))
        xs, is_ = _vertices(single, L, _lib, int(sv.value))
        xb, ib = x_e[vo[m]:vo[m + 1]], i_e[vo[m]:vo[m + 1]] - np.uint32(a)
        matched += int(np.array_equal(is_, ib) and np.array_equal(xs.view(np.uint64), xb.view(np.uint64)))
    L.apo_destroy(single)
    say(f"batch spot check: {matched} of {len(picks)} molecules equal their single builds")
    return {"workload": f"{count} protein-like molecules (C2 generator per molecule, seed (4, m)), atom counts "
                        f"log-uniform [500, 50000], {n} atoms, FP32 A",
            "molecules": count, "sites": n, "cells_per_s": n / (ms / 1000.0), "molecules_per_s": count / (ms / 1000.0),
            "ms_per_step": ms, "device_ms_per_replay": dev_ms / steps, "steps": steps, "path": "HIP graph replay",
            "eager_stage_ms": stage, "vertices": v_eager, "vertices_per_s": v_eager / (ms / 1000.0),
            "graph_equals_eager": same_graph, "spot_molecules": picks, "spot_equal_single_build": matched,
            "empty_sphere_violations": int(cnt.value), "config": tuning,
            "note": "violations are the reference method's own on overlapping atoms (counted, as its precision check does)"}


def _ctx(tuning, proposal=None, graph=False):
    """A context with the default config plus the tuned fields."""
    from apo_amd import _lib
    cfg = _lib.Config()
    _lib.lib().apo_default_config(_lib.byref(cfg))
    for key, val in (tuning or {}).items():
        setattr(cfg, key, val)
    cfg.graph = 1 if graph else 0
    ctx = _lib.create(cfg)
    if proposal is not None:
        p32 = np.ascontiguousarray(proposal, np.float32)
        _lib.check(ctx, _lib.lib().apo_proposal_set(ctx, _lib.fptr(p32)))
    return ctx


def _row_keys(ids):
    """Vertex tangent-site quadruples as sortable 16-byte keys."""
    a = np.ascontiguousarray(np.sort(ids, axis=1))
    return a.view(np.dtype((np.void, 16))).ravel()


def _vertex_rows(xyzw, ids):
    """Every emitted vertex as (4 ids, FP64 bits of x, y, z, w), in a
    canonical order: equal arrays mean identical vertex sets, positions
    included."""
    a = np.concatenate([ids.astype(np.uint64), np.ascontiguousarray(xyzw).view(np.uint64)], axis=1)
    return a[np.lexsort(a.T[::-1])]


def _completion_stats(s):
    return {"ms": s.ms_completion, "rounds": int(s.completion_rounds),
            "updated_cells": int(s.completion_updated_cells), "unrepairable": int(s.completion_unrepairable),
            "validation_tasks": int(s.completion_validation_tasks)}


def completion_section(dev, n, plain_tuning, steps, warmup, say):
    """The plain build with the completion mode (config.complete = 4): one
    HIP graph replay per step, then an eager build for its counts, the
    FP64 empty-sphere check and the interval validator.  Returns the section
    and the vertex rows the learned path is compared with."""
    import torch
    from apo_amd import _lib
    tuning = dict(plain_tuning, complete=4)
    gr = graph_replay(dev, n, tuning, steps, warmup, say)
    ctx = _ctx(tuning)
    L = _lib.lib()
    ptr = _lib.ctypes.c_void_p(dev.data_ptr())
    for _ in range(max(warmup, 1)):
        _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
    wall = 0.0
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
    eager_ms = 1000.0 * wall / steps
    s = _lib.Stats()
    _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(s)))
    xyzw, ids = _vertices(ctx, L, _lib, int(s.vertices))
    cnt = _lib.c_uint64()
    _lib.check(ctx, L.apo_check_empty_spheres(ctx, None, None, 0, 1e-9, _lib.byref(cnt)))
    prec = _precision(ctx)
    cell_radii = np.zeros(n, np.float32)
    _lib.check(ctx, L.apo_cell_radii(ctx, _lib.fptr(cell_radii)))
    L.apo_destroy(ctx)
    torch.cuda.synchronize()
    say(f"completion: {gr['ms_per_step']:.1f} ms per graph replay, {int(s.vertices)} vertices, "
        f"{int(cnt.value)} violations")
    return ({"cells_per_s": gr["cells_per_s"], "ms_per_step": gr["ms_per_step"], "vertices": int(s.vertices),
             "eager": {"cells_per_s": n / (eager_ms / 1000.0), "ms_per_step": eager_ms},
             "empty_sphere_violations": int(cnt.value), "precision": prec, "completion": _completion_stats(s),
             "config": tuning, "path": "plain + completion, one HIP graph replay per step"},
            _vertex_rows(xyzw, ids), _row_keys(ids), cell_radii)


def train_proposal(say):
    """The committed proposal (apo_amd/models/proposal_c5.f32), and its recipe
    run again: labels from the plain method on two 1M log-uniform sets
    (seeds 501/502: 2M labelled sites) and 8 epochs of the HIP training
    kernels, timed; the training is deterministic, so the parameters must
    come out bit for bit as committed."""
    from apo_amd import proposal
    from apo_amd.models import train_c5
    committed = proposal.load()
    params, loss, t_labels, t_train = train_c5.train()
    same = bool(np.array_equal(params.view(np.uint32), committed.view(np.uint32)))
    say(f"proposal trained on 2M labelled sites: labels {t_labels:.1f} s, training {t_train:.2f} s, "
        f"loss {loss[0]:.3f} -> {loss[-1]:.3f}, equal to the committed weights: {same}")
    return committed, {"labelled_sites": 2_000_000, "epochs": 8, "loss_first_epoch": float(loss[0]),
                       "loss_last_epoch": float(loss[-1]), "label_s": t_labels, "train_s": t_train,
                       "weights": os.path.relpath(proposal.DEFAULT_WEIGHTS, ROOT), "retrained_equals_committed": same}


def _precision(ctx):
    """Interval validator (app/precision/precision.cpp:14-163) on the build's
    vertices and closed edges."""
    from apo_amd import _lib
    r = _lib.Precision()
    _lib.check(ctx, _lib.lib().apo_precision_check(ctx, _lib.byref(r)))
    return {name: int(getattr(r, name)) for name, _ in r._fields_}


def run_learned(params, dev, n, tuning, steps, warmup, plain_keys, classes=None, plain_rows=None, radii=None):
    """Time the workload with the proposal.  Cells the plain test does not
    certify after the proposed count go through the plain validation.
    classes: per-site classes given instead of the MLP's.  radii: per-site
    search radii (range-query cell pass) instead of either.  plain_rows: the
    plain build's vertex rows (_vertex_rows) for an exact comparison."""
    import torch
    from apo_amd import _lib
    ctx = _ctx(tuning, proposal=params)
    if classes is not None:
        cls = np.ascontiguousarray(classes, np.uint8)
        _lib.check(ctx, _lib.lib().apo_proposal_set_classes(ctx, _lib.u8ptr(cls), len(cls)))
    if radii is not None:
        rad = np.ascontiguousarray(radii, np.float32)
        _lib.check(ctx, _lib.lib().apo_proposal_set_radii(ctx, _lib.fptr(rad), len(rad)))
    L = _lib.lib()
    ptr = _lib.ctypes.c_void_p(dev.data_ptr())
    for _ in range(warmup):
        _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
    wall = 0.0
    s = _lib.Stats()
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
    _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(s)))
    v = int(s.vertices)
    xyzw, ids = _vertices(ctx, L, _lib, v)
    cnt = _lib.c_uint64()
    _lib.check(ctx, L.apo_check_empty_spheres(ctx, None, None, 0, 1e-9, _lib.byref(cnt)))
    prec = _precision(ctx)
    L.apo_destroy(ctx)
    keys = _row_keys(ids)
    same = None if plain_rows is None else bool(np.array_equal(_vertex_rows(xyzw, ids), plain_rows))
    ms = 1000.0 * wall / steps
    non_buried = n - int(s.buried)
    return {"cells_per_s": n / (ms / 1000.0), "ms_per_step": ms, "config": dict(tuning or {}),
            "cells_certified_by_cell_pass": int(s.cells_certified),
            "fallback_rate": 1.0 - int(s.cells_certified) / max(non_buried, 1),
            "stage_ms": {"knn": s.ms_knn, "proposal": s.ms_proposal, "cells": s.ms_cells,
                         "validation": s.ms_validation, "bisector": s.ms_bisector},
            "vertices": v, "vertices_only_in_plain": int(len(np.setdiff1d(plain_keys, keys))),
            "vertices_only_in_learned": int(len(np.setdiff1d(keys, plain_keys))),
            "identical_vertex_rows": same,
            "empty_sphere_violations": int(cnt.value), "precision": prec,
            "completion": _completion_stats(s) if (tuning or {}).get("complete") else None}


def c5_evaluation(params, plain_tuning, learned_tuning, say, n=5_000_000):
    """Inference + certified construction on 5M sites (uniform centres,
    log-uniform radii [1e-4, 1e-2], seed 55) with the proposal trained on
    the 2M labelled sites, beside the plain path on the same set."""
    import torch
    from apo_amd import _lib, datasets
    sites = datasets.loguniform_sites(n, seed=55)
    dev = torch.from_numpy(sites).to("cuda")
    torch.cuda.synchronize()
    L = _lib.lib()
    ptr = _lib.ctypes.c_void_p(dev.data_ptr())
    ctx = _ctx(plain_tuning)
    _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
    torch.cuda.synchronize()
    ms_plain = 1000.0 * (time.perf_counter() - t0)
    s = _lib.Stats()
    _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(s)))
    _, ids = _vertices(ctx, L, _lib, int(s.vertices))
    cnt = _lib.c_uint64()
    _lib.check(ctx, L.apo_check_empty_spheres(ctx, None, None, 0, 1e-9, _lib.byref(cnt)))
    plain_violations = int(cnt.value)
    plain_precision = _precision(ctx)
    L.apo_destroy(ctx)
    say(f"C5 plain: {ms_plain:.1f} ms for {n} sites, {plain_violations} empty-sphere violations")
    lr = run_learned(params, dev, n, learned_tuning, 1, 1, _row_keys(ids))
    say(f"C5 learned: {lr['ms_per_step']:.1f} ms, fallback rate {lr['fallback_rate']:.3f}")
    lr["plain_cells_per_s"] = n / (ms_plain / 1000.0)
    lr["speedup_vs_plain"] = lr["cells_per_s"] / lr["plain_cells_per_s"]
    lr["plain_empty_sphere_violations"] = plain_violations
    lr["plain_precision"] = plain_precision
    lr["workload"] = f"{n} sites, centres U[0,1]^3, radii log-uniform [1e-4,1e-2], seed 55"
    # completion mode on the same set: plain, then learned against its rows
    comp, crows, ckeys, _ = completion_section(dev, n, plain_tuning, 1, 1, say)
    lcc = run_learned(params, dev, n, dict(learned_tuning, complete=4), 1, 1, ckeys, plain_rows=crows)
    lcc["speedup_vs_plain"] = lcc["cells_per_s"] / comp["eager"]["cells_per_s"]
    lr["complete"] = comp
    lr["learned_complete"] = lcc
    say(f"C5 completion: plain {comp['eager']['ms_per_step']:.1f} ms, {comp['empty_sphere_violations']} violations; "
        f"learned {lcc['ms_per_step']:.1f} ms, identical rows {lcc['identical_vertex_rows']}")
    del dev
    torch.cuda.empty_cache()
    return lr


def cavity_section(ctx, vertices, say):
    """Hit ratios of every vertex of the 1M build (64 rays each through the
    LBVH, app/cavity/utils.cu:17-65) and the reference app's default filter
    (hit ratio above 0.5, radius above 5: nothing at this scale) and one at
    this workload's scale."""
    import torch
    from apo_amd import _lib
    L = _lib.lib()
    _lib.check(ctx, L.apo_cavity_hit_ratios(ctx, 64, None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(ctx, L.apo_cavity_hit_ratios(ctx, 64, None))
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0)
    kept = {}
    for hit, rad in ((0.5, 5.0), (0.9, 0.004)):
        c = _lib.c_uint64()
        _lib.check(ctx, L.apo_cavity_vertices(ctx, hit, rad, None, 0, _lib.byref(c)))
        kept[f"hit>{hit},r>{rad}"] = int(c.value)
    say(f"cavity hit ratios: {ms:.1f} ms for {vertices} vertices x 64 rays")
    return {"ms": ms, "vertices": vertices, "rays_per_s": vertices * 64 / (ms / 1000.0), "kept": kept,
            "note": "one call of apo_cavity_hit_ratios (64 Halton directions, any-hit through the LBVH), "
                    "wall time between device synchronisations"}


def protein_section(say):
    """The reference's five smallest sample proteins (MMTF, joggled by 1e-3 as
    its apps do): one eager build each after a warm-up."""
    import torch
    from apo_amd import _lib, molecules
    L = _lib.lib()
    out = {}
    for name in ("5ZCK", "1AGA", "3DIK", "101M", "1A3F"):
        s = molecules.joggle(molecules.load_mmtf(os.path.join(ROOT, "tests", "golden", "mmtf", f"{name}.mmtf")), 1e-3)
        ctx = _ctx({})
        _lib.check(ctx, L.apo_build(ctx, _lib.fptr(s), len(s)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(ctx, L.apo_build(ctx, _lib.fptr(s), len(s)))
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0)
        st = _lib.Stats()
        _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(st)))
        L.apo_destroy(ctx)
        out[name] = {"atoms": int(len(s)), "ms": ms, "device_ms": st.ms_total, "vertices": int(st.vertices)}
    say("proteins: " + ", ".join(f"{k} {v['atoms']} atoms {v['ms']:.2f} ms" for k, v in out.items()))
    return out


def lbvh_sizes(say, sizes=(1_000_000, 4_000_000, 16_000_000)):
    """The LBVH stage alone (scene box, Morton codes, sort, Karras, refit) at
    growing site counts of the same distribution: eager builds stopped after
    the kNN pass, LBVH time from the stage events, 440 algorithmic bytes per
    site (the Morton, sort, leaf, hierarchy and refit traffic)."""
    import torch
    from apo_amd import _lib, datasets
    L = _lib.lib()
    out = {}
    for n in sizes:
        sites = datasets.loguniform_sites(n, seed=3)
        dev = torch.from_numpy(sites).to("cuda")
        torch.cuda.synchronize()
        ctx = _ctx({"stage_mask": 8})
        ptr = _lib.ctypes.c_void_p(dev.data_ptr())
        times = []
        for _ in range(4):
            _lib.check(ctx, L.apo_build_device(ctx, ptr, n))
            s = _lib.Stats()
            _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(s)))
            times.append(s.ms_lbvh)
        L.apo_destroy(ctx)
        del dev
        torch.cuda.empty_cache()
        ms = float(np.median(times[1:]))
        gbs = 440.0 * n / (ms * 1e-3) / 1e9
        out[str(n)] = {"ms": ms, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS}
        say(f"LBVH at {n} sites: {ms:.3f} ms, {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak)")
    return out


def cpu_baseline(sites: np.ndarray, target: int):
    """Time the oracle (host C++/OpenMP restatement of the reference's method,
    reference edge searches) on a sub-cube of the workload holding about
    `target` sites: same density and radius distribution, bounded CPU time."""
    import oracle
    edge = (target / len(sites)) ** (1.0 / 3.0)
    sub = sites[np.all(sites[:, :3] < edge, axis=1)]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    d = oracle.Diagram(sub, threads=threads)
    dt = time.perf_counter() - t0
    v = d.stats()["vertices"]
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": len(sub) / dt, "unit": "cells/s", "cores": threads, "kind": "port", "cpu_model": model,
            "sample": f"{len(sub)} sites of the workload inside [0,{edge:.4f}]^3 (same density and radii), "
                      f"full construction by the oracle with the reference's edge searches, {dt:.2f} s, "
                      f"{v} vertices, {threads} threads on {model}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)
    import torch
    import apo_amd
    from apo_amd import _lib, datasets

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.lbvh_only:
        print(json.dumps({"lbvh_sizes": lbvh_sizes(lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))}),
              flush=True)
        return
    n = args.n
    sites = datasets.loguniform_sites(n, seed=3)
    dev = torch.from_numpy(sites).to("cuda")
    torch.cuda.synchronize()

    from apo_amd import autotune
    tuned = autotune.load("loguniform_1m") if n == 1_000_000 and args.k == 16 else None
    plain_tuning = dict(tuned["plain"]) if tuned else {}
    learned_tuning = dict(tuned.get("learned", tuned["plain"])) if tuned else {"proposal_cost": 400.0}
    plain_tuning["k"] = args.k
    learned_tuning["k"] = args.k
    ctx = _ctx(plain_tuning)
    L = _lib.lib()
    ptr = _lib.ctypes.c_void_p(dev.data_ptr())

    def step():
        _lib.check(ctx, L.apo_build_device(ctx, ptr, n))

    def say(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    say(f"{n} sites resident on {torch.cuda.get_device_name(0)}; {args.warmup} warm-up steps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    say(f"timing {args.steps} steps")
    # Each step is bracketed by device synchronisations; per-kernel times
    # and byte counts are read between steps, outside the timed spans, and
    # summed over all timed steps.
    # (the bisector entry is the whole stage: the plan, search and per-cell
    # kernels, timed and counted together)
    kern = {"k_edge_validation": [0.0, 0, 0], "k_vertex_validation": [0.0, 0, 0], "k_cells": [0.0, 0, 0],
            "bisector_stage": [0.0, 0, 0]}
    s = _lib.Stats()
    wall = 0.0
    # every build launches each validation kernel once per round (rounds
    # whose task list is empty end at once)
    rounds = 10
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        _lib.check(ctx, L.apo_get_stats(ctx, _lib.byref(s)))
        for name, ms, b, launches in (
                ("k_edge_validation", s.ms_edge_validation_kernel, s.bytes_edge_validation, rounds),
                ("k_vertex_validation", s.ms_vertex_validation_kernel, s.bytes_vertex_validation, rounds),
                ("k_cells", s.ms_cells, s.bytes_cells, 1),
                ("bisector_stage", s.ms_bisector, s.bytes_bisector_validation, 1)):
            kern[name][0] += ms
            kern[name][1] += b
            kern[name][2] += launches
    ms_per_step = 1000.0 * wall / args.steps
    say(f"{ms_per_step:.1f} ms per step")

    # Calibration copy for the HBM counters: one elementwise device copy of
    # exactly 256 MiB read + 256 MiB written (profiled beside the kernels).
    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst

    vertices = int(s.vertices)
    stage = {k: 0.0 for k in ("ms_total", "ms_lbvh", "ms_knn", "ms_cells", "ms_validation", "ms_bisector",
                              "ms_output")}
    for k in stage:
        stage[k] = float(getattr(s, k))
    # Roofline of the dominant kernel: its algorithmic bytes (counted by the
    # kernel itself: records, sites and per-cell scalars it must read or
    # write once) over its launch time (HIP events around each launch on the
    # library's stream), both summed over the timed steps.
    dominant = max(kern, key=lambda k: kern[k][0])
    ms_k, bytes_k, launches = kern[dominant]
    achieved = bytes_k / (ms_k * 1e-3) / 1e9
    lbvh_gbs = 440.0 * n / (s.ms_lbvh * 1e-3) / 1e9
    # HBM bytes per launch of that kernel from the two rocprofv3 counter
    # passes (FETCH_SIZE, WRITE_SIZE) committed under profiles/, when they
    # were taken on this workload
    traffic, traffic_src = None, None
    tp = os.path.join(ROOT, "profiles", f"pmc_traffic_{dominant}_1M_loguniform.json")
    if n == 1_000_000 and args.k == 16 and os.path.exists(tp):
        with open(tp) as f:
            traffic = float(json.load(f)["traffic_bytes_per_launch"])
        traffic_src = os.path.relpath(tp, ROOT)
    # VALU instructions per launch of that kernel from the SQ counter passes
    # committed under profiles/ (same workload): the issue rate it reaches
    # with this run's launch time, against 4 SIMDs per CU issuing one wave64
    # VALU instruction every 4 cycles at the 2.4 GHz peak clock
    valu_issue = None
    vp = os.path.join(ROOT, "profiles", "pmc_valu_validation_kernels_1M_loguniform.json")
    if n == 1_000_000 and args.k == 16 and os.path.exists(vp):
        with open(vp) as f:
            vk = json.load(f)["kernels"].get(dominant)
        if vk:
            cus = torch.cuda.get_device_properties(0).multi_processor_count
            peak = cus * 4 * 2.4e9 / 4.0
            rate = vk["valu_insts_per_launch"] / (ms_k / max(int(launches), 1) * 1e-3)
            valu_issue = {"achieved": rate, "peak": peak, "unit": "wave64 VALU insts/s", "frac": rate / peak,
                          "lanes_per_inst": vk["lanes_per_valu_inst"], "source": os.path.relpath(vp, ROOT)}

    violations, precision = None, None
    if not args.no_check:
        say("FP64 empty-sphere check of every vertex through the LBVH")
        cnt = _lib.c_uint64()
        _lib.check(ctx, L.apo_check_empty_spheres(ctx, None, None, 0, 1e-9, _lib.byref(cnt)))
        violations = int(cnt.value)
        say("interval precision check of the vertices and closed edges")
        precision = _precision(ctx)

    if not args.no_check:
        out_cavity = cavity_section(ctx, vertices, say)
    say("graph replay of the same build (the timed steps of value)")
    gr = graph_replay(dev, n, plain_tuning, args.steps, args.warmup, say)
    out = {
        "metric": "cells/sec at 1M sites: learned-k vs plain on 1 GPU, 0 empty-sphere violations",
        "value": gr["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gr["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n} sites, centres U[0,1]^3, radii log-uniform [1e-4,1e-2], seed 3, plain k={args.k}",
                   "path": "plain, one HIP graph replay per step (the learned-k proposal is reported under "
                           "learned)", "sites": n, "k": args.k},
        "vertices_per_s": vertices / (gr["ms_per_step"] / 1000.0),
        "graph_replay": gr,
        "eager": {"cells_per_s": n / (ms_per_step / 1000.0), "ms_per_step": ms_per_step,
                  "note": "stage events and per-kernel events on; the roofline figures come from these steps"},
        "vertices": vertices,
        "stage_ms": stage,
        "buried_sites": int(s.buried),
        "flagged_cells": int(s.flagged_cells),
        "empty_sphere_violations": violations,
        "precision": precision,
        "cavity": out_cavity if not args.no_check else None,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "launches": int(launches), "avg_launch_ms": ms_k / max(int(launches), 1),
                     "bytes_per_launch": bytes_k / max(int(launches), 1),
                     "lbvh_achieved": lbvh_gbs, "lbvh_frac": lbvh_gbs / HBM_PEAK_GBS,
                     "valu_issue": valu_issue},
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kern.items()},
        # every persistent kernel's algorithmic bytes and launch time (the
        # counter passes of each one are compared against these)
        "kernels": {k: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[2] / args.steps,
                        "bytes_per_launch": v[1] / max(v[2], 1),
                        "achieved_gbs": v[1] / (v[0] * 1e-3) / 1e9 if v[0] > 0 else None}
                    for k, v in kern.items()},
    }
    out["tuning"] = {"source": autotune.cache_path("loguniform_1m") if tuned else None,
                     "plain": plain_tuning, "learned": learned_tuning}
    if not args.no_learned:
        _, ids_p = _vertices(ctx, L, _lib, vertices)
        plain_keys = _row_keys(ids_p)
        say("learned neighbour-count proposal: training, then the same workload")
        params, training = train_proposal(say)
        ltune = dict(learned_tuning)
        if args.proposal_cost is not None:
            ltune["proposal_cost"] = args.proposal_cost
        lc = run_learned(params, dev, n, ltune, args.steps, args.warmup, plain_keys)
        lc["training"] = training
        lc["speedup_vs_plain"] = lc["cells_per_s"] / out["eager"]["cells_per_s"]
        out["learned"] = lc
        # the ceiling of any proposal: the classes of the exact labels of
        # these very sites (the label pass, untimed), same knobs
        from apo_amd import proposal as prop
        _, lab = prop.labels(sites)
        ex = run_learned(params, dev, n, ltune, args.steps, args.warmup, plain_keys, classes=prop.label_class(lab))
        ex["speedup_vs_plain"] = ex["cells_per_s"] / out["eager"]["cells_per_s"]
        ex["note"] = "classes of the exact certified-k labels of the same sites instead of the MLP's"
        out["learned_exact_labels"] = ex
        say(f"learned {lc['ms_per_step']:.1f} ms, exact-label classes {ex['ms_per_step']:.1f} ms")
        # completion mode: plain and learned (MLP, then exact-label classes)
        # must emit identical vertex rows
        say("completion mode: plain, learned and exact-label builds")
        out["complete"], crows, ckeys, cradii = completion_section(dev, n, plain_tuning, args.steps, args.warmup,
                                                                   say)
        ltc = dict(ltune, complete=4)
        lcc = run_learned(params, dev, n, ltc, args.steps, args.warmup, ckeys, plain_rows=crows)
        lcc["speedup_vs_plain"] = lcc["cells_per_s"] / out["complete"]["eager"]["cells_per_s"]
        out["learned_complete"] = lcc
        exc = run_learned(params, dev, n, ltc, args.steps, args.warmup, ckeys, classes=prop.label_class(lab),
                          plain_rows=crows)
        exc["speedup_vs_plain"] = exc["cells_per_s"] / out["complete"]["eager"]["cells_per_s"]
        out["learned_exact_labels_complete"] = exc
        # the ceiling of a search-radius proposal: twice every cell's final
        # radius (apo_cell_radii of the completed plain build), 1e-3 above,
        # through the range-query cell pass
        buried = cradii == np.float32(-np.finfo(np.float32).max)
        exact_r = (2 * np.where(buried, np.float32(0), cradii) * np.float32(1 + 1e-3)).astype(np.float32)
        exr = run_learned(None, dev, n, ltc, args.steps, args.warmup, ckeys, plain_rows=crows, radii=exact_r)
        exr["speedup_vs_plain"] = exr["cells_per_s"] / out["complete"]["eager"]["cells_per_s"]
        exr["speedup_vs_plain_without_completion"] = exr["cells_per_s"] / out["eager"]["cells_per_s"]
        exr["note"] = ("search radius 2 x (1 + 1e-3) x the final cell radius of every site (range query, at most "
                       f"{max(ltc.get('proposal_max_k', 64), ltc.get('proposal_fallback_k', 0))} sites per list)")
        out["learned_exact_radius"] = exr
        # the same radii with lists of up to 64 sites: more lists end below
        # their radius, so more cells certify in the cell pass
        t64 = dict(ltc, proposal_max_k=64, proposal_fallback_k=64)
        ex64 = run_learned(None, dev, n, t64, args.steps, args.warmup, ckeys, plain_rows=crows, radii=exact_r)
        ex64["speedup_vs_plain"] = ex64["cells_per_s"] / out["complete"]["eager"]["cells_per_s"]
        ex64["note"] = "as learned_exact_radius with at most 64 sites per list"
        out["learned_exact_radius_64"] = ex64
        # the radius head of the same MLP (proposal_radius): each site's
        # predicted search radius through the range-query cell pass
        tr = dict(ltc, proposal_radius=1)
        lr_ = run_learned(params, dev, n, tr, args.steps, args.warmup, ckeys, plain_rows=crows)
        lr_["speedup_vs_plain"] = lr_["cells_per_s"] / out["complete"]["eager"]["cells_per_s"]
        lr_["note"] = ("search radius from the MLP's radius head (d16 x 2^output), range query of at most "
                       f"{max(tr.get('proposal_max_k', 64), tr.get('proposal_fallback_k', 0))} sites per list")
        out["learned_radius"] = lr_
        say(f"completion: learned {lcc['ms_per_step']:.1f} ms (identical rows {lcc['identical_vertex_rows']}), "
            f"exact labels {exc['ms_per_step']:.1f} ms (identical rows {exc['identical_vertex_rows']}), "
            f"exact radii {exr['ms_per_step']:.1f} ms (identical rows {exr['identical_vertex_rows']}, "
            f"{exr['cells_certified_by_cell_pass']} certified by the cell pass), with 64-site lists "
            f"{ex64['ms_per_step']:.1f} ms ({ex64['cells_certified_by_cell_pass']} certified); radius head "
            f"{lr_['ms_per_step']:.1f} ms (identical rows {lr_['identical_vertex_rows']})")
        if not args.no_c5:
            say("C5: the proposal on a 5M-site set")
            out["learned_5m"] = c5_evaluation(params, plain_tuning, ltune, say)
    if not args.no_learned:
        # plain builds at other neighbour counts (one eager build each): the
        # cost split that bounds what a per-cell neighbour count can gain
        sweep = {}
        for kk in (24, 32, 48, 64):
            c = _ctx(dict(plain_tuning, k=kk))
            _lib.check(c, L.apo_build_device(c, ptr, n))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _lib.check(c, L.apo_build_device(c, ptr, n))
            torch.cuda.synchronize()
            ms = 1000.0 * (time.perf_counter() - t0)
            s2 = _lib.Stats()
            _lib.check(c, L.apo_get_stats(c, _lib.byref(s2)))
            L.apo_destroy(c)
            sweep[str(kk)] = {"ms": ms, "knn": s2.ms_knn, "cells": s2.ms_cells, "validation": s2.ms_validation,
                              "bisector": s2.ms_bisector, "certified": int(s2.cells_certified),
                              "vertices": int(s2.vertices)}
        sweep["16"] = {"ms": ms_per_step, "knn": stage["ms_knn"], "cells": stage["ms_cells"],
                       "validation": stage["ms_validation"], "bisector": stage["ms_bisector"],
                       "certified": int(s.cells_certified), "vertices": vertices}
        out["plain_k_sweep"] = sweep
        say("plain k sweep: " + ", ".join(f"k={k} {v['ms']:.0f} ms" for k, v in sorted(sweep.items(), key=lambda t: int(t[0]))))
    if not args.no_batch:
        out["batch"] = batch_workload(args.batch_molecules, min(args.steps, 3), 1, say)
        out["proteins"] = protein_section(say)
        out["lbvh_sizes"] = lbvh_sizes(say)
    if not args.no_cpu_baseline:
        say(f"CPU baseline: oracle on a sub-cube of about {args.cpu_sample} sites")
        out["cpu_baseline"] = cpu_baseline(sites, args.cpu_sample)
    L.apo_destroy(ctx)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
