"""Autotuner of the construction's launch and search parameters.

The plain path's output does not depend on any knob tuned here; only the
time to produce it does:

* waves_cells, waves_vertex, waves_edge, waves_bisector -- waves per CU of
  the persistent work-queue grids (the wave "tiles" each CU runs at once),
  from 1 to each kernel's full occupancy;
* cover_ratio -- the known-radius multiple above which an edge search is
  pruned by the cover of the edge's empty spheres (exact pruning);
* cells_per_wave -- the cell pass's tile: a whole wave64 per cell, or half
  a wave per cell (two cells per wave, twice the LDS per wave).

The learned path adds the neighbours per pass of its fallback class
(proposal_fallback_k: the neighbour count of sites predicted not to
certify), the most neighbours any other class may run (proposal_max_k,
which also sets the length of its kNN lists) and the validation cost of its
class choice (proposal_cost); these change which cells reach validation,
never the certification test, so its output stays the certified diagram.

Fixed on purpose: the plain path's 16 neighbours (the reference's default,
part of its output) and the traversal stack depth of 32 entries (a
shallower stack would change results where it overflows).

Search: coordinate descent.  Each knob in turn is swept with the others
held, scored by the device time of the kernel it drives (apo_stats, median
of `repeats` builds), and the best value kept; the learned knobs, which
interact (the list length bounds every class), get two passes.  The tuned plain
configuration is then checked to emit exactly the default configuration's
vertices.  Results are cached as JSON keyed by device and workload.
"""
from __future__ import annotations

import json
import os
import statistics
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _lib

TUNED_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned")

# knob -> stats field it is scored by (None: the whole build)
PLAIN_KNOBS = {
    "cells_per_wave": "ms_cells",
    "waves_cells": "ms_cells",
    "waves_vertex": "ms_vertex_validation_kernel",
    "waves_edge": "ms_edge_validation_kernel",
    "waves_bisector": "ms_bisector",
    "cover_ratio": "ms_edge_validation_kernel",
}
LEARNED_KNOBS = {"proposal_max_k": None, "proposal_fallback_k": None, "proposal_cost": None}
# (a negative ratio, never covering, is valid but leaves large edge searches
# unpruned: minutes per build at 1M sites, so it is not a candidate)
COVER_RATIOS = (0.0, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0, 32.0)
FALLBACK_K = (16, 24, 32, 48, 64)
MAX_K = (24, 32, 48, 64)
PROPOSAL_COSTS = (0.0, 100.0, 200.0, 400.0, 800.0)


def coordinate_descent(knobs: Sequence[str], candidates: Dict[str, Sequence], start: Dict[str, object],
                       score: Callable[[Dict[str, object], str], float],
                       passes: int = 1) -> (Dict[str, object], List[dict]):
    """`passes` passes over `knobs`: for each, try every candidate value with
    the other knobs at their current best and keep the lowest score.  `score`
    receives the full configuration and the knob being tuned."""
    best = dict(start)
    log = []
    for _ in range(passes):
        for knob in knobs:
            trial = {}
            for value in candidates[knob]:
                cfg = dict(best)
                cfg[knob] = value
                trial[value] = score(cfg, knob)
                log.append({"knob": knob, "value": value, "score": trial[value]})
            best[knob] = min(trial, key=lambda v: (trial[v], str(v)))
    return best, log


def _wave_candidates(occupancy: int) -> List[int]:
    vals = sorted({v for v in (1, 2, 3, 4, 6, 8, 12, 16) if v <= occupancy} | {occupancy})
    return vals


def _device_name() -> str:
    """The library's target and the CU count of device 0 (e.g. gfx950_256cu)."""
    L = _lib.lib()
    c = _lib.Config()
    L.apo_default_config(_lib.byref(c))
    ctx = _lib.create(c)
    try:
        info = np.zeros(5, np.uint32)
        _lib.check(ctx, L.apo_launch_occupancy(ctx, _lib.uptr(info)))
        return f"gfx950_{int(info[4])}cu"
    finally:
        L.apo_destroy(ctx)


def cache_path(workload: str, device: Optional[str] = None) -> str:
    dev = (device or _device_name()).replace(" ", "_").replace("/", "_")
    return os.path.join(TUNED_DIR, f"{dev}__{workload}.json")


def load(workload: str, device: Optional[str] = None) -> Optional[dict]:
    p = cache_path(workload, device)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def save(workload: str, result: dict, device: Optional[str] = None) -> str:
    os.makedirs(TUNED_DIR, exist_ok=True)
    p = cache_path(workload, device)
    with open(p, "w") as f:
        json.dump(result, f, indent=1, sort_keys=True)
    return p


class _Runner:
    """Builds a site set with a given configuration (host sites: each build
    copies them to the device before its timed span starts)."""

    def __init__(self, sites, offsets=None, proposal=None, repeats: int = 2):
        self.sites = np.ascontiguousarray(sites, np.float32).reshape(-1, 4)
        self.n = len(self.sites)
        self.offsets = None if offsets is None else np.ascontiguousarray(offsets, np.uint32)
        self.proposal = None if proposal is None else np.ascontiguousarray(proposal, np.float32)
        self.repeats = repeats
        self.L = _lib.lib()

    def _ctx(self, cfg: dict):
        c = _lib.Config()
        self.L.apo_default_config(_lib.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        ctx = _lib.create(c)
        if self.proposal is not None:
            _lib.check(ctx, self.L.apo_proposal_set(ctx, _lib.fptr(self.proposal)))
        return ctx

    def _build(self, ctx):
        if self.offsets is None:
            _lib.check(ctx, self.L.apo_build(ctx, _lib.fptr(self.sites), self.n))
        else:
            _lib.check(ctx, self.L.apo_build_batched(ctx, _lib.fptr(self.sites), _lib.uptr(self.offsets),
                                                     len(self.offsets) - 1))

    def stats(self, cfg: dict) -> List[dict]:
        ctx = self._ctx(cfg)
        try:
            self._build(ctx)  # warm-up
            out = []
            for _ in range(self.repeats):
                self._build(ctx)
                s = _lib.Stats()
                _lib.check(ctx, self.L.apo_get_stats(ctx, _lib.byref(s)))
                out.append({f: getattr(s, f) for f, _ in s._fields_ if f != "reserved"})
            return out
        finally:
            self.L.apo_destroy(ctx)

    def ids(self, cfg: dict) -> np.ndarray:
        ctx = self._ctx(cfg)
        try:
            self._build(ctx)
            v = _lib.c_uint64()
            _lib.check(ctx, self.L.apo_vertex_count(ctx, _lib.byref(v)))
            ids = np.zeros((v.value, 4), np.uint32)
            _lib.check(ctx, self.L.apo_get_vertices(ctx, None, _lib.uptr(ids), v.value))
            return ids
        finally:
            self.L.apo_destroy(ctx)

    def occupancy(self) -> List[int]:
        ctx = self._ctx({})
        try:
            occ = np.zeros(5, np.uint32)
            _lib.check(ctx, self.L.apo_launch_occupancy(ctx, _lib.uptr(occ)))
            return [int(x) for x in occ[:4]]
        finally:
            self.L.apo_destroy(ctx)


def tune(sites, workload: str, offsets=None, proposal=None, repeats: int = 2,
         say: Callable[[str], None] = lambda m: None, cache: bool = True) -> dict:
    """Tune the plain knobs (and the learned knobs when `proposal` holds MLP
    parameters) on `sites` ((n, 4) float32, a batch when `offsets` is given);
    returns the result record (also written under apo_amd/tuned/ when
    `cache`)."""
    run = _Runner(sites, offsets=offsets, repeats=repeats)
    n = run.n
    occ = run.occupancy()
    candidates = {"waves_cells": _wave_candidates(occ[0]), "waves_vertex": _wave_candidates(occ[1]),
                  "waves_edge": _wave_candidates(occ[2]), "waves_bisector": _wave_candidates(occ[3]),
                  "cover_ratio": COVER_RATIOS, "cells_per_wave": (1, 2)}
    start = {"cells_per_wave": 1, "waves_cells": 0, "waves_vertex": 0, "waves_edge": 0, "waves_bisector": 0,
             "cover_ratio": 2.0}

    def plain_score(cfg, knob):
        st = run.stats(cfg)
        t = statistics.median(s[PLAIN_KNOBS[knob]] for s in st)
        say(f"tune {knob}={cfg[knob]}: {t:.2f} ms")
        return t

    default_total = statistics.median(s["ms_total"] for s in run.stats(start))
    best, log = coordinate_descent(list(PLAIN_KNOBS), candidates, start, plain_score)
    tuned_total = statistics.median(s["ms_total"] for s in run.stats(best))
    same = bool(np.array_equal(run.ids(best), run.ids({})))
    result = {"workload": workload, "sites": n, "occupancy": occ, "plain": best, "log": log,
              "ms_total_default": default_total, "ms_total_tuned": tuned_total, "same_vertices_as_default": same}
    if proposal is not None:
        # whole-build scores differ by several ms between runs: more repeats
        lrun = _Runner(sites, offsets=offsets, proposal=proposal, repeats=max(repeats, 5))

        def learned_score(cfg, knob):
            t = statistics.median(s["ms_total"] for s in lrun.stats(cfg))
            say(f"tune {knob}={cfg[knob]}: {t:.2f} ms")
            return t

        lstart = dict(best, proposal_max_k=64, proposal_fallback_k=16, proposal_cost=400.0)
        lbest, llog = coordinate_descent(list(LEARNED_KNOBS), {"proposal_max_k": MAX_K,
                                                               "proposal_fallback_k": FALLBACK_K,
                                                               "proposal_cost": PROPOSAL_COSTS}, lstart,
                                         learned_score, passes=2)
        result["learned"] = lbest
        result["learned_log"] = llog
    if cache:
        result["path"] = os.path.relpath(save(workload, result), os.path.dirname(TUNED_DIR))
    return result


def _main():
    import argparse
    import shutil
    import sys
    import time

    from . import datasets, proposal as prop

    ap = argparse.ArgumentParser(description="tune the construction on a named workload")
    ap.add_argument("workload", choices=["loguniform_1m", "batch_128"])
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--learned", action="store_true", help="also tune the learned path (trains the MLP first)")
    ap.add_argument("--copy-to", default=None, help="also copy the JSON record into this directory")
    args = ap.parse_args()

    def say(m):
        print(f"[autotune] {m}", file=sys.stderr, flush=True)

    offsets = None
    if args.workload == "loguniform_1m":
        sites = datasets.loguniform_sites(1_000_000, seed=3)
    else:
        sites, offsets = datasets.batched_molecules(count=128, seed=4, lo=500, hi=50000)
    params = None
    if args.learned:
        t0 = time.perf_counter()
        fa, la = prop.labels(datasets.loguniform_sites(1_000_000, seed=501))
        fb, lb = prop.labels(datasets.loguniform_sites(1_000_000, seed=502))
        params, _ = prop.train(np.concatenate([fa, fb]), np.concatenate([la, lb]), epochs=8, batch=4096,
                               lr=2e-3, seed=7)
        say(f"proposal trained in {time.perf_counter() - t0:.1f} s")
    res = tune(sites, args.workload, offsets=offsets, proposal=params, repeats=args.repeats, say=say)
    say(f"default {res['ms_total_default']:.1f} ms -> tuned {res['ms_total_tuned']:.1f} ms, "
        f"same vertices: {res['same_vertices_as_default']}")
    if args.copy_to:
        os.makedirs(args.copy_to, exist_ok=True)
        shutil.copy(cache_path(args.workload), args.copy_to)
    print(json.dumps({k: res[k] for k in ("plain", "ms_total_default", "ms_total_tuned", "same_vertices_as_default")}
                     | ({"learned": res["learned"]} if "learned" in res else {})))


if __name__ == "__main__":
    _main()
